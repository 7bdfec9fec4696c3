/*
 * fbh5_writer.c -- FBH5 (filterbank in HDF5) output of the rawspec program,
 * built as a plugin (librawspec_fbh5.so) that the program loads only when
 * FBH5 output is asked for, so that the program itself does not depend on
 * the HDF5 libraries.
 *
 * File layout, as the reference writes it (fbh5_open.c, fbh5_write.c,
 * fbh5_close.c, fbh5_util.c):
 *  - file attributes CLASS = "FILTERBANK", VERSION = "2.0", the software
 *    versions (VERSION_RAWSPEC, VERSION_LIBRAWSPEC, VERSION_CUFFT,
 *    VERSION_HDF5) and BITSHUFFLE = "ENABLED" / "DISABLED";
 *  - dataset "data": [time][nifs][nchans] little-endian float32 (nbits 32),
 *    unlimited along time, chunks of (spectra per dump, 1, fine channels
 *    per coarse channel), compressed with the Bitshuffle filter (LZ4) when
 *    the HDF5 plugin path provides it;
 *  - dataset attributes: the filterbank header fields (ints, doubles and the
 *    two strings; refbeam only when it is not -1);
 *  - dimension labels "time", "feed_id", "frequency", set at close.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hdf5.h"
#include "hdf5_hl.h"

#include "fbh5_writer.h"

#define FBH5_BITSHUFFLE 32008

struct fbh5w_ctx {
  int active;
  hid_t file, dataset, space, type;
  unsigned int elem_size;
  size_t tint_size;        /* bytes of one integration: nifs * nchans * elem_size */
  hsize_t offset[3];       /* where the next integration goes */
  hsize_t size[3];         /* the data set's extent */
  size_t bytes;
};

static void warn(const char *what)
{
  fprintf(stderr, "FBH5-WARNING :: %s\n", what);
}

static void str_attr(hid_t obj, const char *tag, const char *value)
{
  hid_t space = H5Screate(H5S_SCALAR);
  hid_t type = H5Tcopy(H5T_C_S1);
  hid_t attr;
  H5Tset_size(type, strlen(value) ? strlen(value) : 1);
  H5Tset_strpad(type, H5T_STR_NULLTERM);
  attr = H5Acreate2(obj, tag, type, space, H5P_DEFAULT, H5P_DEFAULT);
  if (attr < 0) {
    warn(tag);
  } else {
    if (H5Awrite(attr, type, value) < 0) warn(tag);
    H5Aclose(attr);
  }
  H5Tclose(type);
  H5Sclose(space);
}

static void num_attr(hid_t obj, const char *tag, hid_t type, const void *value)
{
  hid_t space = H5Screate(H5S_SCALAR);
  hid_t attr = H5Acreate2(obj, tag, type, space, H5P_DEFAULT, H5P_DEFAULT);
  if (attr < 0) {
    warn(tag);
  } else {
    if (H5Awrite(attr, type, value) < 0) warn(tag);
    H5Aclose(attr);
  }
  H5Sclose(space);
}

static void header_attrs(hid_t ds, const fb_hdr_t *h)
{
  num_attr(ds, "machine_id", H5T_NATIVE_INT, &h->machine_id);
  num_attr(ds, "telescope_id", H5T_NATIVE_INT, &h->telescope_id);
  num_attr(ds, "data_type", H5T_NATIVE_INT, &h->data_type);
  num_attr(ds, "nchans", H5T_NATIVE_INT, &h->nchans);
  num_attr(ds, "nfpc", H5T_NATIVE_INT, &h->nfpc);
  num_attr(ds, "nbeams", H5T_NATIVE_INT, &h->nbeams);
  num_attr(ds, "ibeam", H5T_NATIVE_INT, &h->ibeam);
  if (h->refbeam != -1) num_attr(ds, "refbeam", H5T_NATIVE_INT, &h->refbeam);
  num_attr(ds, "nbits", H5T_NATIVE_INT, &h->nbits);
  num_attr(ds, "nifs", H5T_NATIVE_INT, &h->nifs);
  num_attr(ds, "src_raj", H5T_NATIVE_DOUBLE, &h->src_raj);
  num_attr(ds, "src_dej", H5T_NATIVE_DOUBLE, &h->src_dej);
  num_attr(ds, "az_start", H5T_NATIVE_DOUBLE, &h->az_start);
  num_attr(ds, "za_start", H5T_NATIVE_DOUBLE, &h->za_start);
  num_attr(ds, "fch1", H5T_NATIVE_DOUBLE, &h->fch1);
  num_attr(ds, "foff", H5T_NATIVE_DOUBLE, &h->foff);
  num_attr(ds, "tstart", H5T_NATIVE_DOUBLE, &h->tstart);
  num_attr(ds, "tsamp", H5T_NATIVE_DOUBLE, &h->tsamp);
  str_attr(ds, "source_name", h->source_name);
  str_attr(ds, "rawdatafile", h->rawdatafile);
}

fbh5w_ctx *fbh5w_open(const fb_hdr_t *h, unsigned int nds, const char *path,
                      const char *version_rawspec, const char *version_lib,
                      const char *version_fft)
{
  fbh5w_ctx *c = (fbh5w_ctx *)calloc(1, sizeof *c);
  hid_t dcpl, type;
  hsize_t maxd[3], chunk[3];
  unsigned maj, min, rel;
  char txt[64];
  int bitshuffle = H5Zfilter_avail(FBH5_BITSHUFFLE) > 0;

  if (!c) return NULL;
  if (h->nbits % 8 != 0 || h->nbits < 8 || h->nbits > 64 || h->nifs < 1 || h->nifs > 4 ||
      h->nfpc < 1 || h->nchans < 1 || nds < 1) {
    fprintf(stderr, "FBH5-ERROR :: bad header for %s (nbits %d, nifs %d, nfpc %d, nchans %d)\n",
            path, h->nbits, h->nifs, h->nfpc, h->nchans);
    free(c);
    return NULL;
  }
  if (!bitshuffle) warn("the Bitshuffle filter is not available; data will not be compressed");
  c->elem_size = (unsigned)h->nbits / 8;
  c->tint_size = (size_t)h->nifs * (size_t)h->nchans * c->elem_size;
  c->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (c->file < 0) {
    fprintf(stderr, "FBH5-ERROR :: cannot create %s\n", path);
    free(c);
    return NULL;
  }
  str_attr(c->file, "CLASS", "FILTERBANK");
  str_attr(c->file, "VERSION", "2.0");
  str_attr(c->file, "VERSION_RAWSPEC", version_rawspec);
  str_attr(c->file, "VERSION_LIBRAWSPEC", version_lib);
  str_attr(c->file, "VERSION_CUFFT", version_fft);
  H5get_libversion(&maj, &min, &rel);
  snprintf(txt, sizeof txt, "%u.%u.%u", maj, min, rel);
  str_attr(c->file, "VERSION_HDF5", txt);
  str_attr(c->file, "BITSHUFFLE", bitshuffle ? "ENABLED" : "DISABLED");

  /* one (empty) integration at creation; the first write grows from it */
  c->size[0] = 1;
  c->size[1] = (hsize_t)h->nifs;
  c->size[2] = (hsize_t)h->nchans;
  maxd[0] = H5S_UNLIMITED;
  maxd[1] = c->size[1];
  maxd[2] = c->size[2];
  c->space = H5Screate_simple(3, c->size, maxd);
  dcpl = H5Pcreate(H5P_DATASET_CREATE);
  chunk[0] = nds;
  chunk[1] = 1;
  chunk[2] = (hsize_t)h->nfpc;
  if (c->space < 0 || dcpl < 0 || H5Pset_chunk(dcpl, 3, chunk) < 0) {
    fprintf(stderr, "FBH5-ERROR :: cannot set up the dataset of %s\n", path);
    if (dcpl >= 0) H5Pclose(dcpl);
    fbh5w_close(c);
    return NULL;
  }
  if (bitshuffle) {
    const unsigned opts[2] = {0, 2};   /* default block size, LZ4 */
    if (H5Pset_filter(dcpl, FBH5_BITSHUFFLE, H5Z_FLAG_MANDATORY, 2, opts) < 0)
      warn("cannot add the Bitshuffle filter; data will not be compressed");
  }
  switch (h->nbits) {
    case 8: type = H5T_NATIVE_B8; break;
    case 16: type = H5T_NATIVE_B16; break;
    case 32: type = H5T_IEEE_F32LE; break;
    default: type = H5T_IEEE_F64LE; break;
  }
  c->type = type;
  c->dataset = H5Dcreate2(c->file, "data", type, c->space, H5P_DEFAULT, dcpl, H5P_DEFAULT);
  H5Pclose(dcpl);
  if (c->dataset < 0) {
    fprintf(stderr, "FBH5-ERROR :: cannot create the data set of %s\n", path);
    fbh5w_close(c);
    return NULL;
  }
  header_attrs(c->dataset, h);
  c->active = 1;
  return c;
}

int fbh5w_write(fbh5w_ctx *c, const void *buf, size_t bytes)
{
  hsize_t sel[3];
  hid_t fspace;
  const size_t n = bytes / c->tint_size;
  if (!c->active) return 1;
  if (n == 0) return 0;
  c->size[0] += c->offset[0] > 0 ? n : n - 1;
  sel[0] = n;
  sel[1] = c->size[1];
  sel[2] = c->size[2];
  if (H5Dset_extent(c->dataset, c->size) < 0 || H5Sset_extent_simple(c->space, 3, sel, c->size) < 0) {
    c->active = 0;
    return 1;
  }
  fspace = H5Dget_space(c->dataset);
  if (fspace < 0 || H5Sselect_hyperslab(fspace, H5S_SELECT_SET, c->offset, NULL, sel, NULL) < 0 ||
      H5Dwrite(c->dataset, c->type, c->space, fspace, H5P_DEFAULT, buf) < 0) {
    if (fspace >= 0) H5Sclose(fspace);
    c->active = 0;
    return 1;
  }
  H5Sclose(fspace);
  c->offset[0] += n;
  c->bytes += bytes;
  return 0;
}

int fbh5w_close(fbh5w_ctx *c)
{
  int bad = 0;
  if (!c) return 0;
  if (c->dataset > 0) {
    if (H5DSset_label(c->dataset, 0, "time") < 0 || H5DSset_label(c->dataset, 1, "feed_id") < 0 ||
        H5DSset_label(c->dataset, 2, "frequency") < 0)
      warn("cannot set the dimension labels");
  }
  if (c->space > 0 && H5Sclose(c->space) < 0) bad = 1;
  if (c->dataset > 0 && H5Dclose(c->dataset) < 0) bad = 1;
  if (c->file > 0 && H5Fclose(c->file) < 0) bad = 1;
  free(c);
  return bad;
}
