/*
 * fb_header.c -- SIGPROC filterbank header utilities (fb_*), the functions
 * the reference's librawspec.so exports from rawspec_fbutils.c.  Host-only C.
 *
 * The byte format (keyword order, the little-endian values, the ddmmss.s
 * angle encoding, the 80-character string limit and the "rawdatafile"
 * padding scheme) is the reference's, so headers written here are
 * byte-identical to its output.  Implementation: every writer goes through
 * one sink that targets either a file descriptor or a memory buffer, so the
 * header layout is stated once.
 *
 * Deliberate differences: fb_fd_write_header / fb_buf_write_header return
 * the byte count / end position (the reference's versions return nothing);
 * the header readers skip the values of the SIGPROC keywords they do not
 * store (nsamples, refdm, period, npuls, nbins, fchannel) instead of
 * reading the value bytes as the next keyword; fb_fd_read_header never
 * parses past the bytes it read and seeks to the end of the header however
 * short the file is.
 */
#define _GNU_SOURCE
#include <endian.h>
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>
#include <unistd.h>

#include "rawspec_fbutils.h"

#define FB_STR_MAX 80

/* ---- angles ------------------------------------------------------------- */

double fb_ddd_to_dms(double ddd)
{
  const int sign = ddd < 0 ? -1 : +1;
  double mm, ss;
  ddd = fabs(ddd);
  mm = 60 * fmod(ddd, 1.0);
  ss = 60 * fmod(mm, 1.0);
  ddd = floor(ddd);
  mm = floor(mm);
  return sign * (10000 * ddd + 100 * mm + ss);
}

double fb_dms_to_ddd(double dms)
{
  const int sign = dms < 0 ? -1 : +1;
  double dd, mm, ss;
  dms = fabs(dms);
  dd = floor(dms / 10000);
  dms -= 10000 * dd;
  mm = floor(dms / 100);
  ss = dms - 100 * mm;
  dd += mm / 60.0 + ss / 3600.0;
  return sign * dd;
}

/* ---- sink: a file descriptor or a buffer -------------------------------- */

typedef struct {
  int fd;              /* >= 0: write to the descriptor */
  char *p;             /* otherwise: next byte of the buffer */
  ssize_t n;           /* bytes accounted (descriptor: sum of write returns) */
} fb_sink;

static ssize_t sink_put(fb_sink *s, const void *data, size_t len)
{
  if (s->fd >= 0) {
    const ssize_t r = write(s->fd, data, len);
    s->n += r;
    return r;
  }
  memcpy(s->p, data, len);
  s->p += len;
  s->n += (ssize_t)len;
  return (ssize_t)len;
}

static ssize_t sink_int(fb_sink *s, int32_t i)
{
  const uint32_t le = htole32((uint32_t)i);
  return sink_put(s, &le, sizeof le);
}

static ssize_t sink_double(fb_sink *s, double d)
{
  uint64_t bits;
  memcpy(&bits, &d, sizeof bits);
  bits = htole64(bits);
  return sink_put(s, &bits, sizeof bits);
}

/* The length word's own write is not counted for descriptors (so a short
 * string write reports as the reference's does). */
static ssize_t sink_string(fb_sink *s, const char *c)
{
  const int32_t len = (int32_t)strnlen(c, FB_STR_MAX);
  if (s->fd >= 0) {
    const uint32_t le = htole32((uint32_t)len);
    ssize_t r;
    (void)!write(s->fd, &le, sizeof le);
    r = write(s->fd, c, (size_t)len) + (ssize_t)sizeof(int32_t);
    s->n += r;
    return r;
  }
  sink_int(s, len);
  return sink_put(s, c, (size_t)len) + (ssize_t)sizeof(int32_t);
}

/* ---- value writers -------------------------------------------------------- */

ssize_t fb_fd_write_int(int fd, int32_t i) { fb_sink s = {fd, NULL, 0}; return sink_int(&s, i); }
void *fb_buf_write_int(void *buf, int32_t i) { fb_sink s = {-1, buf, 0}; sink_int(&s, i); return s.p; }
ssize_t fb_fd_write_double(int fd, double d) { fb_sink s = {fd, NULL, 0}; return sink_double(&s, d); }
void *fb_buf_write_double(void *buf, double d) { fb_sink s = {-1, buf, 0}; sink_double(&s, d); return s.p; }
ssize_t fb_fd_write_angle(int fd, double d) { return fb_fd_write_double(fd, fb_ddd_to_dms(d)); }
void *fb_buf_write_angle(void *buf, double d) { return fb_buf_write_double(buf, fb_ddd_to_dms(d)); }
ssize_t fb_fd_write_string(int fd, const char *c) { fb_sink s = {fd, NULL, 0}; return sink_string(&s, c); }
void *fb_buf_write_string(void *buf, const char *c) { fb_sink s = {-1, buf, 0}; sink_string(&s, c); return s.p; }

/* ---- header writer -------------------------------------------------------- */

static void keyword_int(fb_sink *s, const char *k, int32_t v) { sink_string(s, k); sink_int(s, v); }
static void keyword_double(fb_sink *s, const char *k, double v) { sink_string(s, k); sink_double(s, v); }

static void write_header(fb_sink *s, const fb_hdr_t *hdr, int32_t minlen)
{
  char pad[FB_STR_MAX];
  int32_t padlen;
  sink_string(s, "HEADER_START");
  keyword_int(s, "machine_id", hdr->machine_id);
  keyword_int(s, "telescope_id", hdr->telescope_id);
  keyword_double(s, "src_raj", fb_ddd_to_dms(hdr->src_raj));
  keyword_double(s, "src_dej", fb_ddd_to_dms(hdr->src_dej));
  keyword_double(s, "az_start", hdr->az_start);
  keyword_double(s, "za_start", hdr->za_start);
  keyword_int(s, "data_type", hdr->data_type);
  keyword_double(s, "fch1", hdr->fch1);
  keyword_double(s, "foff", hdr->foff);
  keyword_int(s, "nchans", hdr->nchans);
  keyword_int(s, "nbeams", hdr->nbeams);
  keyword_int(s, "ibeam", hdr->ibeam);
  keyword_int(s, "nbits", hdr->nbits);
  keyword_double(s, "tstart", hdr->tstart);
  keyword_double(s, "tsamp", hdr->tsamp);
  keyword_int(s, "nifs", hdr->nifs);
  if (hdr->barycentric) keyword_int(s, "barycentric", hdr->barycentric);
  if (hdr->pulsarcentric) keyword_int(s, "pulsarcentric", hdr->pulsarcentric);
  sink_string(s, "source_name");
  sink_string(s, hdr->source_name);
  /* Padding: dummy "rawdatafile" entries before the real one.  One entry
   * with a k-character value takes 4 + 11 + 4 + k = 19 + k bytes (k >= 1,
   * k <= 79): while more than 98 bytes remain, take 79 with 60-blank values;
   * then one entry of (remaining - 19) blanks if more than 19 remain. */
  padlen = minlen - (int32_t)s->n -
           (int32_t)(2 * sizeof(int32_t) + strlen("rawdatafile") + strlen(hdr->rawdatafile)) -
           (int32_t)(sizeof(int32_t) + strlen("HEADER_END"));
  memset(pad, ' ', sizeof pad);
  while (padlen > 98) {
    pad[60] = '\0';
    sink_string(s, "rawdatafile");
    sink_string(s, pad);
    pad[60] = ' ';
    padlen -= 79;
  }
  if (padlen > 19) {
    pad[padlen - 19] = '\0';
    sink_string(s, "rawdatafile");
    sink_string(s, pad);
  }
  sink_string(s, "rawdatafile");
  sink_string(s, hdr->rawdatafile);
  sink_string(s, "HEADER_END");
}

ssize_t fb_fd_write_padded_header(int fd, const fb_hdr_t *hdr, int32_t minlen)
{
  fb_sink s = {fd, NULL, 0};
  write_header(&s, hdr, minlen);
  return s.n;
}

ssize_t fb_fd_write_header(int fd, const fb_hdr_t *hdr) { return fb_fd_write_padded_header(fd, hdr, 0); }

void *fb_buf_write_padded_header(void *buf, const fb_hdr_t *hdr, int32_t minlen)
{
  fb_sink s = {-1, buf, 0};
  write_header(&s, hdr, minlen);
  return s.p;
}

void *fb_buf_write_header(void *buf, const fb_hdr_t *hdr) { return fb_buf_write_padded_header(buf, hdr, 0); }

/* ---- value readers -------------------------------------------------------- */

int32_t fb_fd_read_int(int fd, int32_t *i)
{
  uint32_t le = 0;
  int32_t v;
  (void)!read(fd, &le, sizeof le);
  v = (int32_t)le32toh(le);
  if (i) *i = v;
  return v;
}

void *fb_buf_read_int(void *buf, int32_t *i)
{
  if (i) {
    uint32_t le;
    memcpy(&le, buf, sizeof le);
    *i = (int32_t)le32toh(le);
  }
  return (char *)buf + sizeof(int32_t);
}

static double bits_to_double(uint64_t le)
{
  const uint64_t bits = le64toh(le);
  double d;
  memcpy(&d, &bits, sizeof d);
  return d;
}

double fb_fd_read_double(int fd, double *d)
{
  uint64_t le = 0;
  double v;
  (void)!read(fd, &le, sizeof le);
  v = bits_to_double(le);
  if (d) *d = v;
  return v;
}

void *fb_buf_read_double(void *buf, double *d)
{
  if (d) {
    uint64_t le;
    memcpy(&le, buf, sizeof le);
    *d = bits_to_double(le);
  }
  return (char *)buf + sizeof(uint64_t);
}

double fb_fd_read_angle(int fd, double *d)
{
  const double v = fb_dms_to_ddd(fb_fd_read_double(fd, NULL));
  if (d) *d = v;
  return v;
}

void *fb_buf_read_angle(void *buf, double *d)
{
  buf = fb_buf_read_double(buf, d);
  if (d) *d = fb_dms_to_ddd(*d);
  return buf;
}

ssize_t fb_fd_read_string(int fd, char *c, int32_t *n)
{
  uint32_t le = 0;
  int32_t total, want;
  ssize_t got;
  if (!c || !n) return -1;
  (void)!read(fd, &le, sizeof le);
  total = (int32_t)le32toh(le);
  want = total > *n ? *n : total;
  got = read(fd, c, (size_t)want);
  if (got == -1) return -1;
  if (got < *n) c[got] = '\0';
  if (got < total) lseek(fd, total - got, SEEK_CUR);
  *n = (int32_t)got;
  return got;
}

void *fb_buf_peek_string(void *buf, char **c, int32_t *n)
{
  if (!c || !n) return buf;
  *c = (char *)fb_buf_read_int(buf, n);
  return *c + *n;
}

void *fb_buf_read_string(void *buf, char *c, int32_t *n)
{
  char *s;
  int32_t len;
  if (!c || !n) return buf;
  buf = fb_buf_peek_string(buf, &s, &len);
  if (len > *n) len = *n;
  memcpy(c, s, (size_t)len);
  if (len < *n) c[len] = '\0';
  *n = len;
  return buf;
}

/* ---- header readers ------------------------------------------------------- */

enum { V_INT, V_DOUBLE, V_ANGLE, V_STRING, V_SKIP_INT, V_SKIP_DOUBLE, V_NONE };

/* Keywords in the order the reference's reader tries them; a keyword read
 * from the header selects the first entry it is a prefix of (as the
 * reference compares only the keyword's own length). */
static const struct {
  const char *name;
  int kind;
  size_t off;
} kFields[] = {
    {"machine_id", V_INT, offsetof(fb_hdr_t, machine_id)},
    {"telescope_id", V_INT, offsetof(fb_hdr_t, telescope_id)},
    {"data_type", V_INT, offsetof(fb_hdr_t, data_type)},
    {"barycentric", V_INT, offsetof(fb_hdr_t, barycentric)},
    {"pulsarcentric", V_INT, offsetof(fb_hdr_t, pulsarcentric)},
    {"src_raj", V_ANGLE, offsetof(fb_hdr_t, src_raj)},
    {"src_dej", V_ANGLE, offsetof(fb_hdr_t, src_dej)},
    {"az_start", V_DOUBLE, offsetof(fb_hdr_t, az_start)},
    {"za_start", V_DOUBLE, offsetof(fb_hdr_t, za_start)},
    {"fch1", V_DOUBLE, offsetof(fb_hdr_t, fch1)},
    {"foff", V_DOUBLE, offsetof(fb_hdr_t, foff)},
    {"nchans", V_INT, offsetof(fb_hdr_t, nchans)},
    {"nbeams", V_INT, offsetof(fb_hdr_t, nbeams)},
    {"ibeam", V_INT, offsetof(fb_hdr_t, ibeam)},
    {"nbits", V_INT, offsetof(fb_hdr_t, nbits)},
    {"tstart", V_DOUBLE, offsetof(fb_hdr_t, tstart)},
    {"tsamp", V_DOUBLE, offsetof(fb_hdr_t, tsamp)},
    {"nifs", V_INT, offsetof(fb_hdr_t, nifs)},
    {"source_name", V_STRING, offsetof(fb_hdr_t, source_name)},
    {"rawdatafile", V_STRING, offsetof(fb_hdr_t, rawdatafile)},
    /* SIGPROC keywords the header structure does not hold */
    {"nsamples", V_SKIP_INT, 0},
    {"npuls", V_SKIP_INT, 0},
    {"nbins", V_SKIP_INT, 0},
    {"refdm", V_SKIP_DOUBLE, 0},
    {"period", V_SKIP_DOUBLE, 0},
    {"fchannel", V_SKIP_DOUBLE, 0},
};

/* Parses a header from buf; `cap` bounds the bytes examined (0: unbounded).
 * Returns the position after HEADER_END, or buf when buf holds no header. */
static char *read_header(char *buf, size_t cap, fb_hdr_t *hdr, size_t *hdr_len)
{
  char *p, *kw;
  int32_t len;
#define FITS(q, k) (cap == 0 || ((size_t)((q) - buf) + (size_t)(k) <= cap))
  memset(hdr, 0, sizeof *hdr);
  if (!FITS(buf, 4)) return buf;
  p = (char *)fb_buf_peek_string(buf, &kw, &len);
  if (len < 0 || !FITS(kw, len) || strncmp(kw, "HEADER_START", (size_t)len) != 0) return buf;
  if (!FITS(p, 4)) return buf;
  p = (char *)fb_buf_peek_string(p, &kw, &len);
  while (1) {
    size_t f;
    if (len < 0 || !FITS(kw, len)) return buf;
    if (strncmp(kw, "HEADER_END", (size_t)len) == 0) break;
    for (f = 0; f < sizeof kFields / sizeof kFields[0]; ++f)
      if (strncmp(kw, kFields[f].name, (size_t)len) == 0) break;
    if (f < sizeof kFields / sizeof kFields[0]) {
      char *dst = (char *)hdr + kFields[f].off;
      switch (kFields[f].kind) {
        case V_INT:
        case V_SKIP_INT:
          if (!FITS(p, 4)) return buf;
          p = (char *)fb_buf_read_int(p, kFields[f].kind == V_INT ? (int32_t *)dst : NULL);
          break;
        case V_DOUBLE:
        case V_SKIP_DOUBLE:
          if (!FITS(p, 8)) return buf;
          p = (char *)fb_buf_read_double(p, kFields[f].kind == V_DOUBLE ? (double *)dst : NULL);
          break;
        case V_ANGLE:
          if (!FITS(p, 8)) return buf;
          p = (char *)fb_buf_read_angle(p, (double *)dst);
          break;
        case V_STRING: {
          int32_t n = 80;
          char *s;
          int32_t sl;
          if (!FITS(p, 4)) return buf;
          fb_buf_peek_string(p, &s, &sl);
          if (sl < 0 || !FITS(s, sl)) return buf;
          p = (char *)fb_buf_read_string(p, dst, &n);
          dst[n] = '\0';
          break;
        }
      }
    }
    if (!FITS(p, 4)) return buf;
    p = (char *)fb_buf_peek_string(p, &kw, &len);
  }
#undef FITS
  if (hdr_len) *hdr_len = (size_t)(p - buf);
  return p;
}

void *fb_buf_read_header(void *buf, fb_hdr_t *hdr, size_t *hdr_len)
{
  if (!buf || !hdr || !hdr_len) return buf;
  return read_header((char *)buf, 0, hdr, hdr_len);
}

ssize_t fb_fd_read_header(int fd, fb_hdr_t *hdr, size_t *hdr_len)
{
  char buf[4096];
  ssize_t got;
  size_t len;
  char *end;
  if (!hdr) return -1;
  got = read(fd, buf, sizeof buf);
  if (got < 0) return -1;
  len = (size_t)got;
  end = read_header(buf, (size_t)got, hdr, &len);
  if (end != buf) {
    lseek(fd, (off_t)len - (off_t)got, SEEK_CUR);
    if (hdr_len) *hdr_len = len;
  }
  return (ssize_t)len;
}

/* ---- telescopes ----------------------------------------------------------- */

int fb_telescope_id(const char *telescope_name)
{
  /* SIGPROC's telescope ids (its aliases table); ATA shares 9 with 140FT */
  static const struct { const char *name; int id; } kIds[] = {
      {"FAKE", 0}, {"ARECIBO", 1}, {"OOTY", 2}, {"NANCAY", 3}, {"PARKES", 4},
      {"JODRELL", 5}, {"GBT", 6}, {"GMRT", 7}, {"EFFELSBERG", 8}, {"140FT", 9},
      {"ATA", 9}, {"SRT", 10}, {"LEUSCHNER", 11}, {"MEERKAT", 64},
  };
  size_t i;
  for (i = 0; i < sizeof kIds / sizeof kIds[0]; ++i)
    if (strcasecmp(telescope_name, kIds[i].name) == 0) return kIds[i].id;
  return -1;
}
