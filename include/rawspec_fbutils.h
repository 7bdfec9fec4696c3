/*
 * rawspec_fbutils.h -- SIGPROC filterbank header utilities exported by
 * librawspec (same types, names and arguments as the reference's
 * rawspec_fbutils.h:1-174, so clients such as the rawspec CLI link against
 * this library).
 *
 * A filterbank header is a sequence of keywords, each written as a 4-byte
 * little-endian length followed by that many ASCII characters, and each
 * followed by its value: a 4-byte little-endian integer, an 8-byte
 * little-endian IEEE double, or another length-prefixed string.  It opens
 * with HEADER_START and closes with HEADER_END.  Angles (src_raj, src_dej)
 * are stored as sexagesimal numbers ddmmss.s.
 */
#ifndef _FBUTILS_H_
#define _FBUTILS_H_

#include <stdint.h>
#include <sys/types.h>

/* Header fields (reference rawspec_fbutils.h:70-122). */
typedef struct {
  int machine_id;       /* 0 = fake data, 1 = Arecibo, 2 = Ooty, ... */
  int telescope_id;     /* see fb_telescope_id() */
  int data_type;        /* 1 = filterbank, 2 = time series, ... */
  int barycentric;      /* written only when non-zero */
  int pulsarcentric;    /* written only when non-zero */
  double src_raj;       /* hours (stored as hhmmss.s) */
  double src_dej;       /* degrees (stored as ddmmss.s) */
  double az_start;      /* degrees */
  double za_start;      /* degrees */
  double fch1;          /* centre frequency of the first fine channel (MHz) */
  double foff;          /* fine channel width (MHz) */
  int nchans;           /* fine channels */
  int nbeams;
  int ibeam;
  int nbits;            /* bits per sample */
  double tstart;        /* MJD of the first sample */
  double tsamp;         /* sample interval (s) */
  int nifs;             /* IF channels (polarisation products) */
  char source_name[81];
  char rawdatafile[81];
  int nfpc;             /* fine channels per coarse channel (FBH5) */
  int refbeam;          /* beam that tracked the source (FBH5) */
} fb_hdr_t;

#ifdef __cplusplus
extern "C" {
#endif

/* degrees <-> sexagesimal ddmmss.s (rawspec_fbutils.c:14-40) */
double fb_ddd_to_dms(double ddd);
double fb_dms_to_ddd(double dms);

/* Value writers: to a file descriptor (returning the bytes written) or to a
 * buffer (returning the position after the value).  Strings are limited to
 * 80 characters (rawspec_fbutils.c:44-96). */
ssize_t fb_fd_write_int(int fd, int32_t i);
void *fb_buf_write_int(void *buf, int32_t i);
ssize_t fb_fd_write_double(int fd, double d);
void *fb_buf_write_double(void *buf, double d);
ssize_t fb_fd_write_angle(int fd, double d);
void *fb_buf_write_angle(void *buf, double d);
ssize_t fb_fd_write_string(int fd, const char *c);
void *fb_buf_write_string(void *buf, const char *c);

/* Value readers (rawspec_fbutils.c:98-232).  The string readers copy at most
 * *n characters, NUL-terminate when room remains and store the count in *n;
 * fb_buf_peek_string points *c at the (unterminated) string in buf. */
int32_t fb_fd_read_int(int fd, int32_t *i);
void *fb_buf_read_int(void *buf, int32_t *i);
double fb_fd_read_double(int fd, double *d);
void *fb_buf_read_double(void *buf, double *d);
double fb_fd_read_angle(int fd, double *d);
void *fb_buf_read_angle(void *buf, double *d);
ssize_t fb_fd_read_string(int fd, char *c, int32_t *n);
void *fb_buf_peek_string(void *buf, char **c, int32_t *n);
void *fb_buf_read_string(void *buf, char *c, int32_t *n);

/* Header writers: the keywords in the reference's order; the padded forms
 * pad the header to about minlen bytes with extra "rawdatafile" entries
 * before the real one (rawspec_fbutils.c:234-410).  fb_fd_* return the bytes
 * written, fb_buf_* the position after the header. */
ssize_t fb_fd_write_padded_header(int fd, const fb_hdr_t *hdr, int32_t minlen);
ssize_t fb_fd_write_header(int fd, const fb_hdr_t *hdr);
void *fb_buf_write_padded_header(void *buf, const fb_hdr_t *hdr, int32_t minlen);
void *fb_buf_write_header(void *buf, const fb_hdr_t *hdr);

/* Header readers: parse a header, storing its length in *hdr_len; unknown
 * keywords are skipped (rawspec_fbutils.c:412-521). */
ssize_t fb_fd_read_header(int fd, fb_hdr_t *hdr, size_t *hdr_len);
void *fb_buf_read_header(void *buf, fb_hdr_t *hdr, size_t *hdr_len);

/* SIGPROC telescope id of a telescope name, -1 if unknown (rawspec_fbutils.c:523-566). */
int fb_telescope_id(const char *telescope_name);

#ifdef __cplusplus
}
#endif

#endif
