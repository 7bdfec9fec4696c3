/*
 * rawspec_rawutils.h -- GUPPI RAW header utilities exported by librawspec.
 *
 * Same types, names and arguments as the reference's rawspec_rawutils.h
 * (rawspec_rawutils.h:1-70), so a client compiled against either links
 * against this library.  A GUPPI RAW header is a sequence of 80-byte
 * FITS-style cards "KEYWORD = value / comment" ended by an "END" card,
 * optionally padded to a 512-byte boundary (DIRECTIO), followed by one data
 * block of BLOCSIZE bytes.
 */
#ifndef _RAWUTILS_H_
#define _RAWUTILS_H_

#include <stdint.h>
#include <sys/types.h>

/* Observation parameters of one RAW block (reference rawspec_rawutils.h:7-28). */
typedef struct {
  int directio;          /* DIRECTIO: header padded to a multiple of 512 bytes */
  size_t blocsize;       /* BLOCSIZE: data bytes that follow the header */
  unsigned int npol;     /* NPOL (4, the number of cross products, read as 2) */
  unsigned int obsnchan; /* OBSNCHAN: coarse channels (all antennas) */
  unsigned int nbits;    /* NBITS (default 8) */
  int64_t pktidx;        /* PKTIDX: packet index of the block's first sample */
  double obsfreq;        /* OBSFREQ: centre frequency (MHz) */
  double obsbw;          /* OBSBW: bandwidth (MHz, signed) */
  double tbin;           /* TBIN: sample time (s) */
  double ra;             /* RA_STR in hours */
  double dec;            /* DEC_STR in degrees */
  double mjd;            /* STT_IMJD + STT_SMJD / 86400 */
  int beam_id;           /* BEAM_ID, -1 if absent */
  int nbeam;             /* NBEAM, -1 if absent */
  int refbeam;           /* REFBEAM, -1 if absent */
  unsigned int nants;    /* NANTS (default 1) */
  char src_name[81];     /* SRC_NAME (default "Unknown") */
  char telescop[81];     /* TELESCOP (default "Unknown") */
  off_t hdr_pos;         /* file offset of the header */
  size_t hdr_size;       /* header bytes including END, excluding padding */
} rawspec_raw_hdr_t;

/* Largest header read: a multiple of both 80 and 512. */
#define MAX_RAW_HDR_SIZE (25600)

#ifdef __cplusplus
extern "C" {
#endif

/* Value of keyword `key` in the header `buf` converted to the return type,
 * or `def` when the keyword is absent (reference rawspec_rawutils.c:10-83).
 * The header is read up to its END card or the first NUL byte. */
int32_t rawspec_raw_get_s32(const char *buf, const char *key, int32_t def);
uint32_t rawspec_raw_get_u32(const char *buf, const char *key, uint32_t def);
int64_t rawspec_raw_get_s64(const char *buf, const char *key, int64_t def);
uint64_t rawspec_raw_get_u64(const char *buf, const char *key, uint64_t def);
double rawspec_raw_get_dbl(const char *buf, const char *key, double def);

/* String value of `key` into out (len bytes), or def (rawspec_rawutils.c:85-92). */
void rawspec_raw_get_str(const char *buf, const char *key, const char *def, char *out,
                         size_t len);

/* "[+-]d[:m[:s]]" -> decimal degrees (or hours); modifies dmsstr
 * (rawspec_rawutils.c:94-131). */
double rawspec_raw_dmsstr_to_d(char *dmsstr);

#define rawspec_raw_hmsstr_to_h(hmsstr) (rawspec_raw_dmsstr_to_d(hmsstr))

/* Bytes from the header start to just past the END card (plus the DIRECTIO
 * padding when directio != 0); 0 if no END card lies within len bytes
 * (rawspec_rawutils.c:133-152). */
int rawspec_raw_header_size(char *hdr, size_t len, int directio);

/* Parses the rawspec-related parameters of a header (rawspec_rawutils.c:155-187). */
void rawspec_raw_parse_header(const char *buf, rawspec_raw_hdr_t *raw_hdr);

/* Reads and parses the header at fd's current offset and leaves fd at the
 * start of the following data block, whose offset it returns; 0 at end of
 * file, -1 on error or when a required keyword is missing
 * (rawspec_rawutils.c:194-261). */
off_t rawspec_raw_read_header(int fd, rawspec_raw_hdr_t *raw_hdr);

#ifdef __cplusplus
}
#endif

#endif
