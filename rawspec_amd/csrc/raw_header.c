/*
 * raw_header.c -- GUPPI RAW header utilities (rawspec_raw_*), the functions
 * the reference's librawspec.so exports from rawspec_rawutils.c (which reads
 * cards through the WCSTools hget routines, hget.h).  Host-only C.
 *
 * Card lookup and value rules (restated from what hget.h's ksearch / hgetc
 * do for the cards rawspec reads):
 *  * a keyword matches a card when it starts, compared case-insensitively,
 *    in one of the card's first 8 columns, preceded only by blanks and
 *    followed by '=' or a blank or non-printing character; the first
 *    matching card counts;
 *  * the value is the text between the first pair of single (else double)
 *    quotes when the quote precedes any '/', otherwise the text after the
 *    first '=' (column 10 when there is none) up to the first '/' or
 *    column 79; leading blanks and trailing blanks / CR are dropped and "-0"
 *    reads as "0";
 *  * numbers: an optional leading '#' is skipped; s32 rounds the decimal
 *    value toward zero after a 0.001 nudge away from zero and saturates; u32,
 *    s64, u64 use strtoul / strtoll / strtoull with base 0 (so 0x.. is hex and
 *    a leading 0 octal) and fall back to the decimal value when characters
 *    trail the number; a Fortran 'D' / 'd' exponent is read as 'e' in values
 *    that look like floating-point numbers.
 * Deliberate differences: the search stops at the END card, the first NUL
 * byte or 256000 bytes (hget searches up to the first NUL anywhere, i.e.
 * into the data block that follows a header read from a file); "[n]" token
 * suffixes on keywords are not supported (rawspec never uses them).
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <unistd.h>

#include "rawspec_rawutils.h"

#define CARD 80
#define VALUE_MAX 81          /* longest value kept (as hget's VLENGTH) */
#define SCAN_MAX 256000       /* longest header searched */

/* Length of the searchable header: up to a NUL byte or SCAN_MAX bytes, and
 * no further than the end of the END card. */
static size_t header_extent(const char *buf, size_t cap)
{
  size_t n = 0;
  while (n < cap && buf[n] != '\0') {
    if (n % CARD == 0 && n + 4 <= cap && strncmp(buf + n, "END ", 4) == 0) {
      return n + CARD < cap ? n + CARD : cap;
    }
    ++n;
  }
  return n;
}

/* Start of the card holding keyword key, or NULL. */
static const char *find_card(const char *buf, size_t len, const char *key)
{
  const size_t lk = strlen(key);
  size_t card;
  if (lk == 0 || lk > 8) return NULL;
  for (card = 0; card < len; card += CARD) {
    size_t col;
    for (col = 0; col < 8 && card + col + lk <= len; ++col) {
      const char *p = buf + card + col;
      unsigned char next;
      size_t i;
      if (strncasecmp(p, key, lk) != 0) {
        if (p[0] != ' ') break;       /* only blanks may precede the keyword */
        continue;
      }
      next = card + col + lk < len ? (unsigned char)p[lk] : 0;
      if (next != '=' && next > 32 && next < 127) break;   /* a longer keyword */
      for (i = 0; i < col; ++i)
        if (buf[card + i] != ' ') break;
      if (i == col) return buf + card;
      break;
    }
  }
  return NULL;
}

/* Copies the value of the card at `card` (at most len bytes remain) into
 * out (VALUE_MAX + 1 bytes). */
static void card_value(const char *card, size_t len, char *out)
{
  char line[CARD + 1];
  size_t n = len < CARD ? len : CARD;
  const char *q1, *q2 = NULL, *c1, *v1, *v2;
  size_t i;
  for (i = 0; i < n && card[i] != '\0'; ++i) line[i] = card[i];
  for (; i <= CARD; ++i) line[i] = '\0';

  c1 = strchr(line, '/');
  q1 = strchr(line, '\'');
  if (!q1) q1 = strchr(line, '"');
  if (q1 && c1 && c1 < q1) q1 = NULL;          /* a quote inside the comment */
  if (q1) {
    const char quote = *q1;
    q2 = strchr(q1 + 1, quote);
    if (!q2) {
      /* unterminated: up to the comment or the card end, blanks dropped */
      q2 = c1 ? c1 - 1 : line + CARD - 1;
      while (q2 > q1 && *q2 == ' ') --q2;
      ++q2;
    }
    v1 = q1 + 1;
    v2 = q2;
  } else {
    const char *eq = strchr(line, '=');
    v1 = eq ? eq + 1 : line + 9;
    v2 = c1 ? c1 : line + CARD - 1;
  }
  while (v1 < v2 && *v1 == ' ') ++v1;
  while (v2 > v1 && (v2[-1] == ' ' || v2[-1] == '\r')) --v2;
  if (v2 < v1) v2 = v1;
  if (v2 - v1 == 2 && v1[0] == '-' && v1[1] == '0') ++v1;   /* "-0" -> "0" */
  n = (size_t)(v2 - v1);
  if (n > VALUE_MAX) n = VALUE_MAX;
  memcpy(out, v1, n);
  out[n] = '\0';
}

/* Value of key as text (1) or 0 when absent. */
static int lookup(const char *buf, size_t cap, const char *key, char *out)
{
  const size_t len = header_extent(buf, cap);
  const char *card = find_card(buf, len, key);
  if (!card) return 0;
  card_value(card, len - (size_t)(card - buf), out);
  return 1;
}

/* Whether s looks like a floating-point number whose Fortran exponent
 * letter should be read as 'e': digits with '.', 'd' or 'e', signs only at
 * the start or after an exponent letter, a colon or a blank, and no colon
 * (sexagesimal values are not rewritten). */
static int float_like(const char *s)
{
  size_t n = strlen(s), i;
  int digits = 0, colon = 0, fp = 0;
  if (n == 0 || strchr("DdEe", s[0])) return 0;
  while (n > 0 && s[n - 1] == ' ') --n;
  for (i = 0; i < n; ++i) {
    const char c = s[i];
    if (c == '\n') break;
    if (c == ' ' && digits == 0) continue;
    if (!isdigit((unsigned char)c) && !strchr("+-DdEe:.", c)) return 0;
    if (c == '+' || c == '-') {
      if (s[i + 1] == '+' || s[i + 1] == '-') return 0;
      if (i > 0 && !strchr("DdEe: ", s[i - 1])) return 0;
    } else if (isdigit((unsigned char)c)) {
      ++digits;
    } else if (c == ':') {
      ++colon;
    }
    if (c == '.' || c == 'd' || c == 'e') fp = 1;
  }
  return digits > 0 && !colon && fp;
}

/* Decimal value of a numeric card value (after the optional '#'). */
static double decimal_value(const char *v)
{
  char tmp[VALUE_MAX + 1];
  char *p;
  strncpy(tmp, v, VALUE_MAX);
  tmp[VALUE_MAX] = '\0';
  if (float_like(tmp)) {
    if ((p = strchr(tmp, 'D'))) *p = 'e';
    if ((p = strchr(tmp, 'd'))) *p = 'e';
  }
  return atof(tmp);
}

static const char *skip_hash(const char *v) { return v[0] == '#' ? v + 1 : v; }

static void trailing_warning(const char *key, const char *end)
{
  fprintf(stderr, "librawspec: %s: invalid integer character '%c' (%d)\n", key, end[0], end[0]);
}

int32_t rawspec_raw_get_s32(const char *buf, const char *key, int32_t def)
{
  char v[VALUE_MAX + 1];
  double d;
  if (!lookup(buf, SCAN_MAX, key, v)) return def;
  d = decimal_value(skip_hash(v));
  if (d + 0.001 > INT_MAX) return INT_MAX;
  if (d >= 0) return (int32_t)(d + 0.001);
  if (d - 0.001 < (double)INT_MIN) return INT_MIN;
  return (int32_t)(d - 0.001);
}

uint32_t rawspec_raw_get_u32(const char *buf, const char *key, uint32_t def)
{
  char v[VALUE_MAX + 1];
  const char *s;
  char *end = NULL;
  uint32_t r;
  if (!lookup(buf, SCAN_MAX, key, v)) return def;
  s = skip_hash(v);
  r = (uint32_t)strtoul(s, &end, 0);
  if (end && end[0]) {
    trailing_warning(key, end);
    r = (uint32_t)atof(s);
  }
  return r;
}

int64_t rawspec_raw_get_s64(const char *buf, const char *key, int64_t def)
{
  char v[VALUE_MAX + 1];
  const char *s;
  char *end = NULL;
  int64_t r;
  if (!lookup(buf, SCAN_MAX, key, v)) return def;
  s = skip_hash(v);
  r = (int64_t)strtoll(s, &end, 0);
  if (end && end[0]) {
    trailing_warning(key, end);
    r = (int64_t)atof(s);
  }
  return r;
}

uint64_t rawspec_raw_get_u64(const char *buf, const char *key, uint64_t def)
{
  char v[VALUE_MAX + 1];
  const char *s;
  char *end = NULL;
  uint64_t r;
  if (!lookup(buf, SCAN_MAX, key, v)) return def;
  s = skip_hash(v);
  r = (uint64_t)strtoull(s, &end, 0);
  if (end && end[0]) {
    trailing_warning(key, end);
    r = (uint64_t)atof(s);
  }
  return r;
}

double rawspec_raw_get_dbl(const char *buf, const char *key, double def)
{
  char v[VALUE_MAX + 1];
  if (!lookup(buf, SCAN_MAX, key, v)) return def;
  return decimal_value(skip_hash(v));
}

void rawspec_raw_get_str(const char *buf, const char *key, const char *def, char *out,
                         size_t len)
{
  char v[VALUE_MAX + 1];
  size_t n;
  if (len == 0) return;
  if (!lookup(buf, SCAN_MAX, key, v)) {
    strncpy(out, def, len);
    out[len - 1] = '\0';
    return;
  }
  n = strlen(v);
  if (n < len) {
    memcpy(out, v, n + 1);
  } else if (len > 1) {
    memcpy(out, v, len - 1);    /* truncated, unterminated (as the reference) */
  } else {
    out[0] = v[0];
  }
}

double rawspec_raw_dmsstr_to_d(char *dmsstr)
{
  double sign = 1.0, d = 0.0, scale = 1.0;
  char *p = dmsstr;
  int field = 0;
  if (*p == '-') {
    sign = -1.0;
    ++p;
  } else if (*p == '+') {
    ++p;
  }
  /* up to three ':'-separated fields (empty fields are skipped): d + m/60 + s/3600 */
  while (*p && field < 3) {
    char *end;
    if (*p == ':') {
      ++p;
      continue;
    }
    end = p;
    while (*end && *end != ':') ++end;
    {
      const char c = *end;
      *end = '\0';
      d += strtod(p, NULL) / scale;
      *end = c;
    }
    scale *= 60.0;
    ++field;
    p = end;
  }
  return sign * d;
}

int rawspec_raw_header_size(char *hdr, size_t len, int directio)
{
  int i;
  for (i = 0; (size_t)i < len; i += CARD) {
    if ((size_t)i + 4 <= len && strncmp(hdr + i, "END ", 4) == 0) {
      i += CARD;
      if (directio) i += (MAX_RAW_HDR_SIZE - i) % 512;
      return i;
    }
  }
  return 0;
}

/* Parse with an explicit buffer length (the file reader's buffer is not
 * NUL-terminated). */
static void parse_header_n(const char *buf, size_t cap, rawspec_raw_hdr_t *h)
{
  char tmp[80];
  {
    /* the typed getters scan up to a NUL; work on a terminated copy */
    char *t = (char *)malloc(cap + 1);
    if (!t) {
      memset(h, 0, sizeof(*h));
      return;
    }
    memcpy(t, buf, cap);
    t[cap] = '\0';
    h->blocsize = (size_t)rawspec_raw_get_s32(t, "BLOCSIZE", 0);
    h->npol = (unsigned)rawspec_raw_get_s32(t, "NPOL", 0);
    h->obsnchan = (unsigned)rawspec_raw_get_s32(t, "OBSNCHAN", 0);
    h->nbits = rawspec_raw_get_u32(t, "NBITS", 8);
    h->obsfreq = rawspec_raw_get_dbl(t, "OBSFREQ", 0.0);
    h->obsbw = rawspec_raw_get_dbl(t, "OBSBW", 0.0);
    h->tbin = rawspec_raw_get_dbl(t, "TBIN", 0.0);
    h->directio = rawspec_raw_get_s32(t, "DIRECTIO", 0);
    h->pktidx = (int64_t)rawspec_raw_get_u64(t, "PKTIDX", (uint64_t)-1);
    h->beam_id = rawspec_raw_get_s32(t, "BEAM_ID", -1);
    h->nbeam = rawspec_raw_get_s32(t, "NBEAM", -1);
    h->refbeam = rawspec_raw_get_s32(t, "REFBEAM", -1);
    h->nants = rawspec_raw_get_u32(t, "NANTS", 1);
    rawspec_raw_get_str(t, "RA_STR", "0.0", tmp, sizeof tmp);
    h->ra = rawspec_raw_hmsstr_to_h(tmp);
    rawspec_raw_get_str(t, "DEC_STR", "0.0", tmp, sizeof tmp);
    h->dec = rawspec_raw_dmsstr_to_d(tmp);
    h->mjd = (double)rawspec_raw_get_s32(t, "STT_IMJD", 51545) +
             (double)rawspec_raw_get_s32(t, "STT_SMJD", 0) / 86400.0;
    rawspec_raw_get_str(t, "SRC_NAME", "Unknown", h->src_name, 80);
    rawspec_raw_get_str(t, "TELESCOP", "Unknown", h->telescop, 80);
    free(t);
  }
}

void rawspec_raw_parse_header(const char *buf, rawspec_raw_hdr_t *raw_hdr)
{
  parse_header_n(buf, header_extent(buf, SCAN_MAX), raw_hdr);
}

off_t rawspec_raw_read_header(int fd, rawspec_raw_hdr_t *raw_hdr)
{
  /* 512-byte aligned so files opened with O_DIRECT can be read */
  static __thread char hdr[MAX_RAW_HDR_SIZE] __attribute__((aligned(512)));
  const off_t pos = lseek(fd, 0, SEEK_CUR);
  const ssize_t got = read(fd, hdr, MAX_RAW_HDR_SIZE);
  int full;
  if (got == -1) return -1;
  if (got < CARD) return 0;
  parse_header_n(hdr, header_extent(hdr, (size_t)got), raw_hdr);
  if (raw_hdr->blocsize == 0) { fprintf(stderr, " BLOCSIZE not found in header\n"); return -1; }
  if (raw_hdr->npol == 0) { fprintf(stderr, "NPOL not found in header\n"); return -1; }
  if (raw_hdr->obsnchan == 0) { fprintf(stderr, "OBSNCHAN not found in header\n"); return -1; }
  if (raw_hdr->obsfreq == 0.0) { fprintf(stderr, "OBSFREQ not found in header\n"); return -1; }
  if (raw_hdr->obsbw == 0.0) { fprintf(stderr, "OBSBW not found in header\n"); return -1; }
  if (raw_hdr->tbin == 0.0) { fprintf(stderr, "TBIN not found in header\n"); return -1; }
  if (raw_hdr->pktidx == -1) { fprintf(stderr, "PKTIDX not found in header\n"); return -1; }
  /* NPOL 4 counts the cross products of 2 polarisations */
  if (raw_hdr->npol == 4) raw_hdr->npol = 2;
  raw_hdr->hdr_pos = pos;
  raw_hdr->hdr_size = (size_t)rawspec_raw_header_size(hdr, (size_t)got, 0);
  full = rawspec_raw_header_size(hdr, (size_t)got, raw_hdr->directio);
  return lseek(fd, pos + full, SEEK_SET);
}
