/*
 * rawspec_oracle.c -- FP64 CPU restatement of librawspec's spectrometer
 * semantics.  TEST INFRASTRUCTURE ONLY (checker for tests/, the smoke check
 * and the CPU-baseline leg of bench.py); the product library never links it.
 *
 * Every function cites the reference lines whose behaviour it restates.
 * Arithmetic is IEEE double throughout; twiddles come from libm cos/sin of
 * exactly reduced angles, so the oracle's own error is ~1e-15 relative.
 *
 * Built with -DRO_REAL=float (liboracle32.so) the transforms, detection and
 * sums run in single precision instead: the same algorithm as an FP32 CPU
 * port, which the CPU baseline times beside the FP64 one.  The oracle the
 * tests check against is always the double build.
 */
#include "rawspec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define RO_ABS(x) ((x) < 0 ? -(x) : (x))

#ifndef RO_REAL
#define RO_REAL double
#endif
typedef RO_REAL ro_real;   /* working precision of transforms and sums */

/* ------------------------------------------------------------------------ */
/* Validation and derived counts: rawspec_gpu.cu:471-641 (checks, in the    */
/* same order) and :732-760 (Nss, Nds, Nis, h_pwrbuf_size).                  */
/* ------------------------------------------------------------------------ */
int ro_derive(ro_cfg *c)
{
  unsigned i;
  /* :472 No in [1, MAX_OUTPUTS] */
  if (c->No == 0 || c->No > RO_MAX_OUTPUTS) return RO_E_NO;
  /* :480 Np in [1, 2] */
  if (c->Np == 0 || c->Np > 2) return RO_E_NP;
  /* :488-492 Npolout forced to 1 unless |Npolout| == 4 and Np == 2 */
  for (i = 0; i < c->No; i++)
    if (RO_ABS(c->Npolout[i]) != 4 || c->Np != 2) c->Npolout[i] = 1;
  /* :495 Ntpb != 0 */
  if (c->Ntpb == 0) return RO_E_NTPB;
  /* :506-516 Nbps: 0 -> 8; 4 -> 8 with expansion; else 8 or 16 */
  if (c->Nbps == 0) c->Nbps = 8;
  if (c->Nbps != 4 && c->Nbps != 8 && c->Nbps != 16) return RO_E_NBPS;
  c->expand4 = (c->Nbps == 4);
  if (c->Nbps == 4) c->Nbps = 8;
  /* :519-539 Ntmax, every Nt non-zero and dividing Ntmax */
  c->Ntmax = 0;
  for (i = 0; i < c->No; i++) {
    if (c->Nts[i] == 0) return RO_E_NT0;
    if (c->Ntmax < c->Nts[i]) c->Ntmax = c->Nts[i];
  }
  for (i = 0; i < c->No; i++)
    if (c->Ntmax % c->Nts[i] != 0) return RO_E_NTDIV;
  /* :543-574 Nb given (Ntmax | Nb*Ntpb) or derived */
  if (c->Nb != 0) {
    if (((unsigned long long)c->Nb * c->Ntpb) % c->Ntmax != 0) return RO_E_NB;
  } else if (c->Ntmax < c->Ntpb) {
    if (c->Ntpb % c->Ntmax != 0) return RO_E_NTPBDIV;
    c->Nb = 1;
  } else {
    if (c->Ntmax % c->Ntpb != 0) return RO_E_NTMAXDIV;
    c->Nb = c->Ntmax / c->Ntpb;
  }
  /* :577-584 Nb_host */
  if (c->Nb_host == 0 && c->have_blkbufs) return RO_E_NBHOST;
  if (c->Nb_host == 0) c->Nb_host = c->Nb;
  /* :587-613 Nas: non-zero; Nt*Na and Nb*Ntpb divide one another */
  {
    unsigned long long nbuf = (unsigned long long)c->Nb * c->Ntpb;
    for (i = 0; i < c->No; i++) {
      unsigned long long ntna = (unsigned long long)c->Nts[i] * c->Nas[i];
      if (c->Nas[i] == 0) return RO_E_NA0;
      if (ntna < nbuf) { if (nbuf % ntna != 0) return RO_E_NADIV; }
      else             { if (ntna % nbuf != 0) return RO_E_NADIV; }
    }
  }
  /* :615 Nant defaults to 1 */
  if (c->Nant == 0) c->Nant = 1;
  /* :617-641 channel batching (validated only; results do not depend on it) */
  if (c->Nbc == 0) {
    c->Nbc = c->Nc;
  } else {
    if (c->Nbc <= 1) {
      if (c->Nant > 1) {
        c->Nbc = c->Nc / c->Nant;
      } else {
        for (i = 1; i <= 10; i++)
          if (c->Nc % i == 0) c->Nbc = c->Nc / i;
      }
    }
    if (c->Nbc == 0 || c->Nc % c->Nbc != 0) return RO_E_NBC;
  }
  /* :732-760 derived counts and dump sizes */
  for (i = 0; i < c->No; i++) {
    unsigned q = (unsigned)RO_ABS(c->Npolout[i]);
    c->Nss[i] = (unsigned)(((unsigned long long)c->Nb * c->Ntpb) / c->Nts[i]);
    c->Nds[i] = c->Nss[i] / c->Nas[i];
    if (c->Nds[i] == 0) c->Nds[i] = 1;
    c->Nis[i] = c->Nas[i] / c->Nss[i];
    if (c->Nis[i] == 0) c->Nis[i] = 1;
    c->pwr_len[i] = (size_t)q * c->Nds[i] * c->Nts[i] * c->Nc;
    c->ics_len[i] = c->pwr_len[i] / c->Nant;
  }
  /* :1040-1055 ICS weights: one broadcast weight or one per antenna */
  if (c->incoherently_sum) {
    if (!((c->Naws == 1 && (unsigned)c->Naws < c->Nant) ||
          (c->Naws > 0 && (unsigned)c->Naws == c->Nant)))
      return RO_E_NAWS;
  }
  return RO_OK;
}

unsigned long ro_dump_events(const ro_cfg *c, int product, unsigned long nbuf)
{
  /* rawspec_gpu.cu:1543,1572: a dump fires on buffer n iff n % Nis == 0 */
  return nbuf / c->Nis[product];
}

/* ------------------------------------------------------------------------ */
/* Sample conversion.  The reference reads samples through a texture in     */
/* cudaReadModeNormalizedFloat (rawspec_gpu.cu:845-863): a signed n-bit      */
/* integer v maps to max(v / (2^(n-1) - 1), -1).                             */
/* ------------------------------------------------------------------------ */
double ro_snorm8(int8_t v)
{
  double x = (double)v / 127.0;
  return x < -1.0 ? -1.0 : x;
}

double ro_snorm16(int16_t v)
{
  double x = (double)v / 32767.0;
  return x < -1.0 ? -1.0 : x;
}

/* rawspec_gpu.cu:342-350: high nibble -> real, low nibble -> imaginary,
 * each sign-extended. */
void ro_expand4(uint8_t b, int8_t *re, int8_t *im)
{
  int hi = (b >> 4) & 0xf, lo = b & 0xf;
  *re = (int8_t)(hi >= 8 ? hi - 16 : hi);
  *im = (int8_t)(lo >= 8 ? lo - 16 : lo);
}

/* ------------------------------------------------------------------------ */
/* FP64 FFT.  The reference delegates to cuFFT C2C (rawspec_gpu.cu:1182-    */
/* 1195, 1560): unnormalised, exp(-2 pi i k m/n) for CUFFT_FORWARD and       */
/* exp(+...) for CUFFT_INVERSE.  Restated from the definition: iterative     */
/* radix-2 for powers of two, recursive mixed radix otherwise, direct DFT   */
/* for prime lengths.                                                        */
/* ------------------------------------------------------------------------ */
typedef struct {
  size_t n;
  ro_real *tw;  /* forward twiddles exp(-2 pi i m / n), m < n (interleaved) */
} ro_fftplan;

static void plan_init(ro_fftplan *p, size_t n)
{
  size_t m;
  p->n = n;
  p->tw = (ro_real *)malloc(2 * n * sizeof(ro_real));
  for (m = 0; m < n; m++) {
    /* angle reduced exactly: 2 pi m / n */
    double a = -2.0 * M_PI * (double)m / (double)n;
    p->tw[2 * m] = (ro_real)cos(a);
    p->tw[2 * m + 1] = (ro_real)sin(a);
  }
}

static void plan_free(ro_fftplan *p)
{
  free(p->tw);
  p->tw = NULL;
}

static int is_pow2(size_t n) { return n && !(n & (n - 1)); }

/* twiddle exp(sign * 2 pi i e / n) from the forward table */
static inline void tw_get(const ro_fftplan *p, size_t e, int sign,
                          ro_real *wr, ro_real *wi)
{
  e %= p->n;
  *wr = p->tw[2 * e];
  *wi = sign < 0 ? p->tw[2 * e + 1] : -p->tw[2 * e + 1];
}

static void fft_pow2(const ro_fftplan *p, ro_real *d, int sign)
{
  size_t n = p->n, i, j, len;
  /* bit reversal */
  for (i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      ro_real tr = d[2 * i], ti = d[2 * i + 1];
      d[2 * i] = d[2 * j]; d[2 * i + 1] = d[2 * j + 1];
      d[2 * j] = tr; d[2 * j + 1] = ti;
    }
  }
  for (len = 2; len <= n; len <<= 1) {
    size_t half = len >> 1, step = n / len, k, b;
    for (b = 0; b < n; b += len) {
      for (k = 0; k < half; k++) {
        ro_real wr, wi;
        tw_get(p, k * step, sign, &wr, &wi);
        ro_real *u = d + 2 * (b + k), *v = d + 2 * (b + k + half);
        ro_real vr = v[0] * wr - v[1] * wi;
        ro_real vi = v[0] * wi + v[1] * wr;
        v[0] = u[0] - vr; v[1] = u[1] - vi;
        u[0] += vr;       u[1] += vi;
      }
    }
  }
}

static size_t smallest_factor(size_t n)
{
  size_t f;
  if (n % 2 == 0) return 2;
  for (f = 3; f * f <= n; f += 2)
    if (n % f == 0) return f;
  return n;
}

/* Recursive mixed-radix DIT on a strided input: out[k] = sum_m in[m*is] w^(mk)
 * where the base twiddle table belongs to length N_root and this sub-FFT has
 * length n = N_root / tstride. */
static void fft_rec(const ro_fftplan *p, const ro_real *in, size_t is, ro_real *out,
                    size_t n, size_t tstride, int sign, ro_real *scratch)
{
  size_t r, k, f, m;
  if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
  f = smallest_factor(n);
  if (f == n) {
    /* prime length: direct DFT */
    for (k = 0; k < n; k++) {
      ro_real sr = 0, si = 0;
      for (r = 0; r < n; r++) {
        ro_real wr, wi;
        tw_get(p, ((r * k) % n) * tstride, sign, &wr, &wi);
        const ro_real *x = in + 2 * r * is;
        sr += x[0] * wr - x[1] * wi;
        si += x[0] * wi + x[1] * wr;
      }
      out[2 * k] = sr; out[2 * k + 1] = si;
    }
    return;
  }
  m = n / f;
  /* sub-transforms of the f decimated sequences, stored consecutively */
  for (r = 0; r < f; r++)
    fft_rec(p, in + 2 * r * is, is * f, scratch + 2 * r * m, m, tstride * f,
            sign, scratch + 2 * n);
  for (k = 0; k < n; k++) {
    ro_real sr = 0, si = 0;
    for (r = 0; r < f; r++) {
      ro_real wr, wi;
      const ro_real *y = scratch + 2 * (r * m + (k % m));
      tw_get(p, ((r * k) % n) * tstride, sign, &wr, &wi);
      sr += y[0] * wr - y[1] * wi;
      si += y[0] * wi + y[1] * wr;
    }
    out[2 * k] = sr; out[2 * k + 1] = si;
  }
}

static size_t rec_scratch_len(size_t n)
{
  /* each level needs 2n doubles for its sub-results; depth <= log2 n */
  size_t tot = 0;
  while (n > 1) {
    size_t f = smallest_factor(n);
    tot += 2 * n;
    if (f == n) break;
    n /= f;
  }
  return tot + 2;
}

static void plan_exec(const ro_fftplan *p, ro_real *d, int sign, ro_real *work)
{
  if (is_pow2(p->n)) {
    fft_pow2(p, d, sign);
  } else {
    size_t n = p->n;
    ro_real *out = work, *scr = work + 2 * n;
    fft_rec(p, d, 1, out, n, 1, sign, scr);
    memcpy(d, out, 2 * n * sizeof(ro_real));
  }
}

static size_t plan_work_len(size_t n)
{
  return is_pow2(n) ? 0 : 2 * n + rec_scratch_len(n);
}

void ro_fft(double *data, size_t n, int sign)
{
  ro_fftplan p;
  ro_real *x, *work;
  size_t m;
  if (n <= 1) return;
  plan_init(&p, n);
  x = (ro_real *)malloc(2 * n * sizeof(ro_real));
  work = (ro_real *)malloc((plan_work_len(n) + 2) * sizeof(ro_real));
  for (m = 0; m < 2 * n; m++) x[m] = (ro_real)data[m];
  plan_exec(&p, x, sign, work);
  for (m = 0; m < 2 * n; m++) data[m] = (double)x[m];
  free(work);
  free(x);
  plan_free(&p);
}

void ro_dft_naive(const double *in, double *out, size_t n, int sign)
{
  size_t k, m;
  for (k = 0; k < n; k++) {
    long double sr = 0, si = 0;
    for (m = 0; m < n; m++) {
      /* exact reduction of the angle index before the libm call */
      size_t e = (k * m) % n;
      long double a = (long double)sign * 2.0L * (long double)M_PI *
                      (long double)e / (long double)n;
      long double c = cosl(a), s = sinl(a);
      sr += in[2 * m] * c - in[2 * m + 1] * s;
      si += in[2 * m] * s + in[2 * m + 1] * c;
    }
    out[2 * k] = (double)sr;
    out[2 * k + 1] = (double)si;
  }
}

/* ------------------------------------------------------------------------ */
/* Streaming processing: rawspec_start_processing (rawspec_gpu.cu:1517-1747) */
/* ------------------------------------------------------------------------ */
struct ro_state {
  ro_cfg c;
  ro_fftplan plan[RO_MAX_OUTPUTS];
  ro_real *acc[RO_MAX_OUTPUTS];  /* integration in progress, host layout */
  double *w;                     /* Nant antenna weights */
  unsigned long inbuf_count;
  int *chans;
  int nchans;
  int nthreads;
};

ro_state *ro_create(const ro_cfg *c, const float *weights)
{
  unsigned i;
  ro_state *s = (ro_state *)calloc(1, sizeof(ro_state));
  s->c = *c;
  for (i = 0; i < c->No; i++) {
    plan_init(&s->plan[i], c->Nts[i]);
    s->acc[i] = (ro_real *)calloc(c->pwr_len[i], sizeof(ro_real));
  }
  s->w = (double *)calloc(c->Nant, sizeof(double));
  for (i = 0; i < c->Nant; i++) {
    if (!weights) s->w[i] = 1.0;
    else if (c->Naws == 1) s->w[i] = weights[0];     /* :1040-1044 */
    else s->w[i] = weights[i];                        /* :1046-1049 */
  }
  s->nchans = (int)c->Nc;
  s->chans = (int *)malloc(c->Nc * sizeof(int));
  for (i = 0; i < c->Nc; i++) s->chans[i] = (int)i;
  return s;
}

void ro_destroy(ro_state *s)
{
  unsigned i;
  if (!s) return;
  for (i = 0; i < s->c.No; i++) {
    plan_free(&s->plan[i]);
    free(s->acc[i]);
  }
  free(s->w);
  free(s->chans);
  free(s);
}

void ro_reset(ro_state *s)
{
  unsigned i;
  /* rawspec_gpu.cu:1762-1792: partial integrations dropped, count to 0 */
  for (i = 0; i < s->c.No; i++)
    memset(s->acc[i], 0, s->c.pwr_len[i] * sizeof(ro_real));
  s->inbuf_count = 0;
}

void ro_set_channels(ro_state *s, const int *chans, int nchans)
{
  int i;
  if (!chans || nchans <= 0) {
    s->nchans = (int)s->c.Nc;
    for (i = 0; i < s->nchans; i++) s->chans[i] = i;
  } else {
    s->nchans = nchans;
    for (i = 0; i < nchans; i++) s->chans[i] = chans[i];
  }
}

void ro_set_threads(ro_state *s, int nthreads) { s->nthreads = nthreads; }

/* Unpacks Nt samples of pol p starting at time index t0 of channel ch of a
 * device buffer [Nc][Nb*Ntpb][Np][re,im] (rawspec_gpu.cu:133-148 addressing,
 * :1454-1460 layout). */
static void unpack(const ro_cfg *c, const void *buf, unsigned ch, size_t t0,
                   unsigned p, size_t nt, ro_real *x)
{
  size_t m, chan_samples = (size_t)c->Nb * c->Ntpb;
  size_t base = ((size_t)ch * chan_samples + t0) * c->Np + p; /* complex idx */
  if (c->Nbps == 16) {
    const int16_t *b = (const int16_t *)buf;
    for (m = 0; m < nt; m++) {
      size_t e = 2 * (base + m * c->Np);
      x[2 * m] = (ro_real)ro_snorm16(b[e]);
      x[2 * m + 1] = (ro_real)ro_snorm16(b[e + 1]);
    }
  } else {
    const int8_t *b = (const int8_t *)buf;
    for (m = 0; m < nt; m++) {
      size_t e = 2 * (base + m * c->Np);
      x[2 * m] = (ro_real)ro_snorm8(b[e]);
      x[2 * m + 1] = (ro_real)ro_snorm8(b[e + 1]);
    }
  }
}

/* Detection for one spectrum, accumulated into the host-layout accumulator
 * at shifted fine index.  rawspec_gpu.cu:153-267 (store callbacks) and
 * :1659-1691 (the two-copy FFT shift: D[k] lands at (k + Nt/2) mod Nt).
 * Pol-mode mapping follows the code: +4 -> XX,YY,Re XY*,Im XY*
 * (store_callback_pol0/pol1[_conj]); -4 -> I,Q,U,V (…_iquv[_conj]). */
static void detect_acc(const ro_cfg *c, int i, const ro_real *X0,
                       const ro_real *X1, ro_real *acc_cd, size_t qstride)
{
  size_t nt = c->Nts[i], k;
  size_t half = nt / 2;
  int mode = c->Npolout[i];
  ro_real g = c->input_conjugated ? -1 : 1;
  for (k = 0; k < nt; k++) {
    size_t j = (k + half) % nt;
    ro_real a = X0[2 * k], b = X0[2 * k + 1];
    ro_real p0 = a * a + b * b;
    if (mode == 1) {
      ro_real p = p0;
      if (X1) p += X1[2 * k] * X1[2 * k] + X1[2 * k + 1] * X1[2 * k + 1];
      acc_cd[j] += p;
    } else {
      ro_real cr = X1[2 * k], ci = X1[2 * k + 1];
      ro_real p1 = cr * cr + ci * ci;
      ro_real xre = a * cr + b * ci;         /* Re(X0 conj X1) */
      ro_real xim = g * (b * cr - a * ci);   /* Im(X0 conj X1), conj flips */
      if (mode == 4) {
        acc_cd[j] += p0;
        acc_cd[qstride + j] += p1;
      } else {
        acc_cd[j] += p0 + p1;
        acc_cd[qstride + j] += p0 - p1;
      }
      acc_cd[2 * qstride + j] += xre;
      acc_cd[3 * qstride + j] += xim;
    }
  }
}

int ro_process(ro_state *s, const void *buf, int fft_dir,
               double *const pwr[RO_MAX_OUTPUTS],
               double *const ics[RO_MAX_OUTPUTS])
{
  const ro_cfg *c = &s->c;
  int i, mask = 0;
  int sign = fft_dir <= 0 ? +1 : -1;   /* :1563 */
  s->inbuf_count++;                     /* :1543 */
  for (i = 0; i < (int)c->No; i++) {
    size_t nt = c->Nts[i];
    unsigned nss = c->Nss[i], nis = c->Nis[i], nds = c->Nds[i];
    unsigned q = (unsigned)RO_ABS(c->Npolout[i]);
    /* spectra summed per dump group inside this buffer: min(Na, Nss)
     * (:1588-1601); group d starts at spectrum d*Na */
    unsigned na_buf = c->Nas[i] < nss ? c->Nas[i] : nss;
    unsigned ngroups = nss / na_buf;   /* == Nds when Nis == 1, else 1 */
    long ntask = (long)s->nchans * ngroups;
    size_t wl = plan_work_len(nt);
    long t;
#ifdef _OPENMP
    int nth = s->nthreads > 0 ? s->nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nth)
#endif
    {
      ro_real *x0 = (ro_real *)malloc((2 * nt + wl + 2) * sizeof(ro_real));
      ro_real *x1 = (ro_real *)malloc((2 * nt + wl + 2) * sizeof(ro_real));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
      for (t = 0; t < ntask; t++) {
        unsigned ch = (unsigned)s->chans[t / ngroups];
        unsigned g = (unsigned)(t % ngroups);
        unsigned d = (nis == 1) ? g : 0;
        size_t qstride = (size_t)c->Nc * nt;
        ro_real *acc_cd = s->acc[i] + ((size_t)d * q * c->Nc + ch) * nt;
        unsigned sp;
        for (sp = g * na_buf; sp < (g + 1) * na_buf; sp++) {
          unpack(c, buf, ch, (size_t)sp * nt, 0, nt, x0);
          plan_exec(&s->plan[i], x0, sign, x0 + 2 * nt);
          if (c->Np == 2) {
            unpack(c, buf, ch, (size_t)sp * nt, 1, nt, x1);
            plan_exec(&s->plan[i], x1, sign, x1 + 2 * nt);
          }
          detect_acc(c, i, x0, c->Np == 2 ? x1 : NULL, acc_cd, qstride);
        }
      }
      free(x0);
      free(x1);
    }
    (void)nds;
    if (s->inbuf_count % nis == 0) {    /* :1572 dump */
      if (pwr && pwr[i]) {
        size_t e;
        for (e = 0; e < c->pwr_len[i]; e++) pwr[i][e] = (double)s->acc[i][e];
      }
      if (c->incoherently_sum && ics && ics[i]) {
        /* ICS (:321-340 at Nbc = Nc/Nant): weighted sum over antennas of the
         * integrated, shifted spectra; layout [Nds][|P|][Nc/Nant][Nt]. */
        size_t ncpa = c->Nc / c->Nant, dd, qq, cc, j;
        unsigned a;
        for (dd = 0; dd < c->Nds[i]; dd++)
          for (qq = 0; qq < q; qq++)
            for (cc = 0; cc < ncpa; cc++)
              for (j = 0; j < nt; j++) {
                ro_real sum = 0;
                for (a = 0; a < c->Nant; a++)
                  sum += (ro_real)s->w[a] *
                         s->acc[i][((dd * q + qq) * c->Nc + a * ncpa + cc) * nt + j];
                ics[i][((dd * q + qq) * ncpa + cc) * nt + j] = sum;
              }
      }
      memset(s->acc[i], 0, c->pwr_len[i] * sizeof(ro_real));  /* :1711-1719 */
      mask |= 1 << i;
    }
  }
  return mask;
}

/* ------------------------------------------------------------------------ */
/* Seeded synthetic GUPPI blocks (test input generator, no reference code). */
/* ------------------------------------------------------------------------ */
uint64_t ro_splitmix64(uint64_t x)
{
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void ro_gen_block(void *dst, unsigned Nc, unsigned Ntpb, unsigned Np,
                  unsigned Nbps, uint64_t seed, double sigma, int clip,
                  unsigned long blk, int tone_chan, double tone_amp,
                  double tone_bin, double tone_nfft)
{
  long ch;
  uint64_t key = seed * 0x9E3779B97F4A7C15ULL;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (ch = 0; ch < (long)Nc; ch++) {
    size_t t;
    unsigned p;
    for (t = 0; t < Ntpb; t++) {
      uint64_t gt = (uint64_t)blk * Ntpb + t;
      for (p = 0; p < Np; p++) {
        uint64_t ctr = ((uint64_t)ch << 40) ^ (gt << 1) ^ p;
        uint64_t h1 = ro_splitmix64(key + 2 * ctr);
        uint64_t h2 = ro_splitmix64(key + 2 * ctr + 1);
        double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);
        double r = sqrt(-2.0 * log(u1));
        double re = sigma * r * cos(2.0 * M_PI * u2);
        double im = sigma * r * sin(2.0 * M_PI * u2);
        long vr, vi;
        size_t e = 2 * (((size_t)ch * Ntpb + t) * Np + p);
        if (p == 0 && (long)tone_chan == ch && tone_amp != 0.0) {
          /* phase = 2 pi * frac(bin * gt / nfft), reduced exactly */
          double ph = fmod(tone_bin * (double)gt, tone_nfft) / tone_nfft;
          re += tone_amp * cos(2.0 * M_PI * ph);
          im += tone_amp * sin(2.0 * M_PI * ph);
        }
        vr = lround(re);
        vi = lround(im);
        if (vr > clip) vr = clip;
        if (vr < -clip) vr = -clip;
        if (vi > clip) vi = clip;
        if (vi < -clip) vi = -clip;
        if (Nbps == 16) {
          ((int16_t *)dst)[e] = (int16_t)vr;
          ((int16_t *)dst)[e + 1] = (int16_t)vi;
        } else {
          ((int8_t *)dst)[e] = (int8_t)vr;
          ((int8_t *)dst)[e + 1] = (int8_t)vi;
        }
      }
    }
  }
}
