/*
 * rawspec_oracle.h -- FP64 CPU restatement of librawspec's GPU spectrometer
 * semantics.  TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py as the checker.  The product
 * library never links or calls it.
 *
 * What it restates (reference file:line):
 *   ro_derive        validation + derived counts   rawspec_gpu.cu:471-641, 732-760
 *   ro_unpack_*      sample -> complex value        rawspec_gpu.cu:133-148 (texture
 *                                                   SNORM read, :845-863), :342-374 (4-bit)
 *   ro_fft           unnormalised C2C DFT           rawspec_gpu.cu:1182-1195, 1560-1563
 *   ro_process       detection, integration, dump   rawspec_gpu.cu:153-267, 297-340,
 *                    schedule, FFT shift, ICS       1517-1747
 *   ro_gen_*         seeded synthetic GUPPI blocks  (no reference counterpart: the
 *                                                   generator behind the test inputs)
 *
 * The Makefile also builds the same source with -DRO_REAL=float
 * (liboracle32.so): the algorithm in single precision, an FP32 CPU port that
 * the CPU baseline times.  It is never the checker; the API stays double.
 */
#ifndef RAWSPEC_ORACLE_H_
#define RAWSPEC_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RO_MAX_OUTPUTS 4

/* Configuration: the client-set fields of rawspec_context that matter to the
 * numbers, plus the derived values ro_derive() fills in. */
typedef struct {
  /* client */
  unsigned No, Np, Nant, Nc, Ntpb, Nbps, Nbc, Nb, Nb_host;
  int      Npolout[RO_MAX_OUTPUTS];
  unsigned Nts[RO_MAX_OUTPUTS];
  unsigned Nas[RO_MAX_OUTPUTS];
  int      input_conjugated;
  int      incoherently_sum;
  int      Naws;
  int      have_blkbufs;            /* 1 if the caller manages host blocks */
  /* derived */
  int      expand4;                 /* input was 4-bit complex             */
  unsigned Ntmax;
  unsigned Nss[RO_MAX_OUTPUTS];     /* spectra per channel per buffer      */
  unsigned Nds[RO_MAX_OUTPUTS];     /* spectra per dump                    */
  unsigned Nis[RO_MAX_OUTPUTS];     /* buffers per dump                    */
  size_t   pwr_len[RO_MAX_OUTPUTS]; /* floats per dump (h_pwrbuf_size/4)   */
  size_t   ics_len[RO_MAX_OUTPUTS]; /* floats per ICS dump                 */
} ro_cfg;

/* Validates and derives, following rawspec_initialize's order of checks.
 * Returns 0 on success, otherwise a positive code naming the failed rule
 * (see RO_E_* below).  The reference returns 1 for all of them. */
int ro_derive(ro_cfg *c);

enum {
  RO_OK = 0,
  RO_E_NO = 1, RO_E_NP, RO_E_NTPB, RO_E_NBPS, RO_E_NT0, RO_E_NTDIV,
  RO_E_NB, RO_E_NTPBDIV, RO_E_NTMAXDIV, RO_E_NBHOST, RO_E_NA0, RO_E_NADIV,
  RO_E_NBC, RO_E_NAWS
};

/* Sample conversion (texture read in normalised-float mode). */
double ro_snorm8(int8_t v);
double ro_snorm16(int16_t v);
/* complex4 byte -> (re, im) signed nibbles (rawspec_gpu.cu:342-350). */
void   ro_expand4(uint8_t b, int8_t *re, int8_t *im);

/* In-place unnormalised DFT of length n: sign = -1 gives exp(-2 pi i k m / n)
 * (forward), +1 gives the inverse kernel.  Any n >= 1 (mixed radix with a
 * direct DFT for prime factors).  data is interleaved re,im. */
void ro_fft(double *data, size_t n, int sign);
/* O(n^2) reference DFT, used to pin ro_fft. */
void ro_dft_naive(const double *in, double *out, size_t n, int sign);

/* Streaming state: one per configuration. */
typedef struct ro_state ro_state;
ro_state *ro_create(const ro_cfg *c, const float *weights /* Nant values or NULL */);
void      ro_destroy(ro_state *s);
void      ro_reset(ro_state *s);
/* Restrict processing to a channel subset (NULL / 0 = all channels).
 * Outputs for other channels are left at zero. */
void      ro_set_channels(ro_state *s, const int *chans, int nchans);
/* Number of OpenMP threads used by ro_process (0 = runtime default). */
void      ro_set_threads(ro_state *s, int nthreads);

/* Processes one device input buffer (layout [Nc][Nb*Ntpb][Np][re,im] with
 * Nbps-bit components, i.e. what rawspec_copy_blocks_to_gpu builds).
 * fft_dir follows rawspec_start_processing: > 0 forward, <= 0 inverse.
 * For every product that completes an integration on this buffer, bit i of
 * the return value is set and the dump (host layout, FFT-shifted) is written
 * to pwr[i] (pwr_len[i] doubles) and, with ICS on, ics[i] (ics_len[i]).
 * Returns -1 on error. */
int ro_process(ro_state *s, const void *buf, int fft_dir,
               double *const pwr[RO_MAX_OUTPUTS],
               double *const ics[RO_MAX_OUTPUTS]);

/* Number of dump events a product emits after nbuf buffers since reset. */
unsigned long ro_dump_events(const ro_cfg *c, int product, unsigned long nbuf);

/* ---- seeded synthetic input ------------------------------------------- */

/* splitmix64 finaliser of x (public-domain mixing function). */
uint64_t ro_splitmix64(uint64_t x);

/* Fills one GUPPI host block [Nc][Ntpb][Np][re,im] (Nbps = 8 or 16) for
 * block index blk: components round(N(0, sigma^2)) clipped to +-clip, plus
 * an optional complex tone of amplitude tone_amp at frequency
 * tone_bin/tone_nfft cycles per sample on channel tone_chan, pol 0, with
 * phase continuous in the global sample index blk*Ntpb + t.  tone_chan < 0
 * disables the tone.  Values depend only on (seed, chan, global t, pol). */
void ro_gen_block(void *dst, unsigned Nc, unsigned Ntpb, unsigned Np,
                  unsigned Nbps, uint64_t seed, double sigma, int clip,
                  unsigned long blk, int tone_chan, double tone_amp,
                  double tone_bin, double tone_nfft);

#ifdef __cplusplus
}
#endif

#endif
