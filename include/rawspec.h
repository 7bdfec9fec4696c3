/*
 * rawspec.h -- public C API of librawspec (MI355X / gfx950 implementation).
 *
 * This header is a fresh restatement of the librawspec interface so that
 * existing clients (the rawspec CLI, rawspectest, autoconf-probed programs)
 * compile and link against this library unchanged.  The layout of
 * `struct rawspec_context_s` is binary-identical to the reference
 * (rawspec.h:32-148): same field order, same C types, same array bounds.
 *
 * Reference interface replaced, entry point by entry point:
 *   get_librawspec_version                     rawspec.h:163, rawspec_gpu.cu:431
 *   get_cufft_version                          rawspec.h:165, rawspec_gpu.cu:437
 *   rawspec_initialize                         rawspec.h:174, rawspec_gpu.cu:448-1294
 *   rawspec_cleanup                            rawspec.h:181, rawspec_gpu.cu:1300-1391
 *   rawspec_copy_blocks_to_gpu                 rawspec.h:186, rawspec_gpu.cu:1441-1469
 *   rawspec_copy_blocks_to_gpu_expanding_complex4
 *                                              rawspec.h:193, rawspec_gpu.cu:1395-1437
 *   rawspec_zero_blocks_to_gpu                 rawspec.h:201, rawspec_gpu.cu:1476-1500
 *   rawspec_start_processing                   rawspec.h:220, rawspec_gpu.cu:1517-1747
 *   rawspec_copy_blocks_to_gpu_and_start_processing
 *                                              rawspec.h:224, rawspec_gpu.cu:1749-1758
 *   rawspec_reset_integration                  rawspec.h:228, rawspec_gpu.cu:1762-1792
 *   rawspec_check_for_completion               rawspec.h:233, rawspec_gpu.cu:1795-1807
 *   rawspec_wait_for_completion                rawspec.h:237, rawspec_gpu.cu:1811-1836
 *
 * Semantics that differ from the reference on purpose are listed next to the
 * declaration they concern.
 */
#ifndef RAWSPEC_H_
#define RAWSPEC_H_

#include <stddef.h>
#include <unistd.h>   /* off_t, size_t as in the reference header */

/* FFT direction selectors for rawspec_start_processing() (rawspec.h:7-8).
 * fft_dir > 0 selects the forward transform exp(-2*pi*i*k*m/N);
 * fft_dir <= 0 selects the inverse (unnormalised) transform exp(+2*pi*i*k*m/N). */
#define RAWSPEC_FORWARD_FFT (+1)
#define RAWSPEC_INVERSE_FFT (-1)

/* Maximum number of output products per context (rawspec.h:10). */
#define MAX_OUTPUTS (4)

/* Size in bytes of one GUPPI RAW block as seen by the library
 * (rawspec.h:12-19): Nc * Ntpb * Np * 2 * Nbps / 8. */
#define RAWSPEC_BLOCSIZE(pctx)      \
  (                                 \
    ((pctx)->Nc   *                 \
     (pctx)->Ntpb *                 \
     (pctx)->Np   *                 \
     2            *                 \
     (pctx)->Nbps) / 8              \
  )

/* callback_type values passed to the dump callback (rawspec.h:21-22). */
#define RAWSPEC_CALLBACK_PRE_DUMP  (0)
#define RAWSPEC_CALLBACK_POST_DUMP (1)

typedef struct rawspec_context_s rawspec_context;

/* Dump callback (rawspec.h:27-29).  Called on a runtime-owned thread, in
 * stream order: PRE(i) before integrated spectra of product i are written to
 * h_pwrbuf[i] (and h_icsbuf[i]); POST(i) once they are there.  The data stay
 * valid until the next PRE(i).  The callback must not call GPU APIs. */
typedef void (*rawspec_dump_callback_t)(rawspec_context *ctx,
                                        int output_product,
                                        int callback_type);

struct rawspec_context_s {
  /* ---- set by the client ------------------------------------------------ */
  unsigned int No;        /* number of output products, 1..MAX_OUTPUTS        */
  unsigned int Np;        /* input polarisations, 1 or 2                      */
  unsigned int Nant;      /* antennas; Nc is a multiple of Nant (0 -> 1)      */
  unsigned int Nc;        /* coarse channels per block                        */
  unsigned int Ntpb;      /* time samples per block per channel               */
  unsigned int exit_soon; /* set by the client's callback to request a stop   */

  /* Coarse channels per batch.  0 disables batching, 1 (or less) picks a
   * batch automatically, otherwise it must divide Nc.  This implementation
   * keeps every channel resident, so the value is validated and reported
   * exactly as the reference does but does not change any result. */
  unsigned int Nbc;

  /* Bits per sample component: 4, 8 or 16 (0 is taken as 8).  4-bit data
   * are complex4 bytes (high nibble = real, low nibble = imaginary) that the
   * library expands to 8 bits on the device; the field then reads 8. */
  unsigned int Nbps;

  /* Output polarisation mode per product.  As implemented by the reference
   * code (rawspec_gpu.cu:1218-1233, rawspec.c:120-121):
   *    1  total power  |X0|^2 + |X1|^2
   *   +4  full pol     XX, YY, Re(X0 X1*), Im(X0 X1*)
   *   -4  full Stokes  I, Q, U, V   (U, V without a factor of 2)
   * Any other value, or Np != 2, is forced to 1 by rawspec_initialize(). */
  int Npolout[MAX_OUTPUTS];

  /* FFT length per product; every value must divide the largest one. */
  unsigned int Nts[MAX_OUTPUTS];

  /* Spectra integrated per dump per product.  Nt*Na and Nb*Ntpb must divide
   * one another (whichever is larger is the multiple). */
  unsigned int Nas[MAX_OUTPUTS];

  /* Client dump callback, may be NULL. */
  rawspec_dump_callback_t dump_callback;

  /* Opaque client pointer, never touched by the library. */
  void *user_data;

  /* Blocks per device input buffer; 0 lets the library derive it. */
  unsigned int Nb;

  /* Host block buffers; 0 lets the library derive it (must be set when the
   * caller provides h_blkbufs). */
  unsigned int Nb_host;

  /* Host block buffers.  NULL: the library allocates Nb_host pinned blocks
   * and frees them in rawspec_cleanup().  Otherwise the caller owns them.
   * Deviation: caller-owned blocks are NOT page-locked by the library (the
   * reference calls cudaHostRegister, rawspec_gpu.cu:719-728); copies from
   * them are plain pageable copies.  Give the library pinned memory if
   * copy speed matters. */
  char **h_blkbufs;

  /* Device ordinal to run on. */
  int gpu_index;

  /* 1 if the band is frequency-flipped: negates the imaginary cross term
   * (Im XY* for +4, V for -4).  Read by rawspec_initialize(). */
  int input_conjugated;

  /* Incoherent sum across antennas (1 = on), with Naws weights in Aws
   * (Naws == 1 broadcasts one weight, Naws == Nant gives one per antenna).
   * Deviation: Aws stays owned by the caller and is never freed by the
   * library (the reference frees it with the wrong allocator,
   * rawspec_gpu.cu:1379-1382). */
  int incoherently_sum;
  int Naws;
  float *Aws;

  /* ---- managed by the library ------------------------------------------ */

  /* Pinned host buffers that receive each dump, laid out
   * [Nds][|Npolout|][Nc][Nt] with the fine channels of every coarse channel
   * in FFT-shifted order (DC at index Nt/2). */
  float *h_pwrbuf[MAX_OUTPUTS];
  size_t h_pwrbuf_size[MAX_OUTPUTS];

  /* Incoherent-sum dumps, [Nds][|Npolout|][Nc/Nant][Nt], size
   * h_pwrbuf_size[i]/Nant; only allocated when incoherently_sum == 1. */
  float *h_icsbuf[MAX_OUTPUTS];

  /* Integrated spectra per dump per product. */
  unsigned int Nds[MAX_OUTPUTS];

  unsigned int Ntmax;     /* largest Nts[i]                                   */
  void *gpu_ctx;          /* opaque per-context device state                  */
};

/* Output sink selector used by the CLI (rawspec.h:151-155). */
typedef enum {
  RAWSPEC_FILE,
  RAWSPEC_NET
} rawspec_output_mode_t;

#ifdef __cplusplus
extern "C" {
#endif

/* Version strings (rawspec.h:161-165).  get_cufft_version() is kept for
 * binary compatibility; this library has no cuFFT and returns a string that
 * names its own FFT engine. */
const char *get_librawspec_version();
const char *get_cufft_version();

/* Validates the context, derives Ntmax, Nb, Nb_host, Nbc, Nds, allocates host
 * and device buffers, creates the context's stream.  Returns 0 on success,
 * 1 on error (after releasing whatever it had allocated). */
int rawspec_initialize(rawspec_context *ctx);

/* Releases everything rawspec_initialize() allocated. */
void rawspec_cleanup(rawspec_context *ctx);

/* Copies num_blocks host blocks, starting at h_blkbufs[src_idx % Nb_host],
 * into device blocks (dst_idx + b) % Nb.  Returns 0 on success. */
int rawspec_copy_blocks_to_gpu(rawspec_context *ctx,
                               off_t src_idx, off_t dst_idx, size_t num_blocks);

/* As rawspec_copy_blocks_to_gpu() for complex4 host blocks (half size),
 * expanding every nibble to a signed byte on the device.  Fails when
 * num_blocks > Nb. */
int rawspec_copy_blocks_to_gpu_expanding_complex4(rawspec_context *ctx,
                               off_t src_idx, off_t dst_idx, size_t num_blocks);

/* Zero-fills device blocks (dst_idx + b) % Nb for b < num_blocks. */
int rawspec_zero_blocks_to_gpu(rawspec_context *ctx,
                               off_t dst_idx, size_t num_blocks);

/* Processes the device input buffer asynchronously: FFT, detection,
 * integration; a product whose integration completes is copied to
 * h_pwrbuf[i] between a PRE and a POST callback.  Returns 0 on success. */
int rawspec_start_processing(rawspec_context *ctx, int fft_dir);

/* Copies Nb... blocks from h_blkbufs[0..num_blocks) to device blocks
 * 0..num_blocks (expanding complex4 when asked) then starts processing. */
int rawspec_copy_blocks_to_gpu_and_start_processing(rawspec_context *ctx,
                               size_t num_blocks, char expand4bps_to8bps,
                               int fft_dir);

/* Waits for outstanding work, discards partial integrations and restarts
 * the integration count.  Returns 0 on success. */
int rawspec_reset_integration(rawspec_context *ctx);

/* Non-blocking completion test.  Returns ctx->No when the context is idle and
 * 0 otherwise, as documented by the reference header (rawspec.h:227-230).
 * (The reference code returns 1 when idle, rawspec_gpu.cu:1795-1807; both
 * are non-zero exactly when idle.) */
unsigned int rawspec_check_for_completion(rawspec_context *ctx);

/* Enqueues one extra PRE callback per product, waits for the context's
 * stream and returns ctx->exit_soon (1 if the wait itself failed). */
int rawspec_wait_for_completion(rawspec_context *ctx);

#ifdef __cplusplus
}
#endif

#endif /* RAWSPEC_H_ */
