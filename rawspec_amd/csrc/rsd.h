/*
 * rsd.h -- thin C ABI between librawspec's C host layer (rawspec_api.c) and
 * its HIP device layer (rsd_runtime.hip + kernels).  Plain C types only.
 *
 * The host layer owns the API semantics of rawspec.h (validation, derived
 * counts, dump schedule, callbacks, buffer ownership); the device layer owns
 * device memory, the context's stream and the kernels.  Every call selects
 * the context's device first, so several contexts (on one or several GPUs)
 * can live in one process.
 */
#ifndef RSD_H_
#define RSD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RSD_MAX_PRODUCTS 4

typedef struct rsd_ctx rsd_ctx;

typedef struct {
  unsigned Nt;        /* FFT length                                   */
  unsigned Na;        /* spectra per integration                      */
  unsigned Nss;       /* spectra per channel per buffer               */
  unsigned Nds;       /* spectra per dump                             */
  unsigned Nis;       /* buffers per dump                             */
  int npolout;        /* 1, 4 or -4                                   */
} rsd_product;

typedef struct {
  int gpu_index;
  unsigned Nc, Np, Ntpb, Nb;
  unsigned Nbps;      /* 8 or 16 (after complex4 expansion)           */
  int expand4;        /* host blocks hold complex4 bytes              */
  unsigned Nant;
  int conj;           /* input_conjugated                             */
  int ics;            /* incoherent sum requested                     */
  const float *weights; /* Nant weights (broadcast already resolved)  */
  unsigned No;
  rsd_product prod[RSD_MAX_PRODUCTS];
} rsd_config;

/* Creates the device context: stream, input buffer, integration buffers,
 * scratch, twiddle tables.  Returns 0 on success. */
int  rsd_create(const rsd_config *cfg, rsd_ctx **out);
/* Device bytes rsd_create would allocate for cfg (0 if cfg is unsupported);
 * touches no device. */
size_t rsd_plan_bytes(const rsd_config *cfg);
void rsd_destroy(rsd_ctx *d);

/* Page-locked host memory (write_combined: host writes only). */
int  rsd_host_alloc(int gpu_index, void **p, size_t bytes, int write_combined);
void rsd_host_free(void *p);

/* Synchronous host -> device copies of one host block into device block
 * slot dblk (reference rawspec_gpu.cu:1441-1469 and :1395-1437). */
int  rsd_copy_block(rsd_ctx *d, const void *src, unsigned dblk);
int  rsd_copy_block4(rsd_ctx *d, const void *src, unsigned dblk);
int  rsd_zero_block(rsd_ctx *d, unsigned dblk);

/* Enqueues the kernels of product i over device input d_in (NULL: the
 * context's own input buffer).  first != 0 starts a new integration;
 * last != 0 completes one (the dump follows). */
int  rsd_run_product(rsd_ctx *d, int i, const void *d_in, int inverse, int first, int last);
/* As rsd_run_product over n consecutive buffers d_in[0 .. n-1] of one
 * integration (first: d_in[0] starts it; last: d_in[n-1] completes it). */
int  rsd_run_product_batch(rsd_ctx *d, int i, const void *const *d_in, unsigned n, int inverse,
                           int first, int last);
/* A fused pair under the current engine: 8-point product *a, whose dumps
 * integrate 128 spectra (one per 1024 samples), computed by the kernel of
 * 1024-point product *b from the same read of the samples (engine level 1,
 * int8, two pols, same detection).  Returns 1 and sets a, b when the
 * context has such a pair and the engine uses it, else 0.  The pair's
 * products then run together through rsd_run_fused, once per buffer, in
 * place of rsd_run_product for either. */
int  rsd_fused_pair(rsd_ctx *d, int *a, int *b);
/* Enqueues both products of the fused pair over one buffer d_in (NULL: the
 * context's own input); first_b != 0 starts product b's integration
 * (product a's integration is always complete within the buffer). */
int  rsd_run_fused(rsd_ctx *d, const void *d_in, int inverse, int first_b);
/* Products run on two lanes (streams) forked from the context stream:
 * four-step products on one, small products on the other.  rsd_lanes_begin
 * makes both lanes wait for everything enqueued on the context stream;
 * rsd_lane_end makes the context stream wait for product i's lane (i < 0:
 * both lanes); rsd_lane_resume makes product i's lane wait for the context
 * stream (after a dump that must precede the lane's next writes).  With
 * lanes off, or while kernel timing is on, all are no-ops and everything
 * runs on the context stream. */
int  rsd_lanes_begin(rsd_ctx *d);
int  rsd_lane_end(rsd_ctx *d, int i);
int  rsd_lane_resume(rsd_ctx *d, int i);
/* Enqueues the incoherent sum of product i's integration buffer. */
int  rsd_run_ics(rsd_ctx *d, int i);
/* Enqueues the device -> host copy of product i's dump (and ICS dump when
 * h_ics is not NULL). */
int  rsd_dump_to_host(rsd_ctx *d, int i, float *h_pwr, size_t pwr_bytes,
                      float *h_ics, size_t ics_bytes);
/* Enqueues a host function in stream order. */
int  rsd_host_fn(rsd_ctx *d, void (*fn)(void *), void *arg);
/* 1: idle, 0: busy, -1: error. */
int  rsd_query(rsd_ctx *d);
/* Blocks until the stream drains.  0 on success. */
int  rsd_sync(rsd_ctx *d);

void  *rsd_input_ptr(rsd_ctx *d);
size_t rsd_input_bytes(rsd_ctx *d);
float *rsd_pwr_ptr(rsd_ctx *d, int i);
float *rsd_ics_ptr(rsd_ctx *d, int i);
void  *rsd_stream(rsd_ctx *d);
/* Queues the context's work on `stream` (a hipStream_t of the context's
 * device) from now on, after everything already queued; NULL returns to
 * the context's own stream.  The caller keeps the stream alive while the
 * context uses it. */
int    rsd_set_stream(rsd_ctx *d, void *stream);

/* Human-readable name of the FFT engine (returned by get_cufft_version). */
const char *rsd_engine_name(void);

/* Kernel timing: when enabled, every kernel launch is bracketed by HIP
 * events on the context's stream.  rsd_profile_report() synchronises, then
 * writes one line per kernel, "name launches total_ms", into buf and
 * clears the records.  Returns the number of bytes written, or -1. */
int rsd_profile(rsd_ctx *d, int enable);
int rsd_profile_report(rsd_ctx *d, char *buf, size_t len);

/* Transform engine: 0 = FP32 vector units (default); 1 = the matrix-core
 * engine at its 1e-3 bar (unsplit FP16 operands for int8 1024-point products
 * integrating at least 32 spectra, a 24-bit four-step intermediate, vector
 * kernels everywhere else, both four-step passes included;
 * see rsd_fused_pair for the 8-point product it computes with a 1024-point one);
 * 2 = matrix cores with split FP16 operands everywhere (FP32-level error).
 * Other levels are refused (1).  A four-step product switches at the start
 * of its next integration. */
int rsd_set_mfma(rsd_ctx *d, int level);
/* Two-lane execution on (default where it pays) or off. */
int rsd_set_lanes(rsd_ctx *d, int enable);
int rsd_get_lanes(rsd_ctx *d);
int rsd_get_mfma(rsd_ctx *d);
/* Buffers of first-pass output a four-step product that integrates over
 * several buffers may hold across one-buffer calls (default 8, capped by the
 * product's integration length and 8 GiB per context; below 2: none).
 * Applies from the next integration start. */
int rsd_set_hold(rsd_ctx *d, unsigned max_buffers);
unsigned rsd_get_hold(rsd_ctx *d);
/* Launch-plan candidates of product i (index 0: the defaults).
 * rsd_tune_apply makes candidate k the product's plan (growing scratch
 * buffers as needed, with the device idle); returns 2 when its plan equals
 * an earlier candidate's.  rsd_tune_describe writes the current plan. */
int rsd_tune_count(rsd_ctx *d, int i);
int rsd_tune_apply(rsd_ctx *d, int i, int k);
int rsd_tune_describe(rsd_ctx *d, int i, char *buf, size_t len);
/* Allocation-free processing: between rsd_noalloc_begin and rsd_noalloc_end
 * the context's kernels take their partial sums and four-step intermediate
 * from the device's static scratch (a fixed size, shared by every call on
 * that GPU), never allocate held first-pass output and record no timing
 * events.  begin fails, with a message, when one channel's intermediate or
 * the partial sums do not fit; end restores the context's own buffers. */
int rsd_noalloc_begin(rsd_ctx *d);
void rsd_noalloc_end(rsd_ctx *d);
/* Shrinks the scratch buffers to what the products' current plans need
 * (after tuning, which grows them for larger candidates). */
int rsd_tune_fit(rsd_ctx *d);

#ifdef __cplusplus
}
#endif

#endif
