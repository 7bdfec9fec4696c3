/*
 * rawspec_amd.h -- extensions of librawspec beyond the reference API.
 *
 * None of these has a counterpart in the reference (rawspec.h): they serve
 * GPU-resident producers and consumers that should not pay a PCIe round trip
 * per buffer (a capture pipeline writing voltages straight into device
 * memory, a GPU search consuming spectra in place), and the throughput
 * measurement, which times the device-side path with inputs already in HBM.
 * The reference API in rawspec.h keeps working unchanged alongside them.
 */
#ifndef RAWSPEC_AMD_H_
#define RAWSPEC_AMD_H_

#include <stddef.h>

#include "rawspec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Validates and derives exactly as rawspec_initialize() does (updating the
 * same fields of ctx) but allocates nothing; reports the device bytes and the
 * pinned host bytes rawspec_initialize() would allocate.  For admission
 * control by HBM budget.  The device bytes include the held first-pass
 * output (rawspec_amd_set_held_buffers) the context may allocate at first
 * use with the default setting; a caller that turns holding off gets an
 * upper bound.  Returns 0 when the configuration is valid. */
int rawspec_amd_plan(rawspec_context *ctx, size_t *device_bytes, size_t *host_bytes);

/* Device pointer and size of the context's input buffer, laid out
 * [Nc][Nb*Ntpb][Np][re,im] with Nbps-bit components (the layout
 * rawspec_copy_blocks_to_gpu builds).  NULL / 0 before initialisation. */
void  *rawspec_amd_device_input(rawspec_context *ctx);
size_t rawspec_amd_device_input_size(rawspec_context *ctx);

/* As rawspec_start_processing(), reading the input from d_in, a 16-byte
 * aligned device buffer of rawspec_amd_device_input_size() bytes in the same
 * layout on the context's GPU.  d_in must stay unchanged until the work
 * completes. */
int rawspec_amd_start_processing_device(rawspec_context *ctx, const void *d_in,
                                        int fft_dir);

/* As rawspec_amd_start_processing_device(), but the call allocates nothing
 * (no hipMalloc, hipFree or allocating library routine), for callers that
 * cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation.  No counterpart in the reference.  The
 * kernels keep their partial sums and the four-step intermediate in a static
 * device array of 256 MiB that the library holds on each GPU it runs on until
 * the process ends (the context's own buffers, allocated by
 * rawspec_initialize(), still hold input, integrations and dumps).  Four-step
 * products transform as many channels per launch as that scratch holds, and
 * an integration that starts in this call does not hold first-pass output
 * unless an earlier call allocated the store.  A configuration whose partial
 * sums, or one channel's intermediate, exceed the scratch is refused with a
 * message (returns 1, nothing enqueued).  Every call on a GPU shares that
 * GPU's scratch: calls that overlap on one GPU (contexts on different
 * streams) race; calls on one context are ordered by its stream.  Results
 * equal rawspec_amd_start_processing_device()'s with held output off. */
int rawspec_amd_start_processing_device_noalloc(rawspec_context *ctx, const void *d_in,
                                                int fft_dir);

/* As nbuf calls of rawspec_amd_start_processing_device(), one per buffer of
 * d_in[0 .. nbuf-1] in order: the same integrations, the same dumps, each
 * product's PRE/POST callbacks in the same order.  The device work is
 * batched: a product whose integration spans several buffers transforms
 * consecutive buffers together and adds them into its running sum once per
 * batch (up to 8 buffers per launch) instead of once per buffer.  Two
 * differences from the sequence of single calls: callbacks of different
 * products may interleave differently (each product's own order is kept),
 * and a batched sum adds the buffers in a different order, so results agree
 * to FP32 rounding rather than bit for bit.  Every d_in[k] must stay
 * unchanged until the work completes. */
int rawspec_amd_start_processing_device_batch(rawspec_context *ctx, const void *const *d_in,
                                              unsigned int nbuf, int fft_dir);

/* enable != 0: dumps stay in device memory (no copy into h_pwrbuf/h_icsbuf);
 * the PRE/POST callbacks still fire in stream order, both from one host
 * function per dump (no copy separates them; every host function holds the
 * stream until it returns).  The dumped spectra of
 * product i (same layout as h_pwrbuf[i]) are then valid at
 * rawspec_amd_device_pwrbuf(ctx, i) from POST(i) until the next call that
 * starts processing on this context. */
int rawspec_amd_set_device_output(rawspec_context *ctx, int enable);
float *rawspec_amd_device_pwrbuf(rawspec_context *ctx, int i);
float *rawspec_amd_device_icsbuf(rawspec_context *ctx, int i);

/* The context's HIP stream (a hipStream_t), for callers that order their
 * own device work after the context's. */
void *rawspec_amd_stream(rawspec_context *ctx);

/* Queues the context's work (kernels, dump copies, callbacks) on `stream`, a
 * hipStream_t of the context's GPU, from now on; NULL returns to the
 * context's own stream.  Work already queued is ordered before the work that
 * follows on the new stream.  Several contexts may share one stream (a
 * server keeps its streams within the GPU's hardware queues that way);
 * rawspec_wait_for_completion() then also waits for the other contexts' work
 * on it.  The caller keeps the stream alive while the context uses it.
 * Returns 0 on success. */
int rawspec_amd_set_stream(rawspec_context *ctx, void *stream);

/* Transform engine of the context.  enable = 0: FP32 on the vector units
 * (the default; 1e-5 relative error bar).  enable = 1: the matrix-core engine
 * at its 1e-3 relative error bar, using each kernel where it is faster: FP16
 * matrix products with FP32 accumulation for 1024-point int8 products that
 * integrate at least 32 spectra per dump (unsplit operands: about 2^-12
 * rounding per transform, random from spectrum to spectrum, so the error of
 * an integration falls with its length to a floor near 1.5e-4; shorter
 * integrations stay on the vector units), the four-step intermediate in
 * 24-bit floats (about 2e-5 added error at one spectrum per dump), and the
 * vector units for everything else, the four-step passes included.  An
 * 8-point product whose dumps integrate 128 spectra beside an unsplit
 * 1024-point product of the same detection is computed, in FP32, by that
 * product's kernel from the same read of the samples; batch calls then run
 * the two buffer by buffer, each buffer's dumps in product order.
 * enable = 2: every matrix-core kernel with split operands, the four-step
 * first pass among them (error at the FP32 level).  Other
 * values are refused.  A product longer than 4096 points switches engine at
 * the start of its next integration, so a partial integration is never
 * mixed.  The getter returns the setting, -1 before initialisation. */
int rawspec_amd_set_mfma(rawspec_context *ctx, int enable);
int rawspec_amd_get_mfma(rawspec_context *ctx);

/* Held first-pass output.  A product longer than 4096 points whose
 * integration spans several buffers transforms in two passes; its second
 * pass adds into the running sum of the integration, a read-modify-write of
 * the whole sum.  Called one buffer at a time (rawspec_start_processing(),
 * rawspec_amd_start_processing_device()), such a product keeps each buffer's
 * first-pass output in device memory and runs the second pass once per
 * max_buffers buffers (and at the end of every integration), as a batch call
 * would.  Results, dumps and callbacks are those of the one-buffer path; an
 * input buffer may be reused as soon as its call's work completes.  Default
 * 8; capped by the integration length and by 8 GiB per context, allocated
 * at the first integration that holds and kept until cleanup (1 GiB per
 * held buffer for 64 channels x 2^20 points x 2 pols).  max_buffers < 2
 * turns it off.  A new value applies from the next integration start.
 * Returns 0 on success; the getter returns the setting, -1 before
 * initialisation. */
int rawspec_amd_set_held_buffers(rawspec_context *ctx, unsigned int max_buffers);
int rawspec_amd_get_held_buffers(rawspec_context *ctx);

/* Launch-plan autotuner.  Times candidate launch plans of every product by
 * running the context's own start calls on the device buffer d_in (NULL:
 * the context's input buffer; the data does not affect timing), per_call
 * buffers per call (1: as rawspec_start_processing() and
 * rawspec_amd_start_processing_device(); more: as the batch call), lanes and
 * held output as configured, and keeps for each product in turn the plan
 * under which the whole sequence runs fastest (a candidate must win by 3%).
 * Each timing is `rounds` rounds (0: 3) of 16 buffers after a warm-up round.
 * Small products try the workgroups a launch aims for and the slots per dump
 * group (on the matrix-core engine, the wavefronts a launch aims for);
 * four-step products try the intermediate's budget per launch pair, which
 * sets the channels per launch and the buffers per batch.  Plans change only
 * how work is split over workgroups and launches, so results stay within
 * the same error bars.  The timing runs make no dumps and fire no callbacks;
 * they overwrite the integration buffers, so every integration restarts
 * with the next buffer, as after rawspec_reset_integration().  Candidates
 * may grow the context's scratch buffers while they are timed; on return
 * the scratch buffers are sized for the chosen plans (more than
 * rawspec_amd_plan() reported only where a chosen plan needs more than the
 * default one).  Writes one line per product ("p<i> N<Nt>
 * <plan>: <best> ms (default <ms> ms, <n> candidates)", milliseconds per
 * buffer of the whole sequence) into report (NUL-terminated; NULL with len 0
 * for none).  Returns 0 on success. */
int rawspec_amd_autotune(rawspec_context *ctx, const void *d_in, unsigned int per_call,
                         unsigned int rounds, char *report, size_t len);

/* Two-lane execution: the four-step products' kernels run on one stream and
 * the small products' on another, both forked from and joined back into the
 * context's stream within each start call, so vector-bound small transforms
 * overlap memory-bound four-step passes.  Dumps and callbacks keep their
 * order on the context's stream.  rawspec_initialize() turns it on when the
 * context has both kinds of products and buffers of at least 64 MiB; a
 * process running many contexts at once (more streams than hardware queues)
 * may do better with it off.  Changing it between calls is allowed.
 * rawspec_amd_set_lanes returns 0 on success; rawspec_amd_get_lanes returns
 * 1 or 0 (-1 before initialisation). */
int rawspec_amd_set_lanes(rawspec_context *ctx, int enable);
int rawspec_amd_get_lanes(rawspec_context *ctx);

/* Kernel timing.  With enable != 0 every kernel the context launches is
 * bracketed by HIP events on its stream (the lanes are not used meanwhile,
 * so each pair times one kernel running alone).  rawspec_amd_profile_report()
 * waits for the stream, writes one line per kernel, "p<product>.<kernel>.N<Nt>
 * <launches> <total milliseconds>", into buf (NUL-terminated) and clears the
 * records.  Returns the number of characters written, -1 on error. */
int rawspec_amd_profile(rawspec_context *ctx, int enable);
int rawspec_amd_profile_report(rawspec_context *ctx, char *buf, size_t len);

#ifdef __cplusplus
}
#endif

#endif
