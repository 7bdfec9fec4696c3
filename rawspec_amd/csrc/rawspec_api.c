/*
 * rawspec_api.c -- librawspec's public C API (include/rawspec.h) on top of
 * the gfx950 device layer (rsd.h).
 *
 * This file owns the semantics a drop-in replacement must honour:
 * validation and derived counts (rawspec_gpu.cu:471-641, 732-760), buffer
 * ownership (:694-782, :1300-1391), the per-buffer dump schedule and the
 * PRE/POST callback order (:1517-1747), and completion/reset behaviour
 * (:1762-1836).  It never calls HIP directly.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "rawspec.h"
#include "rawspec_amd.h"
#include "rsd.h"

#define RS_ABS(x) ((x) < 0 ? -(x) : (x))

#ifndef RAWSPEC_AMD_VERSION
#define RAWSPEC_AMD_VERSION "3.0-amd.1"
#endif

typedef struct {
  rawspec_context *ctx;
  int output_product;
  int type;                       /* RAWSPEC_CALLBACK_PRE/POST_DUMP */
} rs_cb_t;

typedef struct {
  rsd_ctx *dev;
  unsigned int inbuf_count;       /* buffers processed since init/reset */
  int caller_managed;             /* h_blkbufs belong to the caller      */
  int expand4;                    /* Nbps was 4                          */
  int device_output;              /* dumps stay in device memory         */
  int tuning;                     /* autotuning: work without dumps      */
  unsigned int Nss[MAX_OUTPUTS];
  unsigned int Nis[MAX_OUTPUTS];
  rs_cb_t pre[MAX_OUTPUTS];
  rs_cb_t post[MAX_OUTPUTS];
} rs_gpu_ctx;

static void rs_err(const char *msg)
{
  fprintf(stderr, "librawspec: %s\n", msg);
  fflush(stderr);
}

const char *get_librawspec_version()
{
  return RAWSPEC_AMD_VERSION;
}

const char *get_cufft_version()
{
  return rsd_engine_name();
}

/* Runs the client callback on the stream's host-function thread
 * (rawspec_gpu.cu:378-402). */
static void rs_cb_trampoline(void *arg)
{
  rs_cb_t *cb = (rs_cb_t *)arg;
  if (cb->ctx->dump_callback)
    cb->ctx->dump_callback(cb->ctx, cb->output_product, cb->type);
}

/* A dump left in device memory has no copies between its PRE and POST
 * callbacks, so one host function makes both calls, in that order.  Each
 * host function holds its stream (and the hardware queue the stream shares
 * with other contexts) until it returns, so this halves those stalls. */
static void rs_cb_pre_post(void *arg)
{
  rs_cb_t *cb = (rs_cb_t *)arg;
  rawspec_context *ctx = cb->ctx;
  if (ctx->dump_callback)
    ctx->dump_callback(ctx, cb->output_product, RAWSPEC_CALLBACK_PRE_DUMP);
  if (ctx->dump_callback)
    ctx->dump_callback(ctx, cb->output_product, RAWSPEC_CALLBACK_POST_DUMP);
}

static unsigned long rs_mod(off_t x, unsigned long n)
{
  long long r = (long long)x % (long long)n;
  return (unsigned long)(r < 0 ? r + (long long)n : r);
}

/* Validation and derived counts, same rules and order as
 * rawspec_gpu.cu:471-641 and :732-760.  Updates the ctx fields the reference
 * updates (Npolout, Nbps, Ntmax, Nb, Nb_host, Nant, Nbc, Nds,
 * h_pwrbuf_size).  Returns 0 when the configuration is valid. */
static int rs_validate(rawspec_context *ctx, int *expand4_out, unsigned long long *nbuf_out)
{
  unsigned int i;
  int expand4;
  unsigned long long nbuf;

  if (ctx->No == 0 || ctx->No > MAX_OUTPUTS) {
    fprintf(stderr, "output products must be in range [1..%d], not %d\n", MAX_OUTPUTS, ctx->No);
    fflush(stderr);
    return 1;
  }
  if (ctx->Np == 0 || ctx->Np > 2) {
    fprintf(stderr, "number of polarizations must be in range [1..2], not %d\n", ctx->Np);
    fflush(stderr);
    return 1;
  }
  for (i = 0; i < ctx->No; i++)
    if (RS_ABS(ctx->Npolout[i]) != 4 || ctx->Np != 2) ctx->Npolout[i] = 1;
  if (ctx->Ntpb == 0) {
    rs_err("number of time samples per block cannot be zero");
    return 1;
  }
  if (ctx->Nbps == 0) ctx->Nbps = 8;
  if (ctx->Nbps != 4 && ctx->Nbps != 8 && ctx->Nbps != 16) {
    fprintf(stderr, "Number of bits per sample in raw header must be 4, 0/8, or 16\n"
                    "Observed a value of %d\n", ctx->Nbps);
    fflush(stderr);
    return 1;
  }
  expand4 = ctx->Nbps == 4;
  if (ctx->Nbps == 4) ctx->Nbps = 8;
  ctx->Ntmax = 0;
  for (i = 0; i < ctx->No; i++) {
    if (ctx->Nts[i] == 0) {
      fprintf(stderr, "Nts[%d] cannot be 0\n", i);
      fflush(stderr);
      return 1;
    }
    if (ctx->Ntmax < ctx->Nts[i]) ctx->Ntmax = ctx->Nts[i];
  }
  for (i = 0; i < ctx->No; i++) {
    if (ctx->Ntmax % ctx->Nts[i] != 0) {
      fprintf(stderr, "Nts[%d] (%u) is not a factor of Ntmax (%u)\n", i, ctx->Nts[i], ctx->Ntmax);
      fflush(stderr);
      return 1;
    }
  }
  if (ctx->Nb != 0) {
    if (((unsigned long long)ctx->Nb * ctx->Ntpb) % ctx->Ntmax != 0) {
      fprintf(stderr, "Ntmax (%u) is not a factor of Nb*Ntpb (%u * %u)\n", ctx->Ntmax, ctx->Nb,
              ctx->Ntpb);
      fflush(stderr);
      return 1;
    }
  } else if (ctx->Ntmax < ctx->Ntpb) {
    if (ctx->Ntpb % ctx->Ntmax != 0) {
      fprintf(stderr, "Ntmax (%u) is not a factor of Ntpb (%u)\n", ctx->Ntmax, ctx->Ntpb);
      fflush(stderr);
      return 1;
    }
    ctx->Nb = 1;
  } else {
    if (ctx->Ntmax % ctx->Ntpb != 0) {
      fprintf(stderr, "Ntpb (%u) is not a factor of Nmax (%u)\n", ctx->Ntpb, ctx->Ntmax);
      fflush(stderr);
      return 1;
    }
    ctx->Nb = ctx->Ntmax / ctx->Ntpb;
  }
  if (ctx->Nb_host == 0 && ctx->h_blkbufs) {
    rs_err("Must specify number of host input blocks when caller-managed");
    return 1;
  } else if (ctx->Nb_host == 0) {
    ctx->Nb_host = ctx->Nb;
  }
  nbuf = (unsigned long long)ctx->Nb * ctx->Ntpb;
  for (i = 0; i < ctx->No; i++) {
    unsigned long long ntna = (unsigned long long)ctx->Nts[i] * ctx->Nas[i];
    if (ctx->Nas[i] == 0) {
      fprintf(stderr, "Nas[%d] cannot be 0\n", i);
      fflush(stderr);
      return 1;
    }
    if (ntna < nbuf ? (nbuf % ntna != 0) : (ntna % nbuf != 0)) {
      fprintf(stderr, "Nts[%d] * Nas[%d] (%u * %u) and Nb * Ntpb (%u * %u) must divide one another\n",
              i, i, ctx->Nts[i], ctx->Nas[i], ctx->Nb, ctx->Ntpb);
      fflush(stderr);
      return 1;
    }
  }
  if (nbuf > 0xffffffffull) {
    rs_err("Nb * Ntpb exceeds 2^32 samples per channel");
    return 1;
  }
  if (ctx->Nant == 0) ctx->Nant = 1;
  if (ctx->Nbc == 0) {
    ctx->Nbc = ctx->Nc;
  } else {
    if (ctx->Nbc <= 1) {
      if (ctx->Nant > 1) {
        ctx->Nbc = ctx->Nc / ctx->Nant;
      } else {
        for (i = 1; i <= 10; i++)
          if (ctx->Nc % i == 0) ctx->Nbc = ctx->Nc / i;
      }
    }
    if (ctx->Nbc == 0 || ctx->Nc % ctx->Nbc != 0) {
      fprintf(stderr, "%d channels cannot be factorised to batches of %d\n", ctx->Nc, ctx->Nbc);
      fflush(stderr);
      return 1;
    }
    printf("Batching %d channels into %d batches of %d.\n", ctx->Nc, ctx->Nc / ctx->Nbc, ctx->Nbc);
  }
  if (ctx->incoherently_sum) {
    if (!((ctx->Naws == 1 && (unsigned)ctx->Naws < ctx->Nant) ||
          (ctx->Naws > 0 && (unsigned)ctx->Naws == ctx->Nant)) || !ctx->Aws) {
      fprintf(stderr, "Not enough antenna-weights provided for the %d antennas: only provided %d.\n",
              ctx->Nant, ctx->Naws);
      fflush(stderr);
      return 1;
    }
  }
  /* derived counts (rawspec_gpu.cu:732-760), before any device work */
  for (i = 0; i < ctx->No; i++) {
    unsigned int nss = (unsigned)(nbuf / ctx->Nts[i]);
    ctx->Nds[i] = nss / ctx->Nas[i];
    if (ctx->Nds[i] == 0) ctx->Nds[i] = 1;
    ctx->h_pwrbuf_size[i] = (size_t)RS_ABS(ctx->Npolout[i]) * ctx->Nds[i] * ctx->Nts[i] *
                            ctx->Nc * sizeof(float);
  }
  *expand4_out = expand4;
  *nbuf_out = nbuf;
  return 0;
}

/* The device layer's view of a validated context (weights excluded). */
static void rs_fill_config(const rawspec_context *ctx, int expand4, unsigned long long nbuf,
                           rsd_config *cfg)
{
  unsigned int i;
  memset(cfg, 0, sizeof(*cfg));
  cfg->gpu_index = ctx->gpu_index;
  cfg->Nc = ctx->Nc;
  cfg->Np = ctx->Np;
  cfg->Ntpb = ctx->Ntpb;
  cfg->Nb = ctx->Nb;
  cfg->Nbps = ctx->Nbps;
  cfg->expand4 = expand4;
  cfg->Nant = ctx->Nant;
  cfg->conj = ctx->input_conjugated ? 1 : 0;
  cfg->ics = ctx->incoherently_sum ? 1 : 0;
  cfg->No = ctx->No;
  for (i = 0; i < ctx->No; i++) {
    unsigned int nss = (unsigned)(nbuf / ctx->Nts[i]);
    unsigned int nis = ctx->Nas[i] / nss;
    cfg->prod[i].Nt = ctx->Nts[i];
    cfg->prod[i].Na = ctx->Nas[i];
    cfg->prod[i].Nss = nss;
    cfg->prod[i].Nds = ctx->Nds[i];
    cfg->prod[i].Nis = nis ? nis : 1;
    cfg->prod[i].npolout = ctx->Npolout[i];
  }
}

int rawspec_initialize(rawspec_context *ctx)
{
  unsigned int i;
  int expand4;
  rs_gpu_ctx *g;
  rsd_config cfg;
  float *weights = NULL;
  unsigned long long nbuf;

  if (rs_validate(ctx, &expand4, &nbuf)) return 1;

  /* ---- allocation (rawspec_gpu.cu:643-782) ---- */
  for (i = 0; i < MAX_OUTPUTS; i++) {
    ctx->h_pwrbuf[i] = NULL;
    ctx->h_icsbuf[i] = NULL;
  }
  ctx->gpu_ctx = NULL;

  g = (rs_gpu_ctx *)calloc(1, sizeof(*g));
  if (!g) {
    rs_err("unable to allocate the context");
    return 1;
  }
  ctx->gpu_ctx = g;
  g->expand4 = expand4;
  for (i = 0; i < MAX_OUTPUTS; i++) {
    g->pre[i].ctx = ctx;
    g->pre[i].output_product = (int)i;
    g->pre[i].type = RAWSPEC_CALLBACK_PRE_DUMP;
    g->post[i].ctx = ctx;
    g->post[i].output_product = (int)i;
    g->post[i].type = RAWSPEC_CALLBACK_POST_DUMP;
  }

  if (!ctx->h_blkbufs) {
    size_t bsz = (size_t)ctx->Ntpb * ctx->Np * ctx->Nc * 2 * (ctx->Nbps / 8);
    g->caller_managed = 0;
    ctx->h_blkbufs = (char **)calloc(ctx->Nb_host, sizeof(char *));
    if (!ctx->h_blkbufs) goto fail;
    for (i = 0; i < ctx->Nb_host; i++)
      if (rsd_host_alloc(ctx->gpu_index, (void **)&ctx->h_blkbufs[i], bsz, 1)) goto fail;
  } else {
    g->caller_managed = 1;
  }

  for (i = 0; i < ctx->No; i++) {
    g->Nss[i] = (unsigned)(nbuf / ctx->Nts[i]);
    g->Nis[i] = ctx->Nas[i] / g->Nss[i];
    if (g->Nis[i] == 0) g->Nis[i] = 1;
    if (rsd_host_alloc(ctx->gpu_index, (void **)&ctx->h_pwrbuf[i], ctx->h_pwrbuf_size[i], 0))
      goto fail;
    if (ctx->incoherently_sum == 1) {
      if (rsd_host_alloc(ctx->gpu_index, (void **)&ctx->h_icsbuf[i],
                         ctx->h_pwrbuf_size[i] / ctx->Nant, 0))
        goto fail;
    }
  }

  rs_fill_config(ctx, expand4, nbuf, &cfg);
  if (cfg.ics) {
    /* one broadcast weight or one per antenna (rawspec_gpu.cu:1040-1055) */
    weights = (float *)malloc(ctx->Nant * sizeof(float));
    if (!weights) goto fail;
    for (i = 0; i < ctx->Nant; i++) weights[i] = ctx->Naws == 1 ? ctx->Aws[0] : ctx->Aws[i];
    cfg.weights = weights;
  }
  if (rsd_create(&cfg, &g->dev)) goto fail;
  free(weights);
  return 0;

fail:
  free(weights);
  rawspec_cleanup(ctx);
  return 1;
}

void rawspec_cleanup(rawspec_context *ctx)
{
  unsigned int i;
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;

  /* finish outstanding work before host buffers disappear under it */
  if (g && g->dev) rsd_sync(g->dev);
  for (i = 0; i < MAX_OUTPUTS; i++) {
    if (ctx->h_pwrbuf[i]) {
      rsd_host_free(ctx->h_pwrbuf[i]);
      ctx->h_pwrbuf[i] = NULL;
    }
    if (ctx->h_icsbuf[i]) {
      rsd_host_free(ctx->h_icsbuf[i]);
      ctx->h_icsbuf[i] = NULL;
    }
  }
  if (g) {
    if (!g->caller_managed && ctx->h_blkbufs) {
      for (i = 0; i < ctx->Nb_host; i++)
        if (ctx->h_blkbufs[i]) rsd_host_free(ctx->h_blkbufs[i]);
      free(ctx->h_blkbufs);
      ctx->h_blkbufs = NULL;
    }
    if (g->dev) rsd_destroy(g->dev);
    free(g);
    ctx->gpu_ctx = NULL;
  }
}

static rs_gpu_ctx *rs_get(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  if (!g || !g->dev) {
    rs_err("context is not initialised");
    return NULL;
  }
  return g;
}

int rawspec_copy_blocks_to_gpu(rawspec_context *ctx, off_t src_idx, off_t dst_idx,
                               size_t num_blocks)
{
  size_t b;
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  for (b = 0; b < num_blocks; b++) {
    unsigned long sblk = rs_mod(src_idx + (off_t)b, ctx->Nb_host);
    unsigned long dblk = rs_mod(dst_idx + (off_t)b, ctx->Nb);
    if (rsd_copy_block(g->dev, ctx->h_blkbufs[sblk], (unsigned)dblk)) return 1;
  }
  return 0;
}

/* Host blocks hold complex4 bytes (half the 8-bit block size); each is
 * expanded on the device into device block (dst_idx + b) % Nb.  The
 * reference expands into positions that assume num_blocks == Nb and
 * dst_idx == 0 (rawspec_gpu.cu:1426-1434), the only way its CLI calls it;
 * this implementation places every block at its own slot. */
int rawspec_copy_blocks_to_gpu_expanding_complex4(rawspec_context *ctx, off_t src_idx,
                                                  off_t dst_idx, size_t num_blocks)
{
  size_t b;
  rs_gpu_ctx *g;
  if (num_blocks > ctx->Nb) {
    fprintf(stderr, "%s: num_blocks (%lu) > Nb (%u)\n", __func__, (unsigned long)num_blocks,
            ctx->Nb);
    fflush(stderr);
    return 1;
  }
  g = rs_get(ctx);
  if (!g) return 1;
  if (!g->expand4) {
    rs_err("complex4 expansion requested but the context was not initialised with Nbps = 4");
    return 1;
  }
  for (b = 0; b < num_blocks; b++) {
    unsigned long sblk = rs_mod(src_idx + (off_t)b, ctx->Nb_host);
    unsigned long dblk = rs_mod(dst_idx + (off_t)b, ctx->Nb);
    if (rsd_copy_block4(g->dev, ctx->h_blkbufs[sblk], (unsigned)dblk)) return 1;
  }
  return 0;
}

int rawspec_zero_blocks_to_gpu(rawspec_context *ctx, off_t dst_idx, size_t num_blocks)
{
  size_t b;
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  for (b = 0; b < num_blocks; b++) {
    unsigned long dblk = rs_mod(dst_idx + (off_t)b, ctx->Nb);
    if (rsd_zero_block(g->dev, (unsigned)dblk)) return 1;
  }
  return 0;
}

/* One buffer through every product, in product order; a product whose
 * integration completes is dumped between PRE and POST
 * (rawspec_gpu.cu:1543-1747).  PRE is enqueued before any copy into the
 * host buffers, so a client that joins its output thread in PRE never races
 * the copy.  Host callbacks stall the stream while they run, so none is
 * enqueued when the context has no dump_callback (the reference would call
 * nothing either). */
/* Enqueues the dump of product i: ICS, PRE callback, copy to h_pwrbuf /
 * h_icsbuf (unless dumps stay on the device), POST callback
 * (rawspec_gpu.cu:1603-1709). */
static int rs_dump(rawspec_context *ctx, rs_gpu_ctx *g, unsigned int i)
{
  if (ctx->incoherently_sum && rsd_run_ics(g->dev, (int)i)) return 1;
  if (g->device_output)
    return ctx->dump_callback && rsd_host_fn(g->dev, rs_cb_pre_post, &g->pre[i]);
  if (ctx->dump_callback && rsd_host_fn(g->dev, rs_cb_trampoline, &g->pre[i])) return 1;
  if (rsd_dump_to_host(g->dev, (int)i, ctx->h_pwrbuf[i], ctx->h_pwrbuf_size[i],
                       ctx->incoherently_sum ? ctx->h_icsbuf[i] : NULL,
                       ctx->h_pwrbuf_size[i] / ctx->Nant))
    return 1;
  if (ctx->dump_callback && rsd_host_fn(g->dev, rs_cb_trampoline, &g->post[i])) return 1;
  return 0;
}

/* Whether a dump queues work on the context stream: an incoherent sum, the
 * copies into the host buffers or a callback.  Device-resident output with
 * no callback and no incoherent sum queues nothing, so a batch call need not
 * join the product's lane into the context stream at each such dump (a join
 * followed by the lane's resume would also hold the lane behind everything
 * the context stream waited for, the other lane's work included). */
static int rs_dump_queues(const rawspec_context *ctx, const rs_gpu_ctx *g)
{
  return ctx->incoherently_sum || !g->device_output || ctx->dump_callback != NULL;
}

/* A start call that fails after forking the lanes joins them back into the
 * context stream (so rawspec_wait_for_completion() also waits for the
 * kernels already queued there, which read the input buffer) and takes back
 * the buffers it counted (count: how many). */
static int rs_abandon(rs_gpu_ctx *g, unsigned int count)
{
  (void)rsd_lane_end(g->dev, -1);
  g->inbuf_count -= count;
  return 1;
}

/* The products' kernels run on the device layer's lanes (four-step and small
 * products side by side); every dump waits for its product's lane and runs,
 * in product order, on the context stream, which waits for both lanes before
 * the call returns. */
static int rs_start(rawspec_context *ctx, const void *d_in, int fft_dir)
{
  unsigned int i;
  int last[MAX_OUTPUTS];
  int fa = -1, fb = -1;
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  const int fused = rsd_fused_pair(g->dev, &fa, &fb);
  if (rsd_lanes_begin(g->dev)) return 1;
  g->inbuf_count++;
  for (i = 0; i < ctx->No; i++) {
    int first = ((g->inbuf_count - 1) % g->Nis[i]) == 0;
    last[i] = (g->inbuf_count % g->Nis[i]) == 0;
    if (fused && ((int)i == fa || (int)i == fb)) {
      /* the pair runs as one, at its first product */
      if ((int)i != (fa < fb ? fa : fb)) continue;
      if (rsd_run_fused(g->dev, d_in, fft_dir <= 0,
                        ((g->inbuf_count - 1) % g->Nis[fb]) == 0)) {
        fprintf(stderr, "librawspec: processing output products %d and %d failed\n", fa, fb);
        fflush(stderr);
        return rs_abandon(g, 1);
      }
      continue;
    }
    if (rsd_run_product(g->dev, (int)i, d_in, fft_dir <= 0, first, last[i])) {
      fprintf(stderr, "librawspec: processing output product %u failed\n", i);
      fflush(stderr);
      return rs_abandon(g, 1);
    }
  }
  for (i = 0; i < ctx->No; i++)
    if (last[i] && !g->tuning && (rsd_lane_end(g->dev, (int)i) || rs_dump(ctx, g, i)))
      return rs_abandon(g, 1);
  return rsd_lane_end(g->dev, -1);
}

/* nbuf consecutive buffers: per product, the run is cut at integration
 * boundaries; each piece goes to the device layer in one call (which
 * batches it where the product's kernels support it), followed by the dump
 * when the piece completes an integration. */
static int rs_start_batch(rawspec_context *ctx, const void *const *d_in, unsigned int nbuf,
                          int fft_dir)
{
  unsigned int i;
  int fa = -1, fb = -1;
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  const int fused = rsd_fused_pair(g->dev, &fa, &fb);
  const int dumps = !g->tuning && rs_dump_queues(ctx, g);
  if (rsd_lanes_begin(g->dev)) return 1;
  for (i = 0; i < ctx->No; i++) {
    unsigned int k = 0;
    if (fused && ((int)i == fa || (int)i == fb)) {
      /* the pair runs as one, buffer by buffer, at its first product; each
       * buffer's dumps follow it in product order */
      const int lo = fa < fb ? fa : fb, hi = fa < fb ? fb : fa;
      if ((int)i != lo) continue;
      for (k = 0; k < nbuf; k++) {
        const unsigned int n = g->inbuf_count + k;   /* buffers before this one */
        int dumped = 0, j;
        if (rsd_run_fused(g->dev, d_in[k], fft_dir <= 0, (n % g->Nis[fb]) == 0)) {
          fprintf(stderr, "librawspec: processing output products %d and %d failed\n", fa, fb);
          fflush(stderr);
          return rs_abandon(g, 0);
        }
        for (j = 0; j < 2; j++) {
          const int x = j ? hi : lo;
          if ((n + 1) % g->Nis[x] != 0 || !dumps) continue;
          if (rsd_lane_end(g->dev, x) || rs_dump(ctx, g, (unsigned int)x)) return rs_abandon(g, 0);
          dumped = 1;
        }
        /* the dumps read integration buffers the next buffer rewrites */
        if (dumped && k + 1 < nbuf && rsd_lane_resume(g->dev, lo)) return rs_abandon(g, 0);
      }
      continue;
    }
    while (k < nbuf) {
      const unsigned int pos = (g->inbuf_count + k) % g->Nis[i];   /* buffers into the integration */
      const unsigned int left = g->Nis[i] - pos;
      const unsigned int len = left < nbuf - k ? left : nbuf - k;
      const int last = len == left;
      if (rsd_run_product_batch(g->dev, (int)i, d_in + k, len, fft_dir <= 0, pos == 0, last)) {
        fprintf(stderr, "librawspec: processing output product %u failed\n", i);
        fflush(stderr);
        return rs_abandon(g, 0);
      }
      /* the dump reads the integration buffer the lane's next piece rewrites */
      if (last && dumps &&
          (rsd_lane_end(g->dev, (int)i) || rs_dump(ctx, g, i) ||
           (k + len < nbuf && rsd_lane_resume(g->dev, (int)i))))
        return rs_abandon(g, 0);
      k += len;
    }
  }
  g->inbuf_count += nbuf;
  return rsd_lane_end(g->dev, -1);
}

int rawspec_start_processing(rawspec_context *ctx, int fft_dir)
{
  return rs_start(ctx, NULL, fft_dir);
}

int rawspec_copy_blocks_to_gpu_and_start_processing(rawspec_context *ctx, size_t num_blocks,
                                                    char expand4bps_to8bps, int fft_dir)
{
  int rc = expand4bps_to8bps ? rawspec_copy_blocks_to_gpu_expanding_complex4(ctx, 0, 0, num_blocks)
                             : rawspec_copy_blocks_to_gpu(ctx, 0, 0, num_blocks);
  if (rc) return rc;
  return rawspec_start_processing(ctx, fft_dir);
}

int rawspec_reset_integration(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  if (!g || !g->dev) return 1;
  rawspec_wait_for_completion(ctx);
  /* The next buffer overwrites (rather than adds to) every integration
   * buffer, so dropping the partial integrations is just the count. */
  g->inbuf_count = 0;
  return 0;
}

unsigned int rawspec_check_for_completion(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  if (!g || !g->dev) return 0;
  return rsd_query(g->dev) == 1 ? ctx->No : 0;
}

int rawspec_wait_for_completion(rawspec_context *ctx)
{
  unsigned int i;
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  for (i = 0; i < ctx->No; i++)
    if (ctx->dump_callback && rsd_host_fn(g->dev, rs_cb_trampoline, &g->pre[i])) return 1;
  if (rsd_sync(g->dev)) return 1;
  if (*(volatile unsigned int *)&ctx->exit_soon)
    fprintf(stderr, "*** rawspec_wait_for_completion detected exit_soon enabled.\n"
                    "*** Cannot continue!\n");
  return (int)*(volatile unsigned int *)&ctx->exit_soon;
}

/* ---- extensions (rawspec_amd.h) ---------------------------------------- */

void *rawspec_amd_device_input(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  return (g && g->dev) ? rsd_input_ptr(g->dev) : NULL;
}

size_t rawspec_amd_device_input_size(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  return (g && g->dev) ? rsd_input_bytes(g->dev) : 0;
}

int rawspec_amd_start_processing_device(rawspec_context *ctx, const void *d_in, int fft_dir)
{
  if (!d_in || ((size_t)d_in & 15) != 0) {
    rs_err("rawspec_amd_start_processing_device: input must be a 16-byte aligned device pointer");
    return 1;
  }
  return rs_start(ctx, d_in, fft_dir);
}

int rawspec_amd_start_processing_device_noalloc(rawspec_context *ctx, const void *d_in,
                                                int fft_dir)
{
  int rc;
  rs_gpu_ctx *g;
  if (!d_in || ((size_t)d_in & 15) != 0) {
    rs_err("rawspec_amd_start_processing_device_noalloc: input must be a 16-byte aligned device pointer");
    return 1;
  }
  g = rs_get(ctx);
  if (!g) return 1;
  if (rsd_noalloc_begin(g->dev)) return 1;
  rc = rs_start(ctx, d_in, fft_dir);
  rsd_noalloc_end(g->dev);
  return rc;
}

int rawspec_amd_start_processing_device_batch(rawspec_context *ctx, const void *const *d_in,
                                              unsigned int nbuf, int fft_dir)
{
  unsigned int k;
  if (!d_in || nbuf == 0) {
    rs_err("rawspec_amd_start_processing_device_batch: need at least one input buffer");
    return 1;
  }
  for (k = 0; k < nbuf; k++) {
    if (!d_in[k] || ((size_t)d_in[k] & 15) != 0) {
      rs_err("rawspec_amd_start_processing_device_batch: inputs must be 16-byte aligned device pointers");
      return 1;
    }
  }
  return rs_start_batch(ctx, d_in, nbuf, fft_dir);
}

int rawspec_amd_set_device_output(rawspec_context *ctx, int enable)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  g->device_output = enable ? 1 : 0;
  return 0;
}

float *rawspec_amd_device_pwrbuf(rawspec_context *ctx, int i)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  if (!g || !g->dev || i < 0 || (unsigned)i >= ctx->No) return NULL;
  return rsd_pwr_ptr(g->dev, i);
}

float *rawspec_amd_device_icsbuf(rawspec_context *ctx, int i)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  if (!g || !g->dev || i < 0 || (unsigned)i >= ctx->No) return NULL;
  return rsd_ics_ptr(g->dev, i);
}

void *rawspec_amd_stream(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  return (g && g->dev) ? rsd_stream(g->dev) : NULL;
}

int rawspec_amd_set_stream(rawspec_context *ctx, void *stream)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  return rsd_set_stream(g->dev, stream);
}

int rawspec_amd_profile(rawspec_context *ctx, int enable)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  return rsd_profile(g->dev, enable);
}

int rawspec_amd_set_mfma(rawspec_context *ctx, int enable)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  return rsd_set_mfma(g->dev, enable);
}

/* Every start call joins the lanes back into the context stream before it
 * returns, so the setting may change between any two calls. */
int rawspec_amd_set_lanes(rawspec_context *ctx, int enable)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  return rsd_set_lanes(g->dev, enable);
}

int rawspec_amd_get_lanes(rawspec_context *ctx)
{
  rs_gpu_ctx *g = (rs_gpu_ctx *)ctx->gpu_ctx;
  return g && g->dev ? rsd_get_lanes(g->dev) : -1;
}

int rawspec_amd_get_mfma(rawspec_context *ctx)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return -1;
  return rsd_get_mfma(g->dev);
}

/* Seconds per buffer of `rounds` rounds of RS_TUNE_BUFS buffers through the
 * start calls a caller feeding `per_call` buffers per call makes, after one
 * round of warm-up; each round starts new integrations. */
#define RS_TUNE_BUFS 16
static double rs_tune_time(rawspec_context *ctx, rs_gpu_ctx *g, const void *d_in,
                           unsigned int per_call, unsigned int rounds)
{
  const void *ins[RS_TUNE_BUFS];
  struct timespec t0, t1;
  unsigned int r, k;
  for (k = 0; k < RS_TUNE_BUFS; k++) ins[k] = d_in;
  for (r = 0; r <= rounds; r++) {
    if (r == 1) {
      if (rsd_sync(g->dev)) return -1.0;
      clock_gettime(CLOCK_MONOTONIC, &t0);
    }
    g->inbuf_count = 0;
    for (k = 0; k < RS_TUNE_BUFS; k += per_call) {
      const unsigned int n = RS_TUNE_BUFS - k < per_call ? RS_TUNE_BUFS - k : per_call;
      if (n == 1 ? rs_start(ctx, d_in, RAWSPEC_FORWARD_FFT)
                 : rs_start_batch(ctx, (const void *const *)(ins + k), n, RAWSPEC_FORWARD_FFT))
        return -1.0;
    }
  }
  if (rsd_sync(g->dev)) return -1.0;
  clock_gettime(CLOCK_MONOTONIC, &t1);
  return ((double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec)) /
         ((double)rounds * RS_TUNE_BUFS);
}

/* Coordinate descent over the products' launch-plan candidates: each
 * product in turn takes the candidate under which the whole call sequence
 * runs fastest, the other products keeping their current plans. */
int rawspec_amd_autotune(rawspec_context *ctx, const void *d_in, unsigned int per_call,
                         unsigned int rounds, char *report, size_t len)
{
  unsigned int i;
  size_t off = 0;
  int rc = 0;
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  if (!d_in) d_in = rsd_input_ptr(g->dev);
  if (per_call == 0) per_call = 1;
  if (per_call > RS_TUNE_BUFS) per_call = RS_TUNE_BUFS;
  if (rounds == 0) rounds = 3;
  if (report && len) report[0] = '\0';
  if (rsd_sync(g->dev)) return 1;
  g->tuning = 1;
  for (i = 0; i < ctx->No && !rc; i++) {
    const int n = rsd_tune_count(g->dev, (int)i);
    int k, best = 0;
    double t_best, t_dflt;
    char plan[160];
    rc = rsd_tune_apply(g->dev, (int)i, 0);
    t_dflt = t_best = rc ? -1.0 : rs_tune_time(ctx, g, d_in, per_call, rounds);
    if (t_best < 0) rc = 1;
    for (k = 1; k < n && !rc; k++) {
      double t;
      const int a = rsd_tune_apply(g->dev, (int)i, k);
      if (a == 2) continue;              /* the same plan as an earlier candidate */
      if (a) { rc = 1; break; }
      t = rs_tune_time(ctx, g, d_in, per_call, rounds);
      if (t < 0) { rc = 1; break; }
      /* a new plan must win clearly: timing noise is a few percent */
      if (t < 0.97 * t_best) { t_best = t; best = k; }
    }
    if (!rc) rc = rsd_tune_apply(g->dev, (int)i, best) == 1;
    if (!rc && report && len > off) {
      int w;
      rsd_tune_describe(g->dev, (int)i, plan, sizeof plan);
      w = snprintf(report + off, len - off, "p%u N%u %s: %.4f ms (default %.4f ms, %d candidates)\n",
                   i, ctx->Nts[i], plan, t_best * 1e3, t_dflt * 1e3, n);
      if (w > 0) off += (size_t)w < len - off ? (size_t)w : len - off - 1;
    }
  }
  g->tuning = 0;
  /* the timing runs overwrote the integration buffers */
  g->inbuf_count = 0;
  /* candidates may have grown the scratch buffers beyond the chosen plans'
   * needs (and beyond what rawspec_amd_plan reports) */
  if (rsd_tune_fit(g->dev)) rc = 1;
  return rc;
}

int rawspec_amd_set_held_buffers(rawspec_context *ctx, unsigned int max_buffers)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return 1;
  return rsd_set_hold(g->dev, max_buffers);
}

int rawspec_amd_get_held_buffers(rawspec_context *ctx)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return -1;
  return (int)rsd_get_hold(g->dev);
}

int rawspec_amd_profile_report(rawspec_context *ctx, char *buf, size_t len)
{
  rs_gpu_ctx *g = rs_get(ctx);
  if (!g) return -1;
  return rsd_profile_report(g->dev, buf, len);
}

int rawspec_amd_plan(rawspec_context *ctx, size_t *device_bytes, size_t *host_bytes)
{
  int expand4;
  unsigned long long nbuf;
  unsigned int i;
  rsd_config cfg;
  size_t d, h = 0;
  if (rs_validate(ctx, &expand4, &nbuf)) return 1;
  rs_fill_config(ctx, expand4, nbuf, &cfg);
  d = rsd_plan_bytes(&cfg);
  if (d == 0) return 1;
  if (!ctx->h_blkbufs)
    h += (size_t)ctx->Nb_host * ctx->Ntpb * ctx->Np * ctx->Nc * 2 * (ctx->Nbps / 8);
  for (i = 0; i < ctx->No; i++) {
    h += ctx->h_pwrbuf_size[i];
    if (ctx->incoherently_sum == 1) h += ctx->h_pwrbuf_size[i] / ctx->Nant;
  }
  if (device_bytes) *device_bytes = d;
  if (host_bytes) *host_bytes = h;
  return 0;
}
