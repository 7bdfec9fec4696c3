/*
 * cli_ingest.c -- moves the blocks of a GUPPI RAW stem into the
 * spectrometer.
 *
 * The reference CLI reads each block with one read() into a pinned host
 * block and, once a buffer of Nb blocks is complete, waits for the GPU,
 * copies the buffer synchronously and starts processing
 * (rawspec.c:905-1019).  Reads, copy and compute then take turns.  Here:
 *
 *  - Reads are deferred until a buffer is complete and then done by a pool
 *    of threads with preadv(), one job per run of consecutive channel rows
 *    (a block's selected channels are contiguous in the file).
 *  - Pipelined mode reads straight into the device layout
 *    [chan][Nb*Ntpb][pol][re,im] of one of several pinned host slots, copies
 *    the whole slot with one hipMemcpyAsync on a copy stream into one of two
 *    device buffers, and starts processing on the library's stream once the
 *    copy has landed (stream-ordered by events).  Reading buffer k+1 thus
 *    overlaps the copy and the kernels of buffer k.
 *  - Synchronous mode keeps the reference's order through the standard API
 *    (wait, copy, start): the library expands 4-bit input during that copy.
 *
 * Gaps in PKTIDX become zero blocks; both modes fill them with zeros in
 * host memory, as the reference does.
 */
#define _GNU_SOURCE 1
#include <errno.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/uio.h>
#include <time.h>
#include <unistd.h>

#include <hip/hip_runtime_api.h>

#include "cli.h"
#include "rawspec_amd.h"

enum { kJobBytes = 4 << 20, kMaxRows = 256, kDevBufs = 2 };

typedef struct {
  int fd;                /* -1: zero rows */
  off_t off;             /* file offset of the first row */
  unsigned char *dst;    /* destination of the first row */
  size_t stride;         /* destination distance between rows */
  unsigned int nrows;
} job;

typedef struct {
  unsigned char *h;      /* pinned, device layout */
  hipEvent_t copied;
  int busy;
} hslot;

struct cli_ingest {
  rawspec_context *ctx;
  int pipelined;
  int expand4;
  size_t cb;             /* file bytes per channel row of one block */
  size_t buf_bytes;      /* Nc * Nb * cb */
  unsigned int Nb, Nc;

  /* deferred work of the buffer being filled */
  job *jobs;
  size_t njobs, cap_jobs;
  int *fds;
  size_t nfds, cap_fds;
  int bad;               /* a job could not be queued */

  /* read pool */
  int nthreads;
  pthread_t *threads;
  pthread_mutex_t mu;
  pthread_cond_t go, done;
  unsigned long gen;
  int idle;              /* readers done with the current generation */
  size_t next;
  int read_error;
  int quit;

  /* pipelined staging */
  hslot *slots;
  int nslots, cur;
  void *dbuf[kDevBufs];
  hipEvent_t consumed[kDevBufs];
  int dused[kDevBufs];
  hipStream_t copy_stream;
  unsigned long nbuf;
  uint64_t bytes;
  double read_s;
  double issue_s;        /* host time spent starting copies and processing */
};

static double now_s(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

#define HIPOK(call)                                                              \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      fprintf(stderr, "rawspec: %s failed: %s\n", #call, hipGetErrorString(e_)); \
      return 1;                                                                  \
    }                                                                            \
  } while (0)

/* ---- read pool ------------------------------------------------------------ */

static int run_job(cli_ingest *g, const job *j)
{
  struct iovec iov[kMaxRows];
  unsigned int r, first = 0;
  size_t want, got = 0;
  if (j->fd < 0) {
    for (r = 0; r < j->nrows; r++) memset(j->dst + r * j->stride, 0, g->cb);
    return 0;
  }
  want = (size_t)j->nrows * g->cb;
  for (r = 0; r < j->nrows; r++) {
    iov[r].iov_base = j->dst + r * j->stride;
    iov[r].iov_len = g->cb;
  }
  while (got < want) {
    ssize_t n = preadv(j->fd, iov + first, (int)(j->nrows - first), j->off + (off_t)got);
    if (n < 0) {
      if (errno == EINTR) continue;
      perror("read");
      return 1;
    }
    if (n == 0) {
      fprintf(stderr, "incomplete block at EOF\n");
      return 1;
    }
    got += (size_t)n;
    /* advance the iovecs past what was read */
    while (n > 0 && first < j->nrows) {
      if ((size_t)n >= iov[first].iov_len) {
        n -= (ssize_t)iov[first].iov_len;
        iov[first].iov_len = 0;
        first++;
      } else {
        iov[first].iov_base = (char *)iov[first].iov_base + n;
        iov[first].iov_len -= (size_t)n;
        n = 0;
      }
    }
  }
  return 0;
}

/* Takes jobs until none is left. */
static void drain(cli_ingest *g)
{
  for (;;) {
    size_t k = __atomic_fetch_add(&g->next, 1, __ATOMIC_RELAXED);
    if (k >= g->njobs) break;
    if (run_job(g, &g->jobs[k])) __atomic_store_n(&g->read_error, 1, __ATOMIC_RELAXED);
  }
}

/* A reader drains each generation of jobs, then reports itself idle; the
 * next generation starts only once every reader is idle, so no reader ever
 * sees the job list while it is being rebuilt. */
static void *reader(void *arg)
{
  cli_ingest *g = (cli_ingest *)arg;
  unsigned long seen = 0;
  for (;;) {
    pthread_mutex_lock(&g->mu);
    while (g->gen == seen && !g->quit) pthread_cond_wait(&g->go, &g->mu);
    if (g->quit) {
      pthread_mutex_unlock(&g->mu);
      return NULL;
    }
    seen = g->gen;
    pthread_mutex_unlock(&g->mu);
    drain(g);
    pthread_mutex_lock(&g->mu);
    if (++g->idle == g->nthreads) pthread_cond_signal(&g->done);
    pthread_mutex_unlock(&g->mu);
  }
}

/* Runs every queued job (pool plus caller) and waits for all of them. */
static int run_jobs(cli_ingest *g)
{
  double t0 = now_s();
  if (g->njobs == 0) return 0;
  g->next = 0;
  g->read_error = 0;
  pthread_mutex_lock(&g->mu);
  g->gen++;
  g->idle = 0;
  pthread_cond_broadcast(&g->go);
  pthread_mutex_unlock(&g->mu);
  drain(g);
  pthread_mutex_lock(&g->mu);
  while (g->idle < g->nthreads) pthread_cond_wait(&g->done, &g->mu);
  pthread_mutex_unlock(&g->mu);
  g->read_s += now_s() - t0;
  return g->read_error;
}

/* ---- queueing --------------------------------------------------------------- */

static int push_job(cli_ingest *g, job j)
{
  if (g->njobs == g->cap_jobs) {
    size_t cap = g->cap_jobs ? 2 * g->cap_jobs : 256;
    job *p = (job *)realloc(g->jobs, cap * sizeof *p);
    if (!p) return 1;
    g->jobs = p;
    g->cap_jobs = cap;
  }
  g->jobs[g->njobs++] = j;
  return 0;
}

static int push_fd(cli_ingest *g, int fd)
{
  if (g->nfds == g->cap_fds) {
    size_t cap = g->cap_fds ? 2 * g->cap_fds : 16;
    int *p = (int *)realloc(g->fds, cap * sizeof *p);
    if (!p) return 1;
    g->fds = p;
    g->cap_fds = cap;
  }
  g->fds[g->nfds++] = fd;
  return 0;
}

static void clear_pending(cli_ingest *g)
{
  size_t k;
  for (k = 0; k < g->nfds; k++) close(g->fds[k]);
  g->nfds = 0;
  g->njobs = 0;
  g->bad = 0;
}

/* Destination of channel row 0 of block b of the buffer being filled, and
 * the distance between its rows. */
static unsigned char *row0(cli_ingest *g, unsigned long bi, size_t *stride)
{
  if (g->pipelined) {
    *stride = (size_t)g->Nb * g->cb;
    return g->slots[g->cur].h + (size_t)(bi % g->Nb) * g->cb;
  }
  *stride = g->cb;
  return (unsigned char *)g->ctx->h_blkbufs[bi % g->ctx->Nb_host];
}

static int queue_rows(cli_ingest *g, unsigned long bi, int fd, off_t off)
{
  size_t stride, per = kJobBytes / g->cb;
  unsigned char *dst = row0(g, bi, &stride);
  unsigned int c = 0;
  if (per < 1) per = 1;
  if (per > kMaxRows) per = kMaxRows;
  while (c < g->Nc) {
    job j;
    j.nrows = (unsigned)((g->Nc - c) < per ? (g->Nc - c) : per);
    j.fd = fd;
    j.off = off + (off_t)((size_t)c * g->cb);
    j.dst = dst + (size_t)c * stride;
    j.stride = stride;
    if (push_job(g, j)) return 1;
    c += j.nrows;
  }
  return 0;
}

int cli_ingest_block(cli_ingest *g, unsigned long bi, int fd, off_t off)
{
  int d = dup(fd);
  if (d < 0 || push_fd(g, d)) {
    if (d >= 0) close(d);
    g->bad = 1;
    return 1;
  }
  if (queue_rows(g, bi, d, off)) g->bad = 1;
  g->bytes += (uint64_t)g->Nc * g->cb;
  return g->bad;
}

int cli_ingest_zero_block(cli_ingest *g, unsigned long bi)
{
  if (queue_rows(g, bi, -1, 0)) g->bad = 1;
  return g->bad;
}

void cli_ingest_discard(cli_ingest *g)
{
  clear_pending(g);
}

/* ---- submission ----------------------------------------------------------------- */

static int submit_pipelined(cli_ingest *g)
{
  hslot *s = &g->slots[g->cur];
  const int d = (int)(g->nbuf % kDevBufs);
  hipStream_t ps = (hipStream_t)rawspec_amd_stream(g->ctx);
  if (g->dused[d]) HIPOK(hipStreamWaitEvent(g->copy_stream, g->consumed[d], 0));
  HIPOK(hipMemcpyAsync(g->dbuf[d], s->h, g->buf_bytes, hipMemcpyHostToDevice, g->copy_stream));
  HIPOK(hipEventRecord(s->copied, g->copy_stream));
  s->busy = 1;
  HIPOK(hipStreamWaitEvent(ps, s->copied, 0));
  if (rawspec_amd_start_processing_device(g->ctx, g->dbuf[d], RAWSPEC_FORWARD_FFT)) return 1;
  HIPOK(hipEventRecord(g->consumed[d], ps));
  g->dused[d] = 1;
  g->cur = (g->cur + 1) % g->nslots;
  return 0;
}

int cli_ingest_submit(cli_ingest *g, int *fatal)
{
  int rc;
  *fatal = 0;
  if (g->bad) {
    clear_pending(g);
    fprintf(stderr, "rawspec: out of memory queueing reads\n");
    return 1;
  }
  if (g->pipelined) {
    hslot *s = &g->slots[g->cur];
    if (s->busy) {
      double t0 = now_s();
      if (hipEventSynchronize(s->copied) != hipSuccess) {
        *fatal = 1;
        clear_pending(g);
        return 1;
      }
      s->busy = 0;
      g->read_s += now_s() - t0;
    }
  }
  rc = run_jobs(g);
  clear_pending(g);
  if (rc) return 1;
  if (*(volatile unsigned int *)&g->ctx->exit_soon) {
    *fatal = 1;
    return 1;
  }
  {
    const double t0 = now_s();
    if (g->pipelined) {
      if (submit_pipelined(g)) {
        *fatal = 1;
        return 1;
      }
    } else {
      if (rawspec_wait_for_completion(g->ctx) != 0 ||
          rawspec_copy_blocks_to_gpu_and_start_processing(g->ctx, g->Nb, (char)g->expand4,
                                                          RAWSPEC_FORWARD_FFT)) {
        *fatal = 1;
        return 1;
      }
    }
    g->issue_s += now_s() - t0;
  }
  g->nbuf++;
  return 0;
}

/* ---- lifetime ------------------------------------------------------------------- */

void cli_ingest_destroy(cli_ingest *g)
{
  int k;
  if (!g) return;
  clear_pending(g);
  if (g->threads) {
    pthread_mutex_lock(&g->mu);
    g->quit = 1;
    pthread_cond_broadcast(&g->go);
    pthread_mutex_unlock(&g->mu);
    for (k = 0; k < g->nthreads; k++) pthread_join(g->threads[k], NULL);
    free(g->threads);
  }
  if (g->pipelined) {
    /* the copies and the kernels that read the device buffers end first */
    if (g->copy_stream) hipStreamSynchronize(g->copy_stream);
    rawspec_wait_for_completion(g->ctx);
    for (k = 0; k < g->nslots; k++) {
      if (g->slots[k].h) hipHostFree(g->slots[k].h);
      if (g->slots[k].copied) hipEventDestroy(g->slots[k].copied);
    }
    for (k = 0; k < kDevBufs; k++) {
      if (g->dbuf[k]) hipFree(g->dbuf[k]);
      if (g->consumed[k]) hipEventDestroy(g->consumed[k]);
    }
    if (g->copy_stream) hipStreamDestroy(g->copy_stream);
    free(g->slots);
  }
  pthread_mutex_destroy(&g->mu);
  pthread_cond_destroy(&g->go);
  pthread_cond_destroy(&g->done);
  free(g->jobs);
  free(g->fds);
  free(g);
}

static int setup_pipeline(cli_ingest *g)
{
  int k;
  HIPOK(hipSetDevice(g->ctx->gpu_index));
  HIPOK(hipStreamCreateWithFlags(&g->copy_stream, hipStreamNonBlocking));
  g->slots = (hslot *)calloc((size_t)g->nslots, sizeof *g->slots);
  if (!g->slots) return 1;
  for (k = 0; k < g->nslots; k++) {
    HIPOK(hipHostMalloc((void **)&g->slots[k].h, g->buf_bytes, hipHostMallocDefault));
    HIPOK(hipEventCreateWithFlags(&g->slots[k].copied, hipEventDisableTiming));
  }
  for (k = 0; k < kDevBufs; k++) {
    HIPOK(hipMalloc(&g->dbuf[k], g->buf_bytes));
    HIPOK(hipEventCreateWithFlags(&g->consumed[k], hipEventDisableTiming));
  }
  return 0;
}

cli_ingest *cli_ingest_create(rawspec_context *ctx, int pipelined, int expand4,
                              size_t chan_bytes, int nreaders, int nslots)
{
  cli_ingest *g = (cli_ingest *)calloc(1, sizeof *g);
  int k;
  if (!g) return NULL;
  g->ctx = ctx;
  g->pipelined = pipelined && !expand4;
  g->expand4 = expand4;
  g->cb = chan_bytes;
  g->Nb = ctx->Nb;
  g->Nc = ctx->Nc;
  g->buf_bytes = (size_t)g->Nc * g->Nb * g->cb;
  g->nslots = nslots < 2 ? 2 : nslots;
  pthread_mutex_init(&g->mu, NULL);
  pthread_cond_init(&g->go, NULL);
  pthread_cond_init(&g->done, NULL);
  if (g->pipelined && (rawspec_amd_device_input_size(ctx) != g->buf_bytes || setup_pipeline(g))) {
    fprintf(stderr, "rawspec: could not set up pipelined ingest\n");
    cli_ingest_destroy(g);
    return NULL;
  }
  g->nthreads = nreaders > 1 ? nreaders - 1 : 0;   /* the caller reads too */
  if (g->nthreads) {
    g->threads = (pthread_t *)calloc((size_t)g->nthreads, sizeof(pthread_t));
    if (!g->threads) {
      cli_ingest_destroy(g);
      return NULL;
    }
    for (k = 0; k < g->nthreads; k++)
      if (pthread_create(&g->threads[k], NULL, reader, g)) {
        g->nthreads = k;
        cli_ingest_destroy(g);
        return NULL;
      }
  }
  return g;
}

uint64_t cli_ingest_bytes(const cli_ingest *g) { return g->bytes; }
unsigned long cli_ingest_buffers(const cli_ingest *g) { return g->nbuf; }
double cli_ingest_read_seconds(const cli_ingest *g) { return g->read_s; }
double cli_ingest_issue_seconds(const cli_ingest *g) { return g->issue_s; }
