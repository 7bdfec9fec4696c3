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
 *    [chan][Nb*Ntpb][pol][re,im] of one of several pinned host slots.  A
 *    slot whose reads are done goes to an issuing thread, which copies it
 *    to one of two device buffers with one hipMemcpyAsync on a copy stream
 *    and starts its processing on the library's stream once the copy has
 *    landed (stream-ordered by events).  The HIP runtime may hold that
 *    thread while the copy engine is busy; the pool meanwhile reads the
 *    next buffers into the other slots.  Reads, the PCIe copies and the
 *    kernels of consecutive buffers thus overlap, and the stem's rate is
 *    that of the slowest of the three.  A buffer is handed over when the
 *    next one is complete or when the stem ends (flush); only the issuing
 *    thread calls the library until the flush has drained it.
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

/* The reads of one buffer, and the file descriptors they use. */
typedef struct {
  job *jobs;
  size_t njobs, cap_jobs;
  int *fds;
  size_t nfds, cap_fds;
  int bad;               /* a job could not be queued */
} batch;

enum { kSlotFree = 0, kSlotQueued = 1, kSlotIssued = 2 };

typedef struct {
  unsigned char *h;      /* pinned, device layout */
  hipEvent_t copied;     /* recorded after its copy (kSlotIssued) */
  int state;             /* kSlot*: guarded by mu */
} hslot;

struct cli_ingest {
  rawspec_context *ctx;
  int pipelined;
  int expand4;
  size_t cb;             /* file bytes per channel row of one block */
  size_t buf_bytes;      /* Nc * Nb * cb */
  unsigned int Nb, Nc;

  batch q[2];            /* q[fill]: the buffer being queued; q[fill ^ 1]: being read */
  int fill;

  /* read pool */
  int nthreads;
  pthread_t *threads;
  pthread_mutex_t mu;
  pthread_cond_t go, done;
  unsigned long gen;
  int idle;              /* readers done with the current generation */
  const batch *work;     /* the generation's jobs */
  size_t next;
  int read_error;
  int quit;
  int reading;           /* a generation is in flight (pipelined mode) */
  int rd_slot;           /* the slot it fills; -1 when none */

  /* pipelined staging */
  hslot *slots;
  int nslots, cur;       /* cur: slot of the buffer being queued */
  pthread_t issuer;
  int have_issuer;
  pthread_cond_t queued, issued;
  int *iq;               /* slots waiting to be issued, FIFO of nslots entries */
  int iq_head, iq_len;
  int issuing;           /* the issuer is working on a slot */
  int fatal;             /* the issuer could not start a buffer */
  void *dbuf[kDevBufs];
  hipEvent_t consumed[kDevBufs];
  int dused[kDevBufs];
  hipStream_t copy_stream;
  unsigned long nbuf;
  uint64_t bytes;
  double read_s;         /* waiting for reads and free slots */
  double issue_s;        /* starting copies and processing */
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

static int run_job(const cli_ingest *g, const job *j)
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

/* Takes jobs of the current generation until none is left. */
static void drain(cli_ingest *g)
{
  const batch *b = g->work;
  for (;;) {
    size_t k = __atomic_fetch_add(&g->next, 1, __ATOMIC_RELAXED);
    if (k >= b->njobs) break;
    if (run_job(g, &b->jobs[k])) __atomic_store_n(&g->read_error, 1, __ATOMIC_RELAXED);
  }
}

/* A reader drains each generation of jobs, then reports itself idle; the
 * next generation starts only once every reader is idle, so no reader ever
 * sees a job list while it is being rebuilt. */
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

/* Starts a generation over b's jobs (every reader is idle). */
static void start_reads(cli_ingest *g, const batch *b)
{
  g->work = b;
  g->next = 0;
  g->read_error = 0;
  pthread_mutex_lock(&g->mu);
  g->gen++;
  g->idle = 0;
  pthread_cond_broadcast(&g->go);
  pthread_mutex_unlock(&g->mu);
}

/* Waits until every reader is idle, i.e. every job of the generation is done. */
static int wait_reads(cli_ingest *g)
{
  pthread_mutex_lock(&g->mu);
  while (g->idle < g->nthreads) pthread_cond_wait(&g->done, &g->mu);
  pthread_mutex_unlock(&g->mu);
  return g->read_error;
}

/* ---- queueing --------------------------------------------------------------- */

static int push_job(batch *b, job j)
{
  if (b->njobs == b->cap_jobs) {
    size_t cap = b->cap_jobs ? 2 * b->cap_jobs : 256;
    job *p = (job *)realloc(b->jobs, cap * sizeof *p);
    if (!p) return 1;
    b->jobs = p;
    b->cap_jobs = cap;
  }
  b->jobs[b->njobs++] = j;
  return 0;
}

static int push_fd(batch *b, int fd)
{
  if (b->nfds == b->cap_fds) {
    size_t cap = b->cap_fds ? 2 * b->cap_fds : 16;
    int *p = (int *)realloc(b->fds, cap * sizeof *p);
    if (!p) return 1;
    b->fds = p;
    b->cap_fds = cap;
  }
  b->fds[b->nfds++] = fd;
  return 0;
}

static void clear_batch(batch *b)
{
  size_t k;
  for (k = 0; k < b->nfds; k++) close(b->fds[k]);
  b->nfds = 0;
  b->njobs = 0;
  b->bad = 0;
}

/* Destination of channel row 0 of block bi of the buffer being queued, and
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
  batch *b = &g->q[g->fill];
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
    if (push_job(b, j)) return 1;
    c += j.nrows;
  }
  return 0;
}

int cli_ingest_block(cli_ingest *g, unsigned long bi, int fd, off_t off)
{
  batch *b = &g->q[g->fill];
  int d = dup(fd);
  if (d < 0 || push_fd(b, d)) {
    if (d >= 0) close(d);
    b->bad = 1;
    return 1;
  }
  if (queue_rows(g, bi, d, off)) b->bad = 1;
  g->bytes += (uint64_t)g->Nc * g->cb;
  return b->bad;
}

int cli_ingest_zero_block(cli_ingest *g, unsigned long bi)
{
  batch *b = &g->q[g->fill];
  if (queue_rows(g, bi, -1, 0)) b->bad = 1;
  return b->bad;
}

void cli_ingest_discard(cli_ingest *g)
{
  clear_batch(&g->q[g->fill]);
}

/* ---- submission ----------------------------------------------------------------- */

/* Copies slot `slot` to the next device buffer and starts its processing
 * (issuing thread only). */
static int issue(cli_ingest *g, int slot)
{
  hslot *s = &g->slots[slot];
  const int d = (int)(g->nbuf % kDevBufs);
  hipStream_t ps = (hipStream_t)rawspec_amd_stream(g->ctx);
  const double t0 = now_s();
  if (g->dused[d]) HIPOK(hipStreamWaitEvent(g->copy_stream, g->consumed[d], 0));
  HIPOK(hipMemcpyAsync(g->dbuf[d], s->h, g->buf_bytes, hipMemcpyHostToDevice, g->copy_stream));
  HIPOK(hipEventRecord(s->copied, g->copy_stream));
  HIPOK(hipStreamWaitEvent(ps, s->copied, 0));
  if (rawspec_amd_start_processing_device(g->ctx, g->dbuf[d], RAWSPEC_FORWARD_FFT)) return 1;
  HIPOK(hipEventRecord(g->consumed[d], ps));
  g->dused[d] = 1;
  g->nbuf++;
  g->issue_s += now_s() - t0;
  return 0;
}

static int stopping(const cli_ingest *g)
{
  return *(volatile unsigned int *)&g->ctx->exit_soon != 0;
}

static void *issuer(void *arg)
{
  cli_ingest *g = (cli_ingest *)arg;
  hipSetDevice(g->ctx->gpu_index);
  for (;;) {
    int slot, rc;
    pthread_mutex_lock(&g->mu);
    while (g->iq_len == 0 && !g->quit) pthread_cond_wait(&g->queued, &g->mu);
    if (g->quit) {
      pthread_mutex_unlock(&g->mu);
      return NULL;
    }
    slot = g->iq[g->iq_head];
    g->iq_head = (g->iq_head + 1) % g->nslots;
    g->iq_len--;
    g->issuing = 1;
    rc = g->fatal;
    pthread_mutex_unlock(&g->mu);
    if (!rc) rc = stopping(g) || issue(g, slot);
    pthread_mutex_lock(&g->mu);
    if (rc) g->fatal = 1;
    g->slots[slot].state = rc ? kSlotFree : kSlotIssued;
    g->issuing = 0;
    pthread_cond_broadcast(&g->issued);
    pthread_mutex_unlock(&g->mu);
  }
}

/* Hands a slot whose reads are done to the issuing thread. */
static void hand_over(cli_ingest *g, int slot)
{
  pthread_mutex_lock(&g->mu);
  g->slots[slot].state = kSlotQueued;
  g->iq[(g->iq_head + g->iq_len) % g->nslots] = slot;
  g->iq_len++;
  pthread_cond_signal(&g->queued);
  pthread_mutex_unlock(&g->mu);
}

/* Waits until slot `slot` may be overwritten; non-zero if the issuer failed. */
static int slot_free(cli_ingest *g, int slot)
{
  hslot *s = &g->slots[slot];
  int st, bad;
  pthread_mutex_lock(&g->mu);
  while (s->state == kSlotQueued && !g->fatal) pthread_cond_wait(&g->issued, &g->mu);
  st = s->state;
  bad = g->fatal;
  pthread_mutex_unlock(&g->mu);
  if (bad) return 1;
  if (st == kSlotIssued) {
    if (hipEventSynchronize(s->copied) != hipSuccess) return 1;
    pthread_mutex_lock(&g->mu);
    s->state = kSlotFree;
    pthread_mutex_unlock(&g->mu);
  }
  return 0;
}

/* Waits until the issuing thread has started every buffer handed to it. */
static int issuer_idle(cli_ingest *g)
{
  int bad;
  pthread_mutex_lock(&g->mu);
  while ((g->iq_len > 0 || g->issuing) && !g->fatal) pthread_cond_wait(&g->issued, &g->mu);
  while (g->issuing) pthread_cond_wait(&g->issued, &g->mu);
  bad = g->fatal;
  pthread_mutex_unlock(&g->mu);
  return bad;
}

static int submit_pipelined(cli_ingest *g, int *fatal)
{
  const double t0 = now_s();
  int rc = 0;
  /* one generation of reads at a time: the previous buffer's end first */
  if (g->reading) {
    rc = wait_reads(g);
    g->reading = 0;
  }
  clear_batch(&g->q[g->fill ^ 1]);     /* the previous buffer's files */
  if (rc) {
    clear_batch(&g->q[g->fill]);
    g->rd_slot = -1;
    g->read_s += now_s() - t0;
    return 1;
  }
  if (g->rd_slot >= 0) hand_over(g, g->rd_slot);
  g->rd_slot = -1;
  /* this buffer's slot is free once its last copy has landed */
  if (slot_free(g, g->cur)) {
    *fatal = 1;
    return 1;
  }
  g->read_s += now_s() - t0;
  start_reads(g, &g->q[g->fill]);
  g->reading = 1;
  g->rd_slot = g->cur;
  g->fill ^= 1;
  g->cur = (g->cur + 1) % g->nslots;
  return 0;
}

int cli_ingest_submit(cli_ingest *g, int *fatal)
{
  batch *b = &g->q[g->fill];
  double t0;
  int rc;
  *fatal = 0;
  if (b->bad) {
    clear_batch(b);
    fprintf(stderr, "rawspec: out of memory queueing reads\n");
    return 1;
  }
  if (g->pipelined) return submit_pipelined(g, fatal);

  /* synchronous: read (pool plus caller), then the reference's wait, copy, start */
  t0 = now_s();
  start_reads(g, b);
  drain(g);
  rc = wait_reads(g);
  g->read_s += now_s() - t0;
  clear_batch(b);
  if (rc) return 1;
  if (stopping(g)) {
    *fatal = 1;
    return 1;
  }
  t0 = now_s();
  if (rawspec_wait_for_completion(g->ctx) != 0 ||
      rawspec_copy_blocks_to_gpu_and_start_processing(g->ctx, g->Nb, (char)g->expand4,
                                                      RAWSPEC_FORWARD_FFT)) {
    *fatal = 1;
    return 1;
  }
  g->issue_s += now_s() - t0;
  g->nbuf++;
  return 0;
}

int cli_ingest_flush(cli_ingest *g, int *fatal)
{
  double t0;
  int rc = 0;
  *fatal = 0;
  if (!g->pipelined) return 0;
  t0 = now_s();
  if (g->reading) {
    rc = wait_reads(g);
    g->reading = 0;
    clear_batch(&g->q[g->fill ^ 1]);
    if (!rc) hand_over(g, g->rd_slot);
  }
  g->rd_slot = -1;
  if (issuer_idle(g)) {
    *fatal = 1;
    rc = 1;
  }
  g->read_s += now_s() - t0;
  return rc;
}

/* ---- lifetime ------------------------------------------------------------------- */

void cli_ingest_destroy(cli_ingest *g)
{
  int k;
  if (!g) return;
  if (g->threads && g->reading) wait_reads(g);
  pthread_mutex_lock(&g->mu);
  g->quit = 1;
  pthread_cond_broadcast(&g->go);
  pthread_cond_broadcast(&g->queued);
  pthread_mutex_unlock(&g->mu);
  if (g->threads) {
    for (k = 0; k < g->nthreads; k++) pthread_join(g->threads[k], NULL);
    free(g->threads);
  }
  if (g->have_issuer) pthread_join(g->issuer, NULL);
  clear_batch(&g->q[0]);
  clear_batch(&g->q[1]);
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
    free(g->iq);
  }
  pthread_mutex_destroy(&g->mu);
  pthread_cond_destroy(&g->go);
  pthread_cond_destroy(&g->done);
  pthread_cond_destroy(&g->queued);
  pthread_cond_destroy(&g->issued);
  for (k = 0; k < 2; k++) {
    free(g->q[k].jobs);
    free(g->q[k].fds);
  }
  free(g);
}

static int setup_pipeline(cli_ingest *g)
{
  int k;
  HIPOK(hipSetDevice(g->ctx->gpu_index));
  HIPOK(hipStreamCreateWithFlags(&g->copy_stream, hipStreamNonBlocking));
  g->slots = (hslot *)calloc((size_t)g->nslots, sizeof *g->slots);
  g->iq = (int *)calloc((size_t)g->nslots, sizeof *g->iq);
  if (!g->slots || !g->iq) return 1;
  for (k = 0; k < g->nslots; k++) {
    /* write-combined: the reads stream into it without pulling its lines
     * into the caches, and the copy engine reads it without snooping */
    HIPOK(hipHostMalloc((void **)&g->slots[k].h, g->buf_bytes, hipHostMallocWriteCombined));
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
  g->rd_slot = -1;
  pthread_mutex_init(&g->mu, NULL);
  pthread_cond_init(&g->go, NULL);
  pthread_cond_init(&g->done, NULL);
  pthread_cond_init(&g->queued, NULL);
  pthread_cond_init(&g->issued, NULL);
  if (g->pipelined && (rawspec_amd_device_input_size(ctx) != g->buf_bytes || setup_pipeline(g))) {
    fprintf(stderr, "rawspec: could not set up pipelined ingest\n");
    cli_ingest_destroy(g);
    return NULL;
  }
  if (g->pipelined) {
    if (pthread_create(&g->issuer, NULL, issuer, g)) {
      cli_ingest_destroy(g);
      return NULL;
    }
    g->have_issuer = 1;
  }
  /* synchronous mode reads with the caller too; pipelined mode leaves the
   * caller free to issue copies while the pool reads */
  if (nreaders < 1) nreaders = 1;
  g->nthreads = g->pipelined ? nreaders : nreaders - 1;
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
