/*
 * Drives the rawspec program's ingest (rawspec_amd/csrc/cli/cli_ingest.c)
 * on the CPU against stand-ins for HIP and for the library's processing
 * calls, which record a checksum of every buffer they are given, in order.
 *
 *   harness FILE NC NB CB NBLOCKS MODE READERS SLOTS DELAY_US ZERO...
 *
 * Block k of the stem is block k of FILE (NC rows of CB bytes each) unless
 * k is listed among the ZERO block numbers (then it is a PKTIDX gap and the
 * file's blocks shift by one).  MODE 1 = pipelined, 0 = synchronous.
 * Prints one line per processed buffer: "buf <n> <checksum>", then
 * "end <rc>" where rc is the first non-zero submit/flush status.
 */
#define _GNU_SOURCE 1
#include <fcntl.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "cli.h"
#include "hip/hip_runtime_api.h"
#include "rawspec_amd.h"

static unsigned g_delay_us;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static unsigned long g_nbuf;
static size_t g_buf_bytes;
static int g_in_library;          /* calls into the "library" must never overlap */

/* ---- HIP stand-ins ---- */
hipError_t hipSetDevice(int dev) { (void)dev; return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned f) { (void)f; *s = (void *)1; return 0; }
hipError_t hipStreamDestroy(hipStream_t s) { (void)s; return 0; }
hipError_t hipStreamSynchronize(hipStream_t s) { (void)s; return 0; }
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f) { (void)s; (void)e; (void)f; return 0; }
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned f) { (void)f; *e = (void *)1; return 0; }
hipError_t hipEventDestroy(hipEvent_t e) { (void)e; return 0; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) { (void)e; (void)s; return 0; }
hipError_t hipEventSynchronize(hipEvent_t e) { (void)e; return 0; }
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, int k, hipStream_t st)
{
  (void)k; (void)st;
  if (g_delay_us) usleep(g_delay_us);
  memcpy(d, s, n);
  return 0;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned f) { (void)f; *p = malloc(n); return *p ? 0 : 2; }
hipError_t hipHostFree(void *p) { free(p); return 0; }
hipError_t hipMalloc(void **p, size_t n) { *p = malloc(n); return *p ? 0 : 2; }
hipError_t hipFree(void *p) { free(p); return 0; }
const char *hipGetErrorString(hipError_t e) { return e ? "error" : "success"; }

/* ---- library stand-ins ---- */
static rawspec_context *g_ctx;

static uint64_t checksum(const unsigned char *p, size_t n)
{
  uint64_t s = 0;
  size_t i;
  for (i = 0; i < n; i++) s += (uint64_t)p[i] * (uint64_t)(i % 65521 + 1);
  return s;
}

static void record(const unsigned char *dev)
{
  pthread_mutex_lock(&g_mu);
  if (g_in_library) {
    fprintf(stderr, "overlapping library calls\n");
    abort();
  }
  g_in_library = 1;
  pthread_mutex_unlock(&g_mu);
  printf("buf %lu %llu\n", g_nbuf++, (unsigned long long)checksum(dev, g_buf_bytes));
  fflush(stdout);
  if (g_delay_us) usleep(g_delay_us / 2);
  pthread_mutex_lock(&g_mu);
  g_in_library = 0;
  pthread_mutex_unlock(&g_mu);
}

void *rawspec_amd_stream(rawspec_context *ctx) { (void)ctx; return (void *)1; }
size_t rawspec_amd_device_input_size(rawspec_context *ctx) { (void)ctx; return g_buf_bytes; }
int rawspec_amd_start_processing_device(rawspec_context *ctx, const void *d, int dir)
{
  (void)ctx; (void)dir;
  record((const unsigned char *)d);
  return 0;
}
int rawspec_wait_for_completion(rawspec_context *ctx) { (void)ctx; return 0; }
int rawspec_copy_blocks_to_gpu_and_start_processing(rawspec_context *ctx, size_t nb, char e4,
                                                    int dir)
{
  /* assemble the device layout [chan][Nb][row] from the host blocks */
  unsigned char *dev = (unsigned char *)malloc(g_buf_bytes);
  const size_t cb = g_buf_bytes / ((size_t)ctx->Nc * ctx->Nb);
  size_t c, b;
  (void)e4; (void)dir;
  for (c = 0; c < ctx->Nc; c++)
    for (b = 0; b < nb; b++)
      memcpy(dev + (c * ctx->Nb + b) * cb, ctx->h_blkbufs[b] + c * cb, cb);
  record(dev);
  free(dev);
  return 0;
}

int main(int argc, char **argv)
{
  rawspec_context ctx;
  cli_ingest *g;
  int fd, mode, readers, slots, rc = 0, fatal = 0, k;
  unsigned nc, nb, nblocks, fblk = 0;
  size_t cb;
  unsigned long bi;
  unsigned char zero[4096] = {0};
  if (argc < 10) return 2;
  fd = open(argv[1], O_RDONLY);
  nc = (unsigned)atoi(argv[2]);
  nb = (unsigned)atoi(argv[3]);
  cb = (size_t)atol(argv[4]);
  nblocks = (unsigned)atoi(argv[5]);
  mode = atoi(argv[6]);
  readers = atoi(argv[7]);
  slots = atoi(argv[8]);
  g_delay_us = (unsigned)atoi(argv[9]);
  for (k = 10; k < argc && k - 10 < 4096; k++) zero[atoi(argv[k])] = 1;
  if (fd < 0) return 2;
  memset(&ctx, 0, sizeof ctx);
  ctx.Nc = nc;
  ctx.Nb = nb;
  ctx.Nb_host = nb;
  ctx.h_blkbufs = (char **)calloc(nb, sizeof(char *));
  for (k = 0; k < (int)nb; k++) ctx.h_blkbufs[k] = (char *)malloc((size_t)nc * cb);
  g_ctx = &ctx;
  g_buf_bytes = (size_t)nc * nb * cb;
  g = cli_ingest_create(&ctx, mode, 0, cb, readers, slots);
  if (!g) return 3;
  for (bi = 0; bi < nblocks; bi++) {
    if (zero[bi])
      cli_ingest_zero_block(g, bi);
    else
      cli_ingest_block(g, bi, fd, (off_t)fblk++ * (off_t)(nc * cb));
    if (bi % nb == nb - 1 && cli_ingest_submit(g, &fatal)) {
      rc = fatal ? 2 : 1;
      break;
    }
  }
  close(fd);   /* the ingest holds its own descriptors */
  if (cli_ingest_flush(g, &fatal) && !rc) rc = fatal ? 2 : 1;
  cli_ingest_discard(g);
  printf("end %d buffers %lu bytes %llu\n", rc, cli_ingest_buffers(g),
         (unsigned long long)cli_ingest_bytes(g));
  cli_ingest_destroy(g);
  for (k = 0; k < (int)nb; k++) free(ctx.h_blkbufs[k]);
  free(ctx.h_blkbufs);
  return 0;
}
