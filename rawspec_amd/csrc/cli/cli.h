/*
 * cli.h -- internals of the rawspec command-line program built on
 * librawspec: GUPPI RAW stems in, SIGPROC filterbank files (or UDP
 * packets) out.  The program follows the reference CLI (rawspec.c,
 * rawspec_file.c, rawspec_socket.c) in options, output names and formats,
 * and in how blocks of a stem feed the spectrometer; its ingest is its own
 * (parallel reads into pinned staging buffers, host-to-device copies on a
 * stream of their own, overlapped with the spectrometer's kernels).
 */
#ifndef RAWSPEC_CLI_H_
#define RAWSPEC_CLI_H_

#include <pthread.h>
#include <stdint.h>
#include <sys/types.h>

#include "fbh5_writer.h"
#include "rawspec.h"
#include "rawspec_fbutils.h"

/* Per output product state shared between the main loop and the dump
 * callback (the reference keeps the same facts in callback_data_t,
 * rawspec_callback.h:24-57). */
typedef struct {
  int *fd;                 /* output files, one per antenna when split (else [0]) */
  int fd_ics;              /* incoherent-sum output file, or -1 */
  unsigned int Nant;       /* antennas the output is split over */
  int per_ant_out;         /* 1: one file per antenna */
  unsigned int nfd;        /* entries of fd */
  unsigned int total_spectra;
  unsigned int total_packets;
  uint64_t total_bytes;
  uint64_t total_ns;
  double rate;             /* network output rate (Gbps) */
  int debug_callback;
  int output_thread_valid;
  pthread_t output_thread;
  float *h_pwrbuf;         /* copies of the context's fields */
  size_t h_pwrbuf_size;
  float *h_icsbuf;
  unsigned int Nds;
  unsigned int Nf;         /* fine channels (Nc * Nt) */
  fb_hdr_t fb_hdr;
  int net;                 /* 1: send UDP packets to fd[0] instead of writing files */
  int fbh5;                /* 1: FBH5 files (h5, h5_ics) instead of SIGPROC ones */
  fbh5w_ctx **h5;          /* one per entry of fd */
  fbh5w_ctx *h5_ics;
  float *gather;           /* one integration of one antenna (FBH5, split output) */
  size_t gather_n;         /* floats gather holds */
  volatile unsigned int exit_soon;
} cli_product;

/* ---- output (cli_output.c) ------------------------------------------------ */

/* Opens "<dest>/<basename(stem)>.rawspec.<idx>.fil" (or "<stem>.rawspec.<idx>.fil"
 * when dest is empty); -1 on failure (rawspec_file.c:15-68). */
int cli_open_output_file(const char *dest, const char *stem, int output_idx);

/* Opens the product's file(s), one per antenna ("<stem>-antNNN") when split,
 * and writes the filterbank header to each (rawspec_file.c:73-98); with
 * p->fbh5, FBH5 files "....h5" instead. */
int cli_open_product_files(cli_product *p, const char *dest, const char *stem, int output_idx);
/* The incoherent-sum file of a product (stem "<stem>-ics"). */
int cli_open_ics_file(cli_product *p, const char *dest, const char *ics_stem, int output_idx);

/* Loads the FBH5 plugin (librawspec_fbh5.so next to the program); 0 on success. */
int cli_fbh5_load(void);

/* Closes the product's files; returns non-zero if a close failed. */
int cli_close_product_files(cli_product *p);

/* Opens "<dest>/<basename(stem)>.rawspec.headers" (rawspec.c:134-162). */
int cli_open_headers_file(const char *dest, const char *stem);

/* Dump callbacks for file and network output (rawspec_file.c:203-242). */
void cli_dump_file_callback(rawspec_context *ctx, int output_product, int callback_type);
void cli_dump_net_callback(rawspec_context *ctx, int output_product, int callback_type);

/* ---- network output (cli_net.c) -------------------------------------------- */

/* Connected UDP socket to host:port, or -1 (rawspec_socket.c:21-94). */
int cli_open_output_socket(const char *host, const char *port);

/* Sends one dump of product p as filterbank packets (rawspec_socket.c:134-325). */
void cli_send_dump(cli_product *p);

/* ---- ingest (cli_ingest.c) ------------------------------------------------- */

typedef struct cli_ingest cli_ingest;

/* Ingest for an initialised context.  pipelined != 0 stages whole buffers
 * through nslots pinned host slots and two device buffers on a copy stream
 * of its own and starts processing with rawspec_amd_start_processing_device;
 * 0 reads into ctx->h_blkbufs and goes through rawspec_wait_for_completion /
 * rawspec_copy_blocks_to_gpu_and_start_processing, as the reference CLI does
 * (required for 4-bit input, which the library expands on upload).
 * chan_bytes: file bytes of one channel of one block; nreaders: read threads. */
cli_ingest *cli_ingest_create(rawspec_context *ctx, int pipelined, int expand4,
                              size_t chan_bytes, int nreaders, int nslots);
void cli_ingest_destroy(cli_ingest *g);

/* Block bi of the stem (bi counts zero-filled blocks too) comes from fd:
 * ctx->Nc channel rows starting at file offset off.  The read is deferred
 * until the block's buffer is complete; fd is duplicated, so the caller may
 * close it at once. */
int cli_ingest_block(cli_ingest *g, unsigned long bi, int fd, off_t off);
/* Block bi is zeros (a gap in PKTIDX). */
int cli_ingest_zero_block(cli_ingest *g, unsigned long bi);
/* Buffer complete (bi % Nb == Nb - 1): read it and start processing (in
 * pipelined mode: start reading it, and start the previous buffer).
 * Returns non-zero on a read error or when processing could not start
 * (*fatal set for the latter). */
int cli_ingest_submit(cli_ingest *g, int *fatal);
/* Starts the last complete buffer still held back (pipelined mode starts a
 * buffer when the next one is complete); call at the end of a stem.
 * Returns as cli_ingest_submit. */
int cli_ingest_flush(cli_ingest *g, int *fatal);
/* Drops the blocks of an incomplete buffer (end of a stem). */
void cli_ingest_discard(cli_ingest *g);
/* Bytes read from files and buffers started so far. */
uint64_t cli_ingest_bytes(const cli_ingest *g);
unsigned long cli_ingest_buffers(const cli_ingest *g);
/* Seconds spent waiting for reads, and (pipelined) for staging slots. */
double cli_ingest_read_seconds(const cli_ingest *g);
/* Seconds spent starting copies and processing (sync mode: waiting for the
 * GPU and copying). */
double cli_ingest_issue_seconds(const cli_ingest *g);

#endif
