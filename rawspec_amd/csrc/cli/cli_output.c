/*
 * cli_output.c -- filterbank file output of the rawspec program
 * (the reference's rawspec_file.c:15-242 and rawspec.c:134-162).
 *
 * Names: "<stem>.rawspec.<NNNN>.fil", with the stem's directory replaced by
 * the destination directory when one is given; per-antenna files use the
 * stem "<stem>-antNNN", the incoherent sum "<stem>-ics".  Each file is a
 * filterbank header followed by every dump of the product: for one antenna
 * the whole [dump][pol][chan][fine] power buffer; split per antenna, for
 * each dump and pol the antenna's own [chan][fine] run.
 *
 * The dump callback hands each dump to a writer thread created at POST and
 * joined at the next PRE of the same product, so the spectrometer's stream
 * waits for a write only when the next dump of that product is due.
 */
#define _GNU_SOURCE 1
#include <dlfcn.h>
#include <errno.h>
#include <fcntl.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "cli.h"

static fbh5w_open_fn h5_open;
static fbh5w_write_fn h5_write;
static fbh5w_close_fn h5_close;

int cli_fbh5_load(void)
{
  char exe[PATH_MAX + 1], lib[PATH_MAX + 32];
  ssize_t n = readlink("/proc/self/exe", exe, PATH_MAX);
  char *slash;
  void *h;
  if (n <= 0) return 1;
  exe[n] = '\0';
  slash = strrchr(exe, '/');
  if (slash) *slash = '\0';
  snprintf(lib, sizeof lib, "%s/librawspec_fbh5.so", exe);
  h = dlopen(lib, RTLD_NOW | RTLD_LOCAL);
  if (!h) {
    fprintf(stderr, "error: FBH5 output is not available: %s\n", dlerror());
    return 1;
  }
  h5_open = (fbh5w_open_fn)dlsym(h, "fbh5w_open");
  h5_write = (fbh5w_write_fn)dlsym(h, "fbh5w_write");
  h5_close = (fbh5w_close_fn)dlsym(h, "fbh5w_close");
  if (!h5_open || !h5_write || !h5_close) {
    fprintf(stderr, "error: %s lacks the FBH5 writer\n", lib);
    return 1;
  }
  return 0;
}

static void output_name(char *out, size_t n, const char *dest, const char *stem,
                        const char *suffix)
{
  if (dest && dest[0]) {
    const char *base = strrchr(stem, '/');
    base = base ? base + 1 : stem;
    snprintf(out, n, "%s/%s%s", dest, base, suffix);
  } else {
    snprintf(out, n, "%s%s", stem, suffix);
  }
}

static int open_trunc(const char *name)
{
  int fd = open(name, O_WRONLY | O_CREAT | O_TRUNC, 0664);
  if (fd < 0)
    perror(name);
  else
    posix_fadvise(fd, 0, 0, POSIX_FADV_DONTNEED);
  return fd;
}

int cli_open_output_file(const char *dest, const char *stem, int output_idx)
{
  char name[PATH_MAX + 1], suffix[32];
  snprintf(suffix, sizeof suffix, ".rawspec.%04d.fil", output_idx);
  output_name(name, sizeof name, dest, stem, suffix);
  return open_trunc(name);
}

static fbh5w_ctx *open_h5(cli_product *p, const char *dest, const char *stem, int output_idx)
{
  char name[PATH_MAX + 1], suffix[32];
  snprintf(suffix, sizeof suffix, ".rawspec.%04d.h5", output_idx);
  output_name(name, sizeof name, dest, stem, suffix);
  return h5_open(&p->fb_hdr, p->Nds, name, get_librawspec_version(), get_librawspec_version(),
                 get_cufft_version());
}

int cli_open_ics_file(cli_product *p, const char *dest, const char *ics_stem, int output_idx)
{
  if (p->fbh5) {
    p->h5_ics = open_h5(p, dest, ics_stem, output_idx);
    if (!p->h5_ics) {
      fprintf(stderr, "cannot open output file, giving up\n");
      return 1;
    }
    return 0;
  }
  p->fd_ics = cli_open_output_file(dest, ics_stem, output_idx);
  if (p->fd_ics < 0) {
    fprintf(stderr, "cannot open output file, giving up\n");
    return 1;
  }
  fb_fd_write_header(p->fd_ics, &p->fb_hdr);
  return 0;
}

int cli_open_headers_file(const char *dest, const char *stem)
{
  char name[PATH_MAX + 1];
  output_name(name, sizeof name, dest, stem, ".rawspec.headers");
  return open_trunc(name);
}

int cli_open_product_files(cli_product *p, const char *dest, const char *stem, int output_idx)
{
  char ant_stem[PATH_MAX + 1];
  unsigned int a;
  const unsigned int n = p->per_ant_out ? p->Nant : 1;
  for (a = 0; a < n; a++) {
    if (p->per_ant_out)
      snprintf(ant_stem, sizeof ant_stem, "%s-ant%03u", stem, a);
    else
      snprintf(ant_stem, sizeof ant_stem, "%s", stem);
    if (p->fbh5) {
      p->h5[a] = open_h5(p, dest, ant_stem, output_idx);
      if (!p->h5[a]) {
        fprintf(stderr, "cannot open output file, giving up\n");
        p->exit_soon = 1;
        return 1;
      }
      continue;
    }
    p->fd[a] = cli_open_output_file(dest, ant_stem, output_idx);
    if (p->fd[a] < 0) {
      fprintf(stderr, "cannot open output file, giving up\n");
      p->exit_soon = 1;
      return 1;
    }
    fb_fd_write_header(p->fd[a], &p->fb_hdr);
  }
  return 0;
}

int cli_close_product_files(cli_product *p)
{
  int bad = 0;
  unsigned int a;
  const unsigned int n = p->per_ant_out ? p->Nant : 1;
  for (a = 0; a < n && a < p->nfd; a++) {
    if (p->h5 && p->h5[a]) {
      if (h5_close(p->h5[a])) {
        fprintf(stderr, "FBH5-CLOSE-ERROR ant %u\n", a);
        bad = 1;
      }
      p->h5[a] = NULL;
    }
    if (p->fd[a] >= 0) {
      if (close(p->fd[a]) < 0) {
        fprintf(stderr, "SIGPROC-CLOSE-ERROR ant %u\n", a);
        bad = 1;
      }
      p->fd[a] = -1;
    }
  }
  if (p->h5_ics) {
    if (h5_close(p->h5_ics)) {
      fprintf(stderr, "FBH5-CLOSE-ERROR ics\n");
      bad = 1;
    }
    p->h5_ics = NULL;
  }
  if (p->fd_ics >= 0) {
    if (close(p->fd_ics) < 0) {
      fprintf(stderr, "SIGPROC-CLOSE-ERROR ics\n");
      bad = 1;
    }
    p->fd_ics = -1;
  }
  return bad;
}

static int write_all(int fd, const void *buf, size_t len)
{
  const char *q = (const char *)buf;
  while (len) {
    ssize_t w = write(fd, q, len);
    if (w < 0) {
      if (errno == EINTR) continue;
      return -1;
    }
    q += w;
    len -= (size_t)w;
  }
  return 0;
}

static void write_failed(cli_product *p)
{
  p->exit_soon = 1;
  fprintf(stderr, p->fbh5 ? "FBH5-WRITE-ERROR\n" : "SIGPROC-WRITE-ERROR\n");
}

/* FBH5 dumps: one antenna writes the whole buffer ([Nds][nifs][nchans]);
 * split per antenna, each spectrum's nifs runs of that antenna are gathered
 * into one integration.  (The reference passes each pol's run on its own,
 * which is less than one integration when nifs > 1.) */
static void fbh5_dump(cli_product *p)
{
  if (p->h_pwrbuf && p->Nant == 1 && p->h5 && p->h5[0]) {
    if (h5_write(p->h5[0], p->h_pwrbuf, p->h_pwrbuf_size)) write_failed(p);
  }
  if (p->h_pwrbuf && p->Nant > 1 && p->per_ant_out && p->h5) {
    const size_t spectrum = p->h_pwrbuf_size / ((size_t)p->Nds * sizeof(float));
    const size_t pol = spectrum / (size_t)p->fb_hdr.nifs;
    const size_t ant = pol / p->Nant;
    const size_t need = (size_t)p->fb_hdr.nifs * ant;
    size_t k, j, a;
    if (p->gather_n < need) {
      free(p->gather);
      p->gather = (float *)malloc(sizeof(float) * need);
      p->gather_n = p->gather ? need : 0;
    }
    if (!p->gather) write_failed(p);
    for (k = 0; k < p->Nds && p->gather; k++)
      for (a = 0; a < p->Nant; a++) {
        if (!p->h5[a]) break;
        for (j = 0; j < (size_t)p->fb_hdr.nifs; j++)
          memcpy(p->gather + j * ant, p->h_pwrbuf + a * ant + j * pol + k * spectrum,
                 ant * sizeof(float));
        if (h5_write(p->h5[a], p->gather, (size_t)p->fb_hdr.nifs * ant * sizeof(float)))
          write_failed(p);
      }
  }
  if (p->h5_ics && p->h_icsbuf) {
    if (h5_write(p->h5_ics, p->h_icsbuf, p->h_pwrbuf_size / p->Nant)) write_failed(p);
  }
}

static void *file_writer(void *arg)
{
  cli_product *p = (cli_product *)arg;
  if (p->fbh5) {
    fbh5_dump(p);
    p->total_spectra += p->Nds;
    return NULL;
  }
  if (p->h_pwrbuf && p->Nant == 1 && p->fd[0] >= 0) {
    if (write_all(p->fd[0], p->h_pwrbuf, p->h_pwrbuf_size)) write_failed(p);
  }
  if (p->h_pwrbuf && p->Nant > 1 && p->per_ant_out) {
    const size_t spectrum = p->h_pwrbuf_size / ((size_t)p->Nds * sizeof(float));
    const size_t pol = spectrum / (size_t)p->fb_hdr.nifs;
    const size_t ant = pol / p->Nant;
    size_t k, j, a;
    for (k = 0; k < p->Nds; k++)
      for (j = 0; j < (size_t)p->fb_hdr.nifs; j++)
        for (a = 0; a < p->Nant; a++) {
          if (p->fd[a] < 0) break;
          if (p->debug_callback)
            printf("file writer: dump %zu pol %zu antenna %zu\n", k, j, a);
          if (write_all(p->fd[a], p->h_pwrbuf + a * ant + j * pol + k * spectrum,
                        ant * sizeof(float)))
            write_failed(p);
        }
  }
  if (p->fd_ics >= 0 && p->h_icsbuf) {
    if (write_all(p->fd_ics, p->h_icsbuf, p->h_pwrbuf_size / p->Nant)) write_failed(p);
  }
  p->total_spectra += p->Nds;
  return NULL;
}

static void *net_writer(void *arg)
{
  cli_send_dump((cli_product *)arg);
  return NULL;
}

static void dump_callback(rawspec_context *ctx, int i, int type, void *(*writer)(void *))
{
  cli_product *p = &((cli_product *)ctx->user_data)[i];
  int rc;
  ctx->exit_soon = p->exit_soon;
  if (type == RAWSPEC_CALLBACK_PRE_DUMP) {
    if (p->output_thread_valid) {
      if ((rc = pthread_join(p->output_thread, NULL)))
        fprintf(stderr, "pthread_join: %s\n", strerror(rc));
      p->output_thread_valid = 0;
    }
  } else if (type == RAWSPEC_CALLBACK_POST_DUMP) {
    if ((rc = pthread_create(&p->output_thread, NULL, writer, p)))
      fprintf(stderr, "pthread_create: %s\n", strerror(rc));
    else
      p->output_thread_valid = 1;
  }
}

void cli_dump_file_callback(rawspec_context *ctx, int i, int type)
{
  dump_callback(ctx, i, type, file_writer);
}

void cli_dump_net_callback(rawspec_context *ctx, int i, int type)
{
  dump_callback(ctx, i, type, net_writer);
}
