/*
 * cli_main.c -- the rawspec program: GUPPI RAW stems to SIGPROC filterbank
 * files (or UDP packets) through librawspec on an MI355X.
 *
 * Options, defaults, output names and formats, and the way the blocks of a
 * stem are sequenced follow the reference CLI (rawspec.c:73-1128):
 *
 *  - a stem is read as the files STEM.0000.raw, STEM.0001.raw, ... until one
 *    is missing; block geometry comes from the first header of the stem;
 *  - PKTIDX steps are learned from the first increase; a block whose PKTIDX
 *    skips k steps is preceded by k-1 zero blocks; a backward or misaligned
 *    step ends the stem; a repeated PKTIDX ends the file;
 *  - buffers of Nb blocks go to the spectrometer; a partial buffer at the
 *    end of a stem is dropped, and the next stem restarts the integrations;
 *  - one filterbank file per output product (per antenna with -S or when the
 *    header says NANTS > 1, plus an incoherent-sum file with -i).
 *
 * Deviations, all in cases where the reference misbehaves: -i weights are
 * parsed one per comma-separated field (the reference reads only the
 * first); channel offsets are computed in 64 bits; zero blocks fill only the
 * selected channels; the headers file is closed at the end of each stem.
 */
#define _GNU_SOURCE 1
#include <errno.h>
#include <fcntl.h>
#include <getopt.h>
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/sendfile.h>
#include <sys/stat.h>
#include <time.h>
#include <unistd.h>

#include "cli.h"
#include "rawspec_amd.h"
#include "rawspec_rawutils.h"

static struct option long_opts[] = {
  {"ant", 1, NULL, 'a'},     {"batch", 1, NULL, 'b'},   {"dest", 1, NULL, 'd'},
  {"ffts", 1, NULL, 'f'},    {"gpu", 1, NULL, 'g'},     {"help", 0, NULL, 'h'},
  {"hdrs", 0, NULL, 'H'},    {"ics", 1, NULL, 'i'},     {"fbh5", 0, NULL, 'j'},
  {"nchan", 1, NULL, 'n'},   {"outidx", 1, NULL, 'o'},  {"pols", 1, NULL, 'p'},
  {"rate", 1, NULL, 'r'},    {"schan", 1, NULL, 's'},   {"splitant", 0, NULL, 'S'},
  {"ints", 1, NULL, 't'},    {"version", 0, NULL, 'v'}, {"debug", 0, NULL, 'z'},
  /* this program's own: ingest tuning and reporting */
  {"readers", 1, NULL, 1000}, {"sync-ingest", 0, NULL, 1001}, {"slots", 1, NULL, 1002},
  {"mfma", 0, NULL, 1003},    {"stats", 0, NULL, 1004},
  {0, 0, 0, 0}};

static void usage(const char *argv0)
{
  const char *b = strrchr(argv0, '/');
  b = b ? b + 1 : argv0;
  fprintf(stderr,
          "Usage: %s [options] STEM [...]\n"
          "\n"
          "Options:\n"
          "  -a, --ant=ANT          The 0-indexed antenna to exclusively process [-1]\n"
          "  -b, --batch=BC         Batch process BC coarse-channels at a time (1: auto, <1: disabled) [0]\n"
          "  -d, --dest=DEST        Destination directory or host:port\n"
          "  -f, --ffts=N1[,N2...]  FFT lengths [1048576, 8, 1024]\n"
          "  -g, --GPU=IDX          Select GPU device to use [0]\n"
          "  -H, --hdrs             Save headers to separate file\n"
          "  -i, --ics=W1[,W2...]   Output incoherent-sum (exclusively, unless with -S)\n"
          "                         specifying per antenna-weights or a singular, uniform weight\n"
          "  -j, --fbh5             Format output as FBH5 (.h5) instead of SIGPROC (.fil)\n"
          "  -n, --nchan=N          Number of coarse channels to process [all]\n"
          "  -o, --outidx=N         First index number for output files [0]\n"
          "  -p  --pols={1|4}[,...] Number of output polarizations [1]\n"
          "                         1=total power, 4=cross pols, -4=full stokes\n"
          "  -r, --rate=GBPS        Desired net data rate in Gbps [6.0]\n"
          "  -s, --schan=C          First coarse channel to process [0]\n"
          "  -S, --splitant         Split output into per antenna files\n"
          "  -t, --ints=N1[,N2...]  Spectra to integrate [51, 128, 3072]\n"
          "  -z, --debug            Turn on selected debug output\n"
          "\n"
          "      --readers=N        Threads reading blocks [12]\n"
          "      --slots=N          Pinned staging buffers for pipelined ingest [2]\n"
          "      --sync-ingest      Read, copy and start each buffer in turn (the library's\n"
          "                         standard copy API) instead of the pipelined ingest\n"
          "      --mfma             Use the matrix-core transform engine at its 1e-3\n"
          "                         relative error bar (engine level 1)\n"
          "      --stats            Print ingest statistics as one JSON line on stderr\n"
          "\n"
          "  -h, --help             Show this message\n"
          "  -v, --version          Show version and exit\n\n",
          b);
}

/* Comma-separated unsigned list into a[0..MAX_OUTPUTS); -1 if too many. */
static int parse_list_u(char *arg, unsigned int *a, const char *what)
{
  int n = 0;
  char *save = NULL, *tok;
  for (tok = strtok_r(arg, ",", &save); tok; tok = strtok_r(NULL, ",", &save)) {
    if (n >= MAX_OUTPUTS) {
      fprintf(stderr, "error: up to %d %s supported.\n", MAX_OUTPUTS, what);
      return -1;
    }
    a[n++] = (unsigned int)strtoul(tok, NULL, 0);
  }
  return n;
}

static int parse_list_i(char *arg, int *a, const char *what)
{
  int n = 0;
  char *save = NULL, *tok;
  for (tok = strtok_r(arg, ",", &save); tok; tok = strtok_r(NULL, ",", &save)) {
    if (n >= MAX_OUTPUTS) {
      fprintf(stderr, "error: up to %d %s supported.\n", MAX_OUTPUTS, what);
      return -1;
    }
    a[n++] = (int)strtoul(tok, NULL, 0);
  }
  return n;
}

static double now_s(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* (Re)allocates a product's file table for n antennas, all closed. */
static int reset_fds(cli_product *p, unsigned int n)
{
  unsigned int a;
  int *fd = (int *)malloc(sizeof(int) * (n ? n : 1));
  fbh5w_ctx **h5 = (fbh5w_ctx **)calloc(n ? n : 1, sizeof(fbh5w_ctx *));
  if (!fd || !h5) {
    free(fd);
    free(h5);
    return 1;
  }
  free(p->fd);
  free(p->h5);
  p->fd = fd;
  p->h5 = h5;
  p->nfd = n ? n : 1;
  for (a = 0; a < p->nfd; a++) p->fd[a] = -1;
  return 0;
}

int main(int argc, char *argv[])
{
  rawspec_context ctx;
  cli_product prod[MAX_OUTPUTS];
  rawspec_raw_hdr_t raw_hdr;
  char fname[PATH_MAX + 1];
  char *dest = NULL, *dest_port = NULL, *ics_stem = NULL, *pc;
  int net = 0, save_headers = 0, per_ant_out = 0, only_ics = 0, debug = 0, fbh5 = 0;
  int ant = -1, opt, i, si, exit_status = 0;
  unsigned int schan = 0, nchan = 0, outidx = 0;
  double rate = 6.0;
  int readers = 12, nslots = 2, sync_ingest = 0, mfma = 0, stats = 0;
  unsigned int Nc = 0, Ncpa, Np = 0, Ntpb = 0, Nbps = 0;
  int expand4 = 0;
  double t_start = 0, t_end = 0, read_s = 0, issue_s = 0;
  uint64_t bytes_in = 0;
  unsigned long buffers = 0, blocks_total = 0, zero_blocks = 0;
  cli_ingest *ing = NULL;
  const char *argv0 = argv[0];

  printf("rawspec %s using librawspec %s and %s\n", get_librawspec_version(),
         get_librawspec_version(), get_cufft_version());

  memset(&ctx, 0, sizeof ctx);
  memset(prod, 0, sizeof prod);
  ctx.Npolout[0] = 1;

  while ((opt = getopt_long(argc, argv, "a:b:d:f:g:HSjzs:i:n:o:p:r:t:hv", long_opts, NULL)) != -1) {
    switch (opt) {
      case 'h': usage(argv[0]); return 0;
      case 'j': fbh5 = 1; break;
      case 'z': debug = 1; break;
      case 'a': ant = (int)strtol(optarg, NULL, 0); break;
      case 'b': ctx.Nbc = (int)strtol(optarg, NULL, 0); break;
      case 'd':
        dest = optarg;
        if ((pc = strrchr(dest, ':')) != NULL) {
          *pc++ = '\0';
          dest_port = pc;
          net = 1;
        }
        break;
      case 'f': if (parse_list_u(optarg, ctx.Nts, "fine channel counts") < 0) return 1; break;
      case 'g':
        ctx.gpu_index = (int)strtol(optarg, NULL, 0);
        printf("using requested GPU: %d\n", ctx.gpu_index);
        break;
      case 'H': save_headers = 1; break;
      case 'i': {
        char *q = optarg, *end;
        int k;
        printf("writing output for incoherent sum over all antennas\n");
        only_ics = 1;
        ctx.incoherently_sum = 1;
        ctx.Naws = 1;
        for (k = 0; optarg[k]; k++) ctx.Naws += optarg[k] == ',';
        free(ctx.Aws);
        ctx.Aws = (float *)malloc(sizeof(float) * (size_t)ctx.Naws);
        for (k = 0; k < ctx.Naws; k++) {
          ctx.Aws[k] = strtof(q, &end);
          q = (*end == ',') ? end + 1 : end;
        }
        break;
      }
      case 'n': nchan = (unsigned int)strtoul(optarg, NULL, 0); break;
      case 'o': outidx = (unsigned int)strtoul(optarg, NULL, 0); break;
      case 'p': if (parse_list_i(optarg, ctx.Npolout, "pol modes") < 0) return 1; break;
      case 'r': rate = strtod(optarg, NULL); break;
      case 's': schan = (unsigned int)strtoul(optarg, NULL, 0); break;
      case 'S': per_ant_out = 1; break;
      case 't': if (parse_list_u(optarg, ctx.Nas, "integration counts") < 0) return 1; break;
      case 'v': printf("librawspec %s, %s\n", get_librawspec_version(), get_cufft_version()); return 0;
      case 1000: readers = (int)strtol(optarg, NULL, 0); break;
      case 1001: sync_ingest = 1; break;
      case 1002: nslots = (int)strtol(optarg, NULL, 0); break;
      case 1003: mfma = 1; break;
      case 1004: stats = 1; break;
      default:
        printf("Unknown CLI option '%c'\n", opt);
        usage(argv[0]);
        return 1;
    }
  }
  argc -= optind;
  argv += optind;
  if (argc == 0) {
    fprintf(stderr, "error: a file stem must be specified\n");
    usage(argv0);
    return 1;
  }
  if (fbh5 && !net && cli_fbh5_load()) return 1;
  if (ctx.incoherently_sum && per_ant_out) {
    fprintf(stderr, "PLEASE NOTE: Currently, there are potential conflicts in running -i and -S concurrently.\n");
    fprintf(stderr, "PLEASE NOTE: -S (split antennas) is being ignored.\n");
    per_ant_out = 0;
  }
  if (!net) printf(fbh5 ? "writing output files in FBH5 format\n"
                        : "writing output files in SIGPROC Filterbank format\n");
  if (schan != 0 && nchan == 0) {
    fprintf(stderr, "error: nchan must be non-zero if schan is non-zero\n");
    return 1;
  }
  if (save_headers && net) {
    fprintf(stderr, "warning: saving headers is only supported for file output\n");
    save_headers = 0;
  }
  for (i = 0; i < MAX_OUTPUTS; i++) {
    if (ctx.Nts[i] == 0 && ctx.Nas[i] == 0) break;
    if (ctx.Nts[i] == 0 || ctx.Nas[i] == 0) {
      fprintf(stderr, "error: must specify same number of FFT and integration lengths\n");
      return 1;
    }
  }
  ctx.No = (unsigned int)i;
  if (ctx.No == 0) {
    printf("using default FFT and integration lengths\n");
    ctx.No = 3;
    ctx.Nts[0] = 1 << 20; ctx.Nts[1] = 1 << 3; ctx.Nts[2] = 1 << 10;
    ctx.Nas[0] = 51;      ctx.Nas[1] = 128;    ctx.Nas[2] = 3072;
  }
  for (i = 0; i < (int)ctx.No; i++) {
    if (ctx.Npolout[i] == 0 && i > 0) {
      ctx.Npolout[i] = ctx.Npolout[i - 1];
    } else if (ctx.Npolout[i] != 1 && abs(ctx.Npolout[i]) != 4) {
      fprintf(stderr, "error: number of output pols must be 1 or +/- 4\n");
      return 1;
    }
    if (ctx.Npolout[i] != 1 && net) {
      fprintf(stderr, "error: full-pol mode is not supported for network output\n");
      return 1;
    }
  }

  ctx.user_data = prod;
  for (i = 0; i < (int)ctx.No; i++) {
    cli_product *p = &prod[i];
    p->debug_callback = debug;
    p->fb_hdr.machine_id = 20;
    p->fb_hdr.telescope_id = -1;
    p->fb_hdr.data_type = 1;
    p->fb_hdr.nbeams = 1;
    p->fb_hdr.ibeam = -1;
    p->fb_hdr.refbeam = -1;
    p->fb_hdr.nbits = 32;
    p->fb_hdr.nifs = abs(ctx.Npolout[i]);
    p->rate = rate;
    p->Nant = 1;
    p->fd_ics = -1;
    p->net = net;
    p->fbh5 = fbh5 && !net;
    if (reset_fds(p, 1)) return 1;
  }
  if (net) {
    ctx.dump_callback = cli_dump_net_callback;
    prod[0].fd[0] = cli_open_output_socket(dest, dest_port);
    if (prod[0].fd[0] < 0) {
      fprintf(stderr, "cannot open output socket, giving up\n");
      return 1;
    }
    for (i = 1; i < (int)ctx.No; i++) prod[i].fd[0] = prod[0].fd[0];
  } else {
    ctx.dump_callback = cli_dump_file_callback;
  }

  for (si = 0; si < argc; si++) {
    int fi, next_stem = 0, fdhdrs = -1;
    unsigned long bi = 0;
    int64_t pktidx = 0, dpktidx = 0;
    size_t cb = 0;          /* file bytes per channel of one block */

    printf("working stem: %s\n", argv[si]);
    if (ctx.incoherently_sum) {
      size_t n = strlen(argv[si]) + 5;
      free(ics_stem);
      ics_stem = (char *)malloc(n);
      snprintf(ics_stem, n, "%s-ics", argv[si]);
    }

    for (fi = 0;; fi++) {
      int fdin;
      off_t pos;
      struct stat st;
      const char *bfname;
      snprintf(fname, sizeof fname, "%s.%04d.raw", argv[si], fi);
      bfname = strrchr(fname, '/');
      bfname = bfname ? bfname + 1 : fname;
      printf("opening file: %s", fname);
      fdin = open(fname, O_RDONLY);
      if (fdin < 0) {
        printf(" [%s]\n", strerror(errno));
        break;
      }
      printf("\n");
      posix_fadvise(fdin, 0, 0, POSIX_FADV_SEQUENTIAL);
      if (fstat(fdin, &st) != 0) st.st_size = 0;

      pos = rawspec_raw_read_header(fdin, &raw_hdr);
      if (pos <= 0) {
        if (pos < 0)
          fprintf(stderr, "error getting obs params from %s\n", fname);
        else
          fprintf(stderr, "no data found in %s\n", fname);
        close(fdin);
        break;
      }

      if (fi == 0) {
        int input_conjugated;
        if (raw_hdr.obsnchan % raw_hdr.nants != 0) {
          fprintf(stderr, "bad obsnchan/nants: %u %% %u != 0\n", raw_hdr.obsnchan, raw_hdr.nants);
          close(fdin);
          break;
        }
        Nc = raw_hdr.obsnchan;
        Ncpa = raw_hdr.obsnchan / raw_hdr.nants;
        Np = raw_hdr.npol;
        Nbps = raw_hdr.nbits;
        Ntpb = (unsigned int)(raw_hdr.blocsize / ((2 * Np * Nc * Nbps) / 8));
        pktidx = raw_hdr.pktidx;
        dpktidx = 0;
        if ((size_t)((2 * Np * Nc * Nbps) / 8) * Ntpb != raw_hdr.blocsize) {
          printf("bad block geometry: 2*%upol*%uchan*%utpb*(%ubps/8) != %lu\n", Np, Nc, Ntpb, Nbps,
                 (unsigned long)raw_hdr.blocsize);
          close(fdin);
          break;
        }
        if (raw_hdr.nants > 1 && !(per_ant_out || ctx.incoherently_sum)) {
          printf("NANTS = %d >1: Enabling --split-ant in lieu of neither --split-ant nor --ics flags.\n",
                 raw_hdr.nants);
          per_ant_out = 1;
        }
        if (per_ant_out) {
          if (!net) {
            if (ant != -1) printf("Ignoring --ant %d option:\n\t", ant);
            printf("Splitting output per %d antennas\n", raw_hdr.nants);
            for (i = 0; i < (int)ctx.No; i++) {
              if (prod[i].Nant != raw_hdr.nants || prod[i].nfd != raw_hdr.nants) {
                if (cli_close_product_files(&prod[i])) exit_status = 1;
                prod[i].per_ant_out = 1;
                if (reset_fds(&prod[i], raw_hdr.nants)) return 1;
              }
            }
          } else {
            printf("Ignoring --splitant flag in network mode\n");
          }
          only_ics = 0;
        }
        if (ant != -1 && !per_ant_out) {
          if (ant > (int)raw_hdr.nants - 1 || ant < 0) {
            printf("bad antenna selection: ant <> {0, nants} (%u <> {0, %d})\n", ant, raw_hdr.nants);
            close(fdin);
            break;
          }
          if (schan >= Ncpa) {
            printf("bad schan specification with antenna selection: "
                   "schan > antnchan {obsnchan/nants} (%u > %u {%d/%d})\n",
                   schan, Ncpa, raw_hdr.obsnchan, raw_hdr.nants);
            close(fdin);
            break;
          }
          printf("Selection of antenna %d equates to a starting channel of %d\n", ant, ant * Ncpa);
          schan += (unsigned int)ant * Ncpa;
          Nc = Ncpa;
        }
        if (nchan != 0) {
          if (ant == -1 && schan + nchan > Nc) {
            printf("bad channel range: schan + nchan > obsnchan (%u + %u > %d)\n", schan, nchan,
                   raw_hdr.obsnchan);
            close(fdin);
            break;
          } else if (ant != -1 && schan + nchan > (unsigned int)(ant + 1) * Ncpa) {
            printf("bad channel range: schan + nchan > antnchan {obsnchan/nants} (%u + %u > %d {%d/%d})\n",
                   schan - ant * Ncpa, nchan, Ncpa, raw_hdr.obsnchan, raw_hdr.nants);
            close(fdin);
            break;
          }
          Nc = nchan;
        }
        input_conjugated = raw_hdr.obsbw < 0 ? 1 : 0;

        if (Nc != ctx.Nc || Np != ctx.Np || Nbps != ctx.Nbps || Ntpb != ctx.Ntpb ||
            input_conjugated != ctx.input_conjugated) {
          if (ing) {
            cli_ingest_destroy(ing);
            ing = NULL;
          }
          if (ctx.Ntpb != 0) rawspec_cleanup(&ctx);
          ctx.Nant = raw_hdr.nants;
          ctx.Nc = Nc;
          ctx.Np = Np;
          ctx.Ntpb = Ntpb;
          ctx.Nbps = Nbps;
          ctx.input_conjugated = input_conjugated;
          ctx.Nb = 0;
          ctx.Nb_host = 0;
          ctx.h_blkbufs = NULL;
          if (rawspec_initialize(&ctx)) {
            fprintf(stderr, "rawspec initialization failed\n");
            return 1;
          }
          if (mfma) rawspec_amd_set_mfma(&ctx, 1);
          expand4 = (ctx.Nbps == 8 && Nbps == 4);
          if (expand4)
            printf("device memory initialised for %d bits per sample,\n\t"
                   "will expand header specified %d bits per sample.\n", ctx.Nbps, Nbps);
          for (i = 0; i < (int)ctx.No; i++) {
            prod[i].h_pwrbuf = ctx.h_pwrbuf[i];
            prod[i].h_pwrbuf_size = ctx.h_pwrbuf_size[i];
            prod[i].h_icsbuf = ctx.h_icsbuf[i];
            prod[i].Nds = ctx.Nds[i];
            prod[i].Nf = ctx.Nts[i] * ctx.Nc;
            if (debug) printf("output %d Nds = %u, Nf = %u\n", i, prod[i].Nds, prod[i].Nf);
            prod[i].Nant = raw_hdr.nants;
          }
        } else {
          printf("resetting integration buffers for new stem\n");
          rawspec_reset_integration(&ctx);
        }
        cb = (size_t)((2 * ctx.Np * Nbps) / 8) * ctx.Ntpb;
        if (!ing) {
          ing = cli_ingest_create(&ctx, !sync_ingest, expand4, cb, readers, nslots);
          if (!ing) return 1;
        }

        for (i = 0; i < (int)ctx.No; i++) {
          cli_product *p = &prod[i];
          const unsigned int ncpa = raw_hdr.obsnchan / raw_hdr.nants;
          p->fb_hdr.telescope_id = fb_telescope_id(raw_hdr.telescop);
          p->fb_hdr.src_raj = raw_hdr.ra;
          p->fb_hdr.src_dej = raw_hdr.dec;
          p->fb_hdr.tstart = raw_hdr.mjd;
          p->fb_hdr.ibeam = raw_hdr.beam_id;
          p->fb_hdr.refbeam = raw_hdr.refbeam;
          strncpy(p->fb_hdr.source_name, raw_hdr.src_name, 80);
          p->fb_hdr.source_name[80] = '\0';
          strncpy(p->fb_hdr.rawdatafile, bfname, 80);
          p->fb_hdr.rawdatafile[80] = '\0';
          /* fine channel width, and the centre frequency of the first fine
           * channel: the first coarse channel's centre, less half the fine
           * channels, moved up by the skipped coarse channels */
          p->fb_hdr.foff = raw_hdr.obsbw / ncpa / ctx.Nts[i];
          p->fb_hdr.fch1 = raw_hdr.obsfreq
                           - raw_hdr.obsbw * (ncpa - 1) / (2 * raw_hdr.obsnchan / raw_hdr.nants)
                           - (ctx.Nts[i] / 2) * p->fb_hdr.foff
                           + (schan % ncpa) * raw_hdr.obsbw / ncpa;
          p->fb_hdr.nfpc = (int)ctx.Nts[i];
          p->fb_hdr.nchans = (int)(ctx.Nc * ctx.Nts[i] / raw_hdr.nants);
          p->fb_hdr.tsamp = raw_hdr.tbin * ctx.Nts[i] * ctx.Nas[i];
          if (!net) {
            if (!only_ics && cli_open_product_files(p, dest, argv[si], (int)outidx + i)) return 1;
            if (ctx.incoherently_sum && cli_open_ics_file(p, dest, ics_stem, (int)outidx + i))
              return 1;
          }
        }
        if (save_headers) {
          fdhdrs = cli_open_headers_file(dest, argv[si]);
          if (fdhdrs < 0) fprintf(stderr, "unable to save headers\n");
        }
        if (net) {
          double sum_inv_na = 0;
          for (i = 0; i < (int)ctx.No; i++) sum_inv_na += 1.0 / ctx.Nas[i];
          for (i = 0; i < (int)ctx.No; i++) {
            prod[i].rate = rate / ctx.Nas[i] / sum_inv_na;
            fprintf(stderr, "output product %d data rate %6.3f Gbps\n", i, prod[i].rate);
          }
        }
      } /* first file */

      if (t_start == 0) t_start = now_s();

      for (;;) {   /* every block of the file */
        const off_t skip = (off_t)schan * (off_t)cb;
        int fatal = 0;
        if (save_headers && fdhdrs >= 0) {
          off_t hp = raw_hdr.hdr_pos;
          if (sendfile(fdhdrs, fdin, &hp, raw_hdr.hdr_size) < 0) perror("sendfile");
        }
        if (dpktidx == 0 && raw_hdr.pktidx > pktidx) dpktidx = raw_hdr.pktidx - pktidx;
        if (raw_hdr.pktidx - pktidx != dpktidx) {
          if (raw_hdr.pktidx < pktidx) {
            printf("got backwards jump in pktidx: %ld -> %ld\n", (long)pktidx, (long)raw_hdr.pktidx);
            next_stem = 1;
            break;
          } else if ((raw_hdr.pktidx - pktidx) % dpktidx != 0) {
            printf("got misaligned jump in pktidx: (%ld - %ld) %% %ld != 0\n", (long)raw_hdr.pktidx,
                   (long)pktidx, (long)dpktidx);
            next_stem = 1;
            break;
          } else if (raw_hdr.pktidx == pktidx) {
            printf("got null jump in pktidx: (%ld - %ld) == 0\n", (long)raw_hdr.pktidx, (long)pktidx);
            break;
          }
          while (raw_hdr.pktidx - pktidx != dpktidx) {
            pktidx += dpktidx;
            cli_ingest_zero_block(ing, bi);
            zero_blocks++;
            if (bi % ctx.Nb == ctx.Nb - 1 && cli_ingest_submit(ing, &fatal)) {
              if (fatal) return 1;
              next_stem = 1;
              break;
            }
            bi++;
          }
          if (next_stem) break;
        }
        /* the block's selected channels must lie within the file */
        if (pos + skip + (off_t)((size_t)ctx.Nc * cb) > st.st_size) {
          fprintf(stderr, "incomplete block at EOF\n");
          next_stem = 1;
          break;
        }
        cli_ingest_block(ing, bi, fdin, pos + skip);
        blocks_total++;
        if (bi % ctx.Nb == ctx.Nb - 1 && cli_ingest_submit(ing, &fatal)) {
          if (fatal) return 1;
          next_stem = 1;
          break;
        }
        pktidx = raw_hdr.pktidx;
        bi++;
        /* next header: past all obsnchan channels of this block */
        if (lseek(fdin, pos + (off_t)raw_hdr.obsnchan * (off_t)cb, SEEK_SET) < 0) break;
        pos = rawspec_raw_read_header(fdin, &raw_hdr);
        if (pos <= 0) {
          if (pos < 0) fprintf(stderr, "error getting obs params from %s [%s]\n", fname, strerror(errno));
          break;
        }
      }
      close(fdin);
      if (next_stem) break;
    } /* files */

    if (ing) {
      int fatal = 0;
      if (cli_ingest_flush(ing, &fatal) && fatal) return 1;
      cli_ingest_discard(ing);
    }
    if (ctx.Nc) rawspec_wait_for_completion(&ctx);
    if (t_start != 0) t_end = now_s();
    if (fdhdrs >= 0) close(fdhdrs);
    if (!net)
      for (i = 0; i < (int)ctx.No; i++)
        if (cli_close_product_files(&prod[i])) exit_status = 1;
  } /* stems */

  if (ing) {
    bytes_in = cli_ingest_bytes(ing);
    buffers = cli_ingest_buffers(ing);
    read_s = cli_ingest_read_seconds(ing);
    issue_s = cli_ingest_issue_seconds(ing);
    cli_ingest_destroy(ing);
  }
  rawspec_cleanup(&ctx);
  free(ics_stem);
  free(ctx.Aws);
  if (net && prod[0].fd && prod[0].fd[0] >= 0) close(prod[0].fd[0]);
  for (i = 0; i < (int)ctx.No; i++) {
    free(prod[i].fd);
    free(prod[i].h5);
    free(prod[i].gather);
  }

  {
    uint64_t tp = 0, tb = 0, tn = 0;
    for (i = 0; i < (int)ctx.No; i++) {
      printf("output product %d: %u spectra", i, prod[i].total_spectra);
      if (prod[i].total_packets > 0) {
        printf(" (%u packets, %.3f Gbps)", prod[i].total_packets,
               8.0 * (double)prod[i].total_bytes / (double)prod[i].total_ns);
        tp += prod[i].total_packets;
        tb += prod[i].total_bytes;
        tn += prod[i].total_ns;
      }
      printf("\n");
    }
    if (tn > 0)
      printf("combined total  : %lu packets, %.3f Gbps\n", (unsigned long)tp, 8.0 * (double)tb / (double)tn);
  }
  if (stats) {
    const double dt = t_end > t_start ? t_end - t_start : 0;
    fprintf(stderr,
            "{\"ingest_bytes\": %llu, \"blocks\": %lu, \"zero_blocks\": %lu, \"buffers\": %lu, "
            "\"seconds\": %.6f, \"read_wait_seconds\": %.6f, \"issue_seconds\": %.6f, "
            "\"GBps\": %.4f, \"mode\": \"%s\", \"readers\": %d}\n",
            (unsigned long long)bytes_in, blocks_total, zero_blocks, buffers, dt, read_s, issue_s,
            dt > 0 ? (double)bytes_in / dt / 1e9 : 0.0,
            (sync_ingest || expand4) ? "sync" : "pipelined", readers);
  }
  if (exit_status) fprintf(stderr, "*** At least one error occured during processing!\n");
  return exit_status;
}
