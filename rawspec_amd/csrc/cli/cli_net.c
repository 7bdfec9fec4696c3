/*
 * cli_net.c -- UDP output of integrated spectra as self-describing
 * filterbank packets (the reference's rawspec_socket.c:21-325).
 *
 * Each packet is a SIGPROC filterbank header describing the packet's slice
 * (nchans = channels in the packet, fch1 = frequency of its first channel,
 * tstart = time of its first spectrum) followed by float32 spectra.  A
 * packet holds at most 2048 floats: a dump of nchans channels goes out as
 * ceil(nchans / 2048) packets per spectrum when nchans >= 2048, otherwise
 * as packets of floor(2048 / nchans) whole spectra.  The header is padded
 * (with extra "rawdatafile" keywords) so that the data starts on a 512-byte
 * boundary of a packet-socket frame, whose UDP payload begins 0x6c bytes
 * into the frame; padding shorter than 20 bytes cannot be expressed, so
 * another 512 bytes are added then.  Sending is throttled to the product's
 * share of the requested rate.
 */
#define _GNU_SOURCE 1
#include <errno.h>
#include <netdb.h>
#include <netinet/in.h>
#include <stdio.h>
#include <string.h>
#include <sys/socket.h>
#include <time.h>
#include <unistd.h>

#include "cli.h"

enum { kMinMtu = 8600, kFloatsPerPacket = 8192 / sizeof(float), kFrameHeader = 0x6c };

int cli_open_output_socket(const char *host, const char *port)
{
  struct addrinfo hints, *res = NULL, *rp;
  int fd = -1, rc;
  unsigned int mtu = 0;
  socklen_t len = sizeof mtu;
  int probe = IP_PMTUDISC_PROBE;

  memset(&hints, 0, sizeof hints);
  hints.ai_family = AF_INET;
  hints.ai_socktype = SOCK_DGRAM;
  rc = getaddrinfo(host, port, &hints, &res);
  if (rc != 0) {
    fprintf(stderr, "getaddrinfo: %s\n", gai_strerror(rc));
    return -1;
  }
  for (rp = res; rp; rp = rp->ai_next) {
    fd = socket(rp->ai_family, rp->ai_socktype, rp->ai_protocol);
    if (fd < 0) continue;
    if (connect(fd, rp->ai_addr, rp->ai_addrlen) == 0) break;
    close(fd);
    fd = -1;
  }
  freeaddrinfo(res);
  if (fd < 0) {
    fprintf(stderr, "could not open output socket\n");
    return -1;
  }
  if (getsockopt(fd, IPPROTO_IP, IP_MTU, &mtu, &len) < 0) {
    perror("getsockopt");
    fprintf(stderr, "warning: could not determine MTU\n");
  } else if (mtu < kMinMtu) {
    fprintf(stderr, "MTU %u is too small, need at least %u\n", mtu, (unsigned)kMinMtu);
    close(fd);
    return -1;
  }
  /* don't-fragment on, path-MTU changes ignored */
  if (setsockopt(fd, IPPROTO_IP, IP_MTU_DISCOVER, &probe, sizeof probe) < 0) {
    perror("setsockopt");
    fprintf(stderr, "warning: could not set IP_MTU_DISCOVER to IP_PMTUDISC_PROBE\n");
  }
  return fd;
}

static uint64_t now_ns(void)
{
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (uint64_t)t.tv_sec * 1000000000ull + (uint64_t)t.tv_nsec;
}

static void sleep_ns(int64_t ns)
{
  struct timespec t;
  if (ns <= 0) return;
  t.tv_sec = ns / 1000000000;
  t.tv_nsec = ns % 1000000000;
  nanosleep(&t, NULL);
}

/* Header bytes, padded so that the data begins on a 512-byte frame boundary. */
static char *packet_header(char *pkt, const fb_hdr_t *h)
{
  char *end = (char *)fb_buf_write_header(pkt, h);
  const long plain = kFrameHeader + (end - pkt);
  long off = (plain + 511) / 512 * 512;
  if (off > plain) {
    if (off < plain + 20) off += 512;
    end = (char *)fb_buf_write_padded_header(pkt, h, (int32_t)(off - kFrameHeader));
  }
  return end;
}

void cli_send_dump(cli_product *p)
{
  static __thread char pkt[90000];
  fb_hdr_t *h = &p->fb_hdr;
  const int nchans = h->nchans;
  const double fch1 = h->fch1;
  const uint64_t t0 = now_ns();
  int chan_per_pkt, spec_per_pkt;
  unsigned int spec_left = p->Nds;
  uint64_t bytes = 0;
  unsigned int packets = 0, errors = 0;

  if ((size_t)nchans >= kFloatsPerPacket) {
    chan_per_pkt = (int)kFloatsPerPacket;
    spec_per_pkt = 1;
  } else {
    chan_per_pkt = nchans;
    spec_per_pkt = (int)(kFloatsPerPacket / (size_t)nchans);
    if ((unsigned)spec_per_pkt > spec_left) spec_per_pkt = (int)spec_left;
  }

  while (spec_left) {
    const int nspec = (unsigned)spec_per_pkt < spec_left ? spec_per_pkt : (int)spec_left;
    const float *pw = p->h_pwrbuf + (size_t)(p->Nds - spec_left) * (size_t)nchans;
    int chan_left = nchans;
    h->fch1 = fch1;
    while (chan_left) {
      const int nch = chan_per_pkt < chan_left ? chan_per_pkt : chan_left;
      char *q;
      size_t len;
      int s;
      int64_t lag;
      h->nchans = nch;
      q = packet_header(pkt, h);
      for (s = 0; s < nspec; s++) {
        memcpy(q, pw + (size_t)s * (size_t)nchans, (size_t)nch * sizeof(float));
        q += (size_t)nch * sizeof(float);
      }
      len = (size_t)(q - pkt);
      packets++;
      bytes += len;
      /* ENOTCONN (no listener yet) is retried once: receivers that capture
       * with packet sockets do not bind the port */
      if (send(p->fd[0], pkt, len, 0) < 0 && errno == ENOTCONN &&
          send(p->fd[0], pkt, len, 0) < 0)
        errors++;
      lag = (int64_t)(8.0 * (double)bytes / p->rate) - (int64_t)(now_ns() - t0);
      if (lag > 100 * 1000) sleep_ns(lag);
      pw += nch;
      h->fch1 += nch * h->foff;
      chan_left -= nch;
    }
    h->tstart += nspec * h->tsamp / 86400.0;
    spec_left -= (unsigned)nspec;
  }
  h->fch1 = fch1;
  h->nchans = nchans;
  if (errors)
    printf("fine channels %10d: error packets %u/%u\n", p->Nf, errors, packets);
  sleep_ns((int64_t)(8.0 * (double)bytes / p->rate) - (int64_t)(now_ns() - t0));
  p->total_spectra += p->Nds;
  p->total_packets += packets;
  p->total_bytes += bytes;
  p->total_ns += now_ns() - t0;
}
