// rsd_runtime.hip -- device layer of librawspec: memory, stream, copies,
// launch planning and kernel dispatch behind the C ABI of rsd.h.
//
// HBM layout per context:
//   d_in    [Nc][Nb*Ntpb][Np][re,im] raw samples (the layout the reference's
//           2-D block copy builds, rawspec_gpu.cu:1454-1460)
//   d_acc   per product, [Nds][|P|][Nc][Nt] float: the integration in
//           progress, already in the host dump layout and FFT-shifted, so a
//           dump is one contiguous device->host copy
//   d_ics   per product, [Nds][|P|][Nc/Nant][Nt] float
//   d_part  partial integrations of groups split over workgroups
//   d_partb partial sums of a four-step pass B whose dump groups are split
//           into runs of spectra (products with few k2 tiles and channels)
//   d_work  four-step intermediate of one channel chunk (several buffers'
//           spectra when batched), tiled for pass B
//   d_held  per four-step product integrating over several buffers, the
//           intermediate of up to hold_k buffers over all channels, so that
//           one-buffer calls run pass B once per hold_k buffers (allocated
//           at the first integration that uses it)
//   d_tw    W_4096^m table; per four-step product the fine table W_N^r,
//           r < N1.  All twiddles are computed in FP64 on the host and
//           rounded once to FP32.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "kernels.h"
#include "rsd.h"

#define RSD_CHECK(call)                                                           \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      fprintf(stderr, "librawspec: %s failed: %s (%s:%d)\n", #call,               \
              hipGetErrorString(e_), __FILE__, __LINE__);                         \
      fflush(stderr);                                                             \
      return 1;                                                                   \
    }                                                                             \
  } while (0)

namespace {

struct ProdPlan {
  rsd_product p;
  int Q;
  int small;               // N <= 4096 (or Bluestein): fused single kernel
  // a length that is not a power of two (<= kBsMaxN): Bluestein's identity
  // in the small kernel at the padded power-of-two length bs_m >= 2 N - 1,
  // with the chirp c[n] = exp(-i pi n^2 / N) and the padded kernel's FFT
  int bs;
  unsigned bs_m;
  float2 *d_chirp;
  float2 *d_bsk;
  // small path
  unsigned na_buf, Ng, chunk, G, gpw, spg, bpc;
  int direct;
  // matrix-core small kernel (N = 1024, two pols): one wavefront per chunk
  int m_ok;
  unsigned m_chunk, m_G;
  int m_unsplit;           // int8 samples, long enough integration: unsplit FP16 operands allowed
  // four-step path
  unsigned N1, N2;
  unsigned fs_chunk;       // channels per pass A/B launch pair (one buffer)
  unsigned fs_bmax;        // buffers per batched launch pair (1: no batching)
  int fs_mfma;             // pass B on the matrix cores for the current integration
  int fs_y24;              // 24-bit intermediate for the current integration
  int fs_amfma;            // pass A on the matrix cores for the current integration
  // launch-plan tunables (defaults below; rsd_autotune times alternatives)
  unsigned t_target;       // small kernel: workgroups a launch should reach
  unsigned t_maxspg;       // small kernel: most slots per dump group
  unsigned t_mtarget;      // matrix-core small kernel: wavefronts a launch should reach
  size_t fs_budget;        // four-step: bytes of Y per pass A/B launch pair
  // held first-pass output (one-buffer calls of an integration over several
  // buffers): capacity hold_k buffers; `holding` is latched per integration
  unsigned hold_k;
  unsigned hold_use;       // slots used in the current integration (<= hold_k)
  float2 *d_held;
  int hold_failed;         // allocation refused: not retried
  int holding;
  unsigned held;           // buffers whose pass B is pending
  int held_first;          // the pending buffers start the integration
  float2 *d_twf;           // W_N^r, r < N1 (pass A's fine twiddle table)
  float *d_acc_int;        // running sum across buffers (Nis > 1)
  // outputs
  float *d_acc;
  size_t acc_len;
  float *d_ics;
  size_t ics_len;
};

// Four-step intermediate budget per pass A/B launch pair (one buffer's or a
// batch's spectra for a chunk of channels).  Larger launches keep more of
// the device busy and outweigh keeping the intermediate inside the 256 MB
// Infinity Cache: measured on 64 channels x 2^20 points, one launch pair over
// all channels (1 GiB) runs 8% faster than 6-channel (96 MiB) chunks, and a
// batch of 8 buffers x 12 channels 6% faster than 8 x 3.
#ifndef RSD_FS_WORK_MB
#define RSD_FS_WORK_MB 1536
#endif
constexpr size_t kFsWorkBytes = (size_t)RSD_FS_WORK_MB << 20;
// smallest input buffer for which a context runs its products on two lanes
constexpr size_t kLaneMinBytes = (size_t)64 << 20;
// Pass-A workgroups per CU: each loops over its share of the launch's items
// (0: one workgroup per item).
#ifndef RSD_FSA_WG_PER_CU
#define RSD_FSA_WG_PER_CU 8
#endif
// Held first-pass output per context, at most (one-buffer calls): 8 buffers
// of a 256 MiB, 2^20-point product
#ifndef RSD_HOLD_MB
#define RSD_HOLD_MB 8192
#endif
constexpr size_t kHoldBytes = (size_t)RSD_HOLD_MB << 20;
// Vector pass B: workgroups per CU a launch aims for, splitting each dump
// group's spectra into runs when its (k2 tile, group, channel) workgroups
// are fewer; the fewest spectra a run integrates; the CU count the planner
// sizes the runs' partial sums for (launches use the device's own count,
// capped by what was allocated).
#ifndef RSD_FSB_MIN_RUN
#define RSD_FSB_MIN_RUN 8
#endif
constexpr unsigned kFsbWgPerCu = 4;
constexpr unsigned kFsbMinRun = RSD_FSB_MIN_RUN;
constexpr unsigned kPlanCus = 256;

bool is_pow2(unsigned n) { return n && !(n & (n - 1)); }
unsigned ilog2u(unsigned n) { unsigned l = 0; while (n > 1) { n >>= 1; ++l; } return l; }

// Longest transform that is not a power of two: Bluestein's identity pads
// it to at least 2 N - 1 points, and the small kernel holds 4096.
constexpr unsigned kBsMaxN = 2048;

// Bluestein tables of an N-point transform padded to M points, computed in
// FP64 and rounded once: chirp[n] = exp(-i pi n^2 / N) (n^2 reduced mod 2N
// exactly), and bsk = FFT_M(b) / M with b[m] = conj(chirp[|m|]) for
// |m| < N (indices mod M), zero elsewhere.
static void bluestein_tables(unsigned N, unsigned M, std::vector<float2> &chirp,
                             std::vector<float2> &bsk) {
  std::vector<double> cr(N), ci(N);
  for (unsigned n = 0; n < N; ++n) {
    const unsigned long long r = ((unsigned long long)n * n) % (2ull * N);
    const double ang = -M_PI * (double)r / (double)N;
    cr[n] = cos(ang);
    ci[n] = sin(ang);
  }
  std::vector<double> br(M, 0.0), bi(M, 0.0);
  for (unsigned m = 0; m < N; ++m) {
    br[m] = cr[m];
    bi[m] = -ci[m];
    if (m) {
      br[M - m] = cr[m];
      bi[M - m] = -ci[m];
    }
  }
  // iterative radix-2 FFT, exp(-2 pi i k m / M)
  for (unsigned i = 1, j = 0; i < M; ++i) {
    unsigned bit = M >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(br[i], br[j]);
      std::swap(bi[i], bi[j]);
    }
  }
  for (unsigned len = 2; len <= M; len <<= 1) {
    for (unsigned k = 0; k < len / 2; ++k) {
      const double ang = -2.0 * M_PI * (double)k / (double)len;
      const double wr = cos(ang), wi = sin(ang);
      for (unsigned i = 0; i < M; i += len) {
        const unsigned u = i + k, v = i + k + len / 2;
        const double tr = br[v] * wr - bi[v] * wi, ti = br[v] * wi + bi[v] * wr;
        br[v] = br[u] - tr;
        bi[v] = bi[u] - ti;
        br[u] += tr;
        bi[u] += ti;
      }
    }
  }
  chirp.resize(N);
  bsk.resize(M);
  for (unsigned n = 0; n < N; ++n) chirp[n] = make_float2((float)cr[n], (float)ci[n]);
  for (unsigned m = 0; m < M; ++m) bsk[m] = make_float2((float)(br[m] / M), (float)(bi[m] / M));
}

// forward twiddles exp(-2 pi i (m * mul) / n) for m < cnt, FP64 -> FP32
std::vector<float2> twiddles(unsigned long long n, unsigned long long cnt, unsigned long long mul) {
  std::vector<float2> t(cnt);
  for (unsigned long long m = 0; m < cnt; ++m) {
    const unsigned long long e = (m * mul) % n;
    const double a = -2.0 * M_PI * (double)e / (double)n;
    t[m] = make_float2((float)cos(a), (float)sin(a));
  }
  return t;
}

}  // namespace

struct rsd_ctx {
  int dev;
  hipStream_t stream;                  // the stream the context's work goes to
  hipStream_t own_stream;              // created with the context (rsd_set_stream may replace `stream`)
  // Products run on two lanes forked from `stream` and joined back into it:
  // four-step products (which share d_work) on lane 0, small products (which
  // share d_part) on lane 1, so the vector-bound small transforms overlap
  // the memory-bound four-step passes.  Dumps stay on `stream`.
  hipStream_t lane[2];
  hipEvent_t ev_main, ev_lane[2];
  bool lanes;
  rsd_config cfg;
  std::vector<float> weights;
  unsigned long long sample_bytes;     // Np * 2 * Nbps/8
  unsigned long long chan_block_bytes; // Ntpb * sample_bytes
  unsigned long long chan_stride;      // Nb * chan_block_bytes
  uint8_t *d_in;
  size_t in_bytes;
  uint8_t *d_stage;                    // one complex4 block
  float *d_part;
  size_t part_len;
  float2 *d_work;
  size_t work_len;
  float *d_partb;
  size_t partb_len;
  unsigned cus;                        // compute units of the device
  unsigned hold_max;                   // buffers a product may hold (rsd_set_hold)
  size_t held_bytes;                   // allocated for held output
  float2 *d_tw;
  float *d_w;
  ProdPlan pp[RSD_MAX_PRODUCTS];
  // a fused pair (-1: none): 8-point product f8_a computed by the kernel of
  // 1024-point product f8_b when the engine runs that one unsplit
  int f8_a, f8_b;
  // kernel timing (rsd_profile)
  int mfma;                            // matrix-core engine requested (rsd_set_mfma)
  // allocation-free call in progress (rsd_noalloc_begin .. rsd_noalloc_end):
  // d_part and d_work point into the device's static scratch, and the
  // context's own buffers and four-step budgets are parked here
  bool noalloc;
  float *sv_part;
  size_t sv_part_len;
  float2 *sv_work;
  size_t sv_work_len;
  float *sv_partb;
  size_t sv_partb_len;
  size_t sv_budget[RSD_MAX_PRODUCTS];
  bool prof;
  struct Rec { std::string name; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
};

static hipEvent_t ev_take(rsd_ctx *d) {
  if (!d->pool.empty()) {
    hipEvent_t e = d->pool.back();
    d->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Lanes are used unless kernel timing is on: timed kernels run one after
// another on the context stream, so each event pair brackets one kernel
// running alone.
static bool use_lanes(const rsd_ctx *d) { return d->lanes && !d->prof; }

static hipStream_t prod_stream(const rsd_ctx *d, const ProdPlan &pp) {
  return use_lanes(d) ? d->lane[pp.small ? 1 : 0] : d->stream;
}

// Brackets one kernel launch with events when timing is on.
struct ProfScope {
  rsd_ctx *d;
  long idx;
  hipStream_t s;
  ProfScope(rsd_ctx *dd, int product, const char *kname, unsigned n)
      : ProfScope(dd, product, kname, n, dd->stream) {}
  ProfScope(rsd_ctx *dd, int product, const char *kname, unsigned n, hipStream_t st)
      : d(dd), idx(-1), s(st) {
    if (!d->prof || d->noalloc) return;
    hipEvent_t a = ev_take(d), b = ev_take(d);
    if (!a || !b) return;
    char nm[64];
    snprintf(nm, sizeof nm, "p%d.%s.N%u", product, kname, n);
    d->recs.push_back({nm, a, b});
    idx = (long)d->recs.size() - 1;
    (void)hipEventRecord(a, s);
  }
  ~ProfScope() {
    if (idx >= 0) (void)hipEventRecord(d->recs[idx].b, s);
  }
};

static void plan_small(rsd_ctx *d, ProdPlan &pp) {
  const unsigned slots = (unsigned)rsk::small_slots((int)(pp.bs ? pp.bs_m : pp.p.Nt));
  const unsigned Nc = d->cfg.Nc;
  pp.na_buf = pp.p.Na < pp.p.Nss ? pp.p.Na : pp.p.Nss;
  pp.Ng = pp.p.Nss / pp.na_buf;
  pp.chunk = pp.na_buf;
  // Split long groups over workgroups until the launch fills the device
  // (by default about 2048 workgroups for 256 CUs), keeping >= one spectrum
  // per slot.
  const unsigned long long target = pp.t_target;
  while ((unsigned long long)Nc * pp.Ng * (pp.na_buf / pp.chunk) < target && pp.chunk % 2 == 0 &&
         pp.chunk / 2 >= slots)
    pp.chunk /= 2;
  pp.G = pp.na_buf / pp.chunk;
  pp.direct = pp.G == 1;
  // Slots per group: a power of two, no larger than the group and at most
  // t_maxspg (default 16), so each slot integrates several spectra in
  // registers and the final cross-slot sum stays short; the remaining slots
  // take further groups (gpw groups per workgroup, only when the group is
  // not split).
  const unsigned kMaxSpg = pp.t_maxspg;
  pp.gpw = 1;
  pp.spg = slots;
  if (pp.G == 1) {
    unsigned spg = 1;
    while (spg < pp.chunk && spg < kMaxSpg) spg <<= 1;
    if (spg < slots) {
      pp.spg = spg;
      pp.gpw = slots / spg;
    }
  }
  pp.bpc = ((pp.Ng + pp.gpw - 1) / pp.gpw) * pp.G;
}

// Shortest integration (spectra per dump) for which the matrix-core engine
// at its 1e-3 bar (rsd_set_mfma level 1) uses unsplit FP16 operands: their
// rounding error is random from spectrum to spectrum and falls as
// 1/sqrt(spectra) from about 6e-3 for one spectrum to a floor near 1.5e-4
// (3e-4 at 32 spectra in a numpy emulation of the roundings on int8 noise
// with a strong tone).  Shorter integrations keep split operands.
constexpr unsigned kUnsplitMinSpectra = 32;

// Matrix-core small kernel plan: halve the chunks until t_mtarget
// wavefronts (default 2048, eight per CU) are busy, keeping >= 8 spectra per
// wavefront.
static void plan_small_mfma(rsd_ctx *d, ProdPlan &pp) {
  pp.m_ok = rsk::small_mfma_ok((int)pp.p.Nt, (int)d->cfg.Np);
  pp.m_unsplit = d->cfg.Nbps == 8 && pp.p.Na >= kUnsplitMinSpectra;
  pp.m_chunk = pp.na_buf;
  while ((unsigned long long)d->cfg.Nc * pp.Ng * (pp.na_buf / pp.m_chunk) < pp.t_mtarget &&
         pp.m_chunk % 2 == 0 && pp.m_chunk / 2 >= 8)
    pp.m_chunk /= 2;
  pp.m_G = pp.na_buf / pp.m_chunk;
}

static size_t part_need(const rsd_ctx *d, const ProdPlan &pp) {
  if (!pp.small) return 0;
  size_t n = pp.direct ? 0 : (size_t)d->cfg.Nc * pp.Ng * pp.G * pp.Q * pp.p.Nt;
  if (pp.m_ok && pp.m_G > 1) {
    const size_t m = (size_t)d->cfg.Nc * pp.Ng * pp.m_G * pp.Q * pp.p.Nt;
    if (m > n) n = m;
  }
  return n;
}

// ---- static scratch of the allocation-free entry point ---------------------
// rawspec_amd_start_processing_device_noalloc runs the same kernels with the
// small products' partial sums and the four-step intermediate in this array
// instead of the context's own scratch buffers, so the call allocates
// nothing.  Each GPU holds its own copy (the code object's) from module load
// until the process ends; its device address is looked up once per device at
// context creation (hipGetSymbolAddress) and only that address is handed to
// kernels.  Nothing in it must start at zero: every kernel of a call reads
// only what an earlier kernel of the same call wrote.  All calls on one GPU
// share it, so allocation-free calls that overlap on one GPU (different
// contexts' streams) race.
#ifndef RSD_SCRATCH_MB
#define RSD_SCRATCH_MB 256
#endif
constexpr size_t kScratchBytes = (size_t)RSD_SCRATCH_MB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kScratchDevices = 64;
static void *g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

// Device address of the scratch on device dev (the current device).
static int scratch_setup(int dev) {
  if (dev < 0 || dev >= kScratchDevices) {
    fprintf(stderr, "librawspec: device %d is beyond the %d devices the scratch table covers\n", dev,
            kScratchDevices);
    fflush(stderr);
    return 1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[dev]) return 0;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    fprintf(stderr, "librawspec: hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
            hipGetErrorString(e));
    fflush(stderr);
    return 1;
  }
  g_scratch_ptr[dev] = p;
  return 0;
}

static int alloc_zero(void **p, size_t bytes) {
  if (bytes == 0) { *p = nullptr; return 0; }
  RSD_CHECK(hipMalloc(p, bytes));
  RSD_CHECK(hipMemset(*p, 0, bytes));
  return 0;
}

extern "C" {

#define RSD_STR1(x) #x
#define RSD_STR(x) RSD_STR1(x)

const char *rsd_engine_name(void) {
  return "rawspec-amd gfx950 HIP FFT kernels (HIP " RSD_STR(HIP_VERSION_MAJOR) "." RSD_STR(
      HIP_VERSION_MINOR) "." RSD_STR(HIP_VERSION_PATCH) ")";
}

int rsd_host_alloc(int gpu_index, void **p, size_t bytes, int write_combined) {
  RSD_CHECK(hipSetDevice(gpu_index));
  RSD_CHECK(hipHostMalloc(p, bytes ? bytes : 1,
                          write_combined ? hipHostMallocWriteCombined : hipHostMallocDefault));
  return 0;
}

void rsd_host_free(void *p) {
  if (p) (void)hipHostFree(p);
}

void rsd_destroy(rsd_ctx *d) {
  if (!d) return;
  (void)hipSetDevice(d->dev);
  if (d->stream) (void)hipStreamSynchronize(d->stream);
  for (int i = 0; i < RSD_MAX_PRODUCTS; ++i) {
    ProdPlan &pp = d->pp[i];
    if (pp.d_acc) (void)hipFree(pp.d_acc);
    if (pp.d_ics) (void)hipFree(pp.d_ics);
    if (pp.d_twf) (void)hipFree(pp.d_twf);
    if (pp.d_chirp) (void)hipFree(pp.d_chirp);
    if (pp.d_bsk) (void)hipFree(pp.d_bsk);
    if (pp.d_acc_int) (void)hipFree(pp.d_acc_int);
    if (pp.d_held) (void)hipFree(pp.d_held);
  }
  if (d->d_in) (void)hipFree(d->d_in);
  if (d->d_stage) (void)hipFree(d->d_stage);
  if (d->d_part) (void)hipFree(d->d_part);
  if (d->d_work) (void)hipFree(d->d_work);
  if (d->d_partb) (void)hipFree(d->d_partb);
  if (d->d_tw) (void)hipFree(d->d_tw);
  if (d->d_w) (void)hipFree(d->d_w);
  for (auto &r : d->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : d->pool) (void)hipEventDestroy(e);
  for (int k = 0; k < 2; ++k) {
    if (d->lane[k]) {
      (void)hipStreamSynchronize(d->lane[k]);
      (void)hipStreamDestroy(d->lane[k]);
    }
    if (d->ev_lane[k]) (void)hipEventDestroy(d->ev_lane[k]);
  }
  if (d->ev_main) (void)hipEventDestroy(d->ev_main);
  if (d->own_stream) {
    (void)hipStreamSynchronize(d->own_stream);
    (void)hipStreamDestroy(d->own_stream);
  }
  delete d;
}

// Channels per pass A/B launch pair for launches of kb buffers: as many as
// keep the chunk's intermediate within the product's fs_budget (at least one).
static unsigned fs_chunk_for(const rsd_ctx *d, const ProdPlan &pp, unsigned kb) {
  const size_t per_chan = (size_t)kb * pp.p.Nss * pp.p.Nt * d->cfg.Np * sizeof(float2);
  const size_t ch = pp.fs_budget / per_chan;
  return (unsigned)(ch < 1 ? 1 : (ch > d->cfg.Nc ? d->cfg.Nc : ch));
}

// Runs each dump group of a vector pass-B launch with na_buf spectra per
// group is split into, for cus compute units: enough for about kFsbWgPerCu
// workgroups per CU where the (k2 tile, group, channel) workgroups of all
// Nc channels are fewer, each run at least kFsbMinRun spectra.  A workgroup
// otherwise integrates its whole group in turn: a 2^13-point product over 8
// channels is 128 workgroups, each summing thousands of spectra.  The count
// depends on the context's channels, not on a launch's channel chunk, so
// every path over the same buffers (batched, held, allocation-free with
// its smaller chunks) adds the same runs in the same order.
static unsigned fsb_runs(const ProdPlan &pp, unsigned cus, unsigned Nc, unsigned na_buf) {
  const unsigned long long base =
      (unsigned long long)(pp.N2 / (unsigned)rsk::fs_tile_b((int)pp.N1)) * pp.Ng * Nc;
  const unsigned long long target = (unsigned long long)cus * kFsbWgPerCu;
  if (base == 0 || base >= target) return 1;
  unsigned long long ns = (target + base - 1) / base;
  const unsigned maxr = na_buf / kFsbMinRun;
  if (ns > maxr) ns = maxr;
  return ns < 1 ? 1u : (unsigned)ns;
}

// floats of partial sums a pass-B launch split into ns runs writes
static size_t fsb_part_floats(const ProdPlan &pp, unsigned ns, unsigned nch) {
  return ns < 2 ? 0 : (size_t)ns * pp.Ng * nch * pp.Q * pp.p.Nt;
}

// The most partial sums product pp's pass-B launches can need: a launch
// over all Nc channels with the most spectra per group any path can give
// it (a batch or held group of min(kMaxBatch, Nis) buffers; the run count
// grows with the spectra).  Launches over a channel chunk or fewer buffers
// need less, so whatever the chunking and batching (autotuned budgets, the
// allocation-free scratch) the run count is never cut by the buffer.
static size_t fsb_part_need(const rsd_ctx *d, const ProdPlan &pp) {
  if (pp.small) return 0;
  const unsigned kb = pp.p.Nis < (unsigned)rsk::kMaxBatch ? pp.p.Nis : (unsigned)rsk::kMaxBatch;
  const unsigned na = (kb > 1 ? kb : 1) * pp.na_buf;
  return fsb_part_floats(pp, fsb_runs(pp, kPlanCus, d->cfg.Nc, na), d->cfg.Nc);
}

// Launch plan of one product from its tunables; returns the float2
// elements of the four-step intermediate it needs (0 for small products).
static size_t plan_product(rsd_ctx *d, ProdPlan &pp) {
  const rsd_config *cfg = &d->cfg;
  const unsigned N = pp.p.Nt;
  size_t work = 0;
  if (pp.small) {
    plan_small(d, pp);
    plan_small_mfma(d, pp);
  } else {
    const unsigned L = ilog2u(N);
    pp.N1 = 1u << ((L + 1) / 2);
    pp.N2 = 1u << (L / 2);
    pp.na_buf = pp.p.Na < pp.p.Nss ? pp.p.Na : pp.p.Nss;
    pp.Ng = pp.p.Nss / pp.na_buf;
    pp.fs_chunk = fs_chunk_for(d, pp, 1);
    pp.hold_k = 0;
    if (pp.p.Nis > 1) {
      const size_t per_buf = (size_t)cfg->Nc * pp.p.Nss * N * cfg->Np * sizeof(float2);
      size_t k = kHoldBytes / per_buf;
      k = k > (size_t)rsk::kMaxBatch ? (size_t)rsk::kMaxBatch : k;
      k = k > pp.p.Nis ? pp.p.Nis : k;
      pp.hold_k = k >= 2 ? (unsigned)k : 0;
    }
    // batches of consecutive buffers only where an integration spans
    // several buffers (then each dump group is the whole buffer)
    pp.fs_bmax = 1;
    if (pp.p.Nis > 1) {
      const size_t per_buf = (size_t)pp.p.Nss * N * cfg->Np * sizeof(float2);
      size_t bm = pp.fs_budget / per_buf;
      bm = bm < 1 ? 1 : bm;
      bm = bm > (size_t)rsk::kMaxBatch ? (size_t)rsk::kMaxBatch : bm;
      pp.fs_bmax = (unsigned)(bm > pp.p.Nis ? pp.p.Nis : bm);
    }
    for (unsigned kb = 1; kb <= pp.fs_bmax; ++kb) {
      const size_t wl = (size_t)fs_chunk_for(d, pp, kb) * kb * pp.p.Nss * N * cfg->Np;
      if (wl > work) work = wl;
    }
  }
  return work;
}

// Sizes and launch plans for a configuration, without touching the device.
// Two products one kernel computes from a single read of the samples
// (k_small_mfma_u with its 8-point section): an 8-point product whose dumps
// integrate 128 spectra, i.e. one dump per 1024 samples and each dump
// within one buffer, and a 1024-point product of the same detection that
// the matrix-core engine runs on unsplit operands (int8, two pols).  The
// 8-point part computes in FP32 as the vector kernel does.  Used at engine
// level 1, where that product runs unsplit.
#ifndef RSD_FUSE8
#define RSD_FUSE8 1
#endif
static bool f8_pair(const rsd_ctx *d, const ProdPlan &a, const ProdPlan &b) {
  return d->cfg.Np == 2 && d->cfg.Nbps == 8 && a.small && !a.bs && a.p.Nt == 8 &&
         a.p.Na == 128 && a.p.Nis == 1 && a.p.Nds == b.p.Nss && b.small && !b.bs &&
         b.p.Nt == 1024 && b.m_ok && b.m_unsplit && a.p.npolout == b.p.npolout;
}
static bool f8_on(const rsd_ctx *d) { return d->f8_a >= 0 && d->mfma == 1; }

static int plan_all(const rsd_config *cfg, rsd_ctx *d) {
  d->dev = cfg->gpu_index;
  d->cfg = *cfg;
  d->cfg.weights = nullptr;
  if (cfg->ics && (cfg->Nant == 0 || cfg->Nc % cfg->Nant != 0)) {
    fprintf(stderr, "librawspec: incoherent sum needs Nc (%u) to be a multiple of Nant (%u)\n",
            cfg->Nc, cfg->Nant);
    fflush(stderr);
    return 1;
  }
  d->sample_bytes = (unsigned long long)cfg->Np * 2 * (cfg->Nbps / 8);
  d->chan_block_bytes = (unsigned long long)cfg->Ntpb * d->sample_bytes;
  d->chan_stride = d->chan_block_bytes * cfg->Nb;
  d->in_bytes = (size_t)(d->chan_stride * cfg->Nc);
  size_t part_max = 0, work_max = 0, partb_max = 0;
  for (unsigned i = 0; i < cfg->No; ++i) {
    ProdPlan &pp = d->pp[i];
    pp.p = cfg->prod[i];
    const unsigned N = pp.p.Nt;
    if (N == 0 || N > (1u << 24) || (!is_pow2(N) && N > kBsMaxN)) {
      fprintf(stderr,
              "librawspec: Nts[%u] = %u is not supported (powers of two up to 2^24, other "
              "lengths up to %u)\n", i, N, kBsMaxN);
      fflush(stderr);
      return 1;
    }
    pp.Q = pp.p.npolout == 1 ? 1 : 4;
    pp.bs = !is_pow2(N);
    pp.bs_m = N;
    if (pp.bs) {
      pp.bs_m = 1;
      while (pp.bs_m < 2 * N - 1) pp.bs_m <<= 1;
    }
    pp.small = pp.bs || N <= (unsigned)rsk::kTwN;
    pp.t_target = 2048;
    pp.t_maxspg = 16;
    pp.t_mtarget = 2048;
    pp.fs_budget = kFsWorkBytes;
    const size_t wl = plan_product(d, pp);
    if (wl > work_max) work_max = wl;
    const size_t pn = part_need(d, pp);
    if (pn > part_max) part_max = pn;
    const size_t pb = fsb_part_need(d, pp);
    if (pb > partb_max) partb_max = pb;
    pp.acc_len = (size_t)pp.Q * pp.p.Nds * cfg->Nc * N;
    pp.ics_len = cfg->ics ? pp.acc_len / cfg->Nant : 0;
  }
  d->part_len = part_max;
  d->work_len = work_max;
  d->partb_len = partb_max;
  d->f8_a = d->f8_b = -1;
  for (unsigned a = 0; RSD_FUSE8 && a < cfg->No && d->f8_a < 0; ++a)
    for (unsigned b = 0; b < cfg->No; ++b)
      if (f8_pair(d, d->pp[a], d->pp[b])) {
        d->f8_a = (int)a;
        d->f8_b = (int)b;
        break;
      }
  return 0;
}

// Device bytes the plan allocates.
static size_t plan_bytes(const rsd_ctx *d) {
  size_t b = d->in_bytes + (d->cfg.expand4 ? (size_t)(d->chan_block_bytes / 2) * d->cfg.Nc : 0);
  b += (size_t)rsk::kTwN * sizeof(float2) + (d->cfg.ics ? d->cfg.Nant * sizeof(float) : 0);
  size_t held = 0;
  for (unsigned i = 0; i < d->cfg.No; ++i) {
    const ProdPlan &pp = d->pp[i];
    b += pp.acc_len * sizeof(float) + pp.ics_len * sizeof(float);
    if (pp.bs) b += ((size_t)pp.p.Nt + pp.bs_m) * sizeof(float2);
    if (!pp.small) {
      b += (size_t)pp.N1 * sizeof(float2);
      if (pp.p.Nis > 1) b += pp.acc_len * sizeof(float);
      // held first-pass output, allocated at first use under the same cap
      // (hold_ready), with the default number of held buffers
      const size_t h = (size_t)pp.hold_k * d->cfg.Nc * pp.p.Nss * pp.p.Nt * d->cfg.Np * sizeof(float2);
      if (pp.hold_k >= 2 && held + h <= kHoldBytes) held += h;
    }
  }
  return b + held + d->part_len * sizeof(float) + d->work_len * sizeof(float2) +
         d->partb_len * sizeof(float);
}

static int create_impl(const rsd_config *cfg, rsd_ctx *d) {
  if (plan_all(cfg, d)) return 1;
  if (cfg->ics) d->weights.assign(cfg->weights, cfg->weights + cfg->Nant);
  RSD_CHECK(hipSetDevice(d->dev));
  if (scratch_setup(d->dev)) return 1;
  {
    int cus = 0;
    RSD_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, d->dev));
    d->cus = cus > 0 ? (unsigned)cus : 1u;
  }
  if (alloc_zero((void **)&d->d_in, d->in_bytes)) return 1;
  if (cfg->expand4 && alloc_zero((void **)&d->d_stage, (size_t)(d->chan_block_bytes / 2) * cfg->Nc))
    return 1;
  // master twiddle table W_4096^m
  {
    std::vector<float2> t = twiddles(rsk::kTwN, rsk::kTwN, 1);
    RSD_CHECK(hipMalloc(&d->d_tw, t.size() * sizeof(float2)));
    RSD_CHECK(hipMemcpy(d->d_tw, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
  }
  if (cfg->ics) {
    RSD_CHECK(hipMalloc(&d->d_w, cfg->Nant * sizeof(float)));
    RSD_CHECK(hipMemcpy(d->d_w, d->weights.data(), cfg->Nant * sizeof(float),
                        hipMemcpyHostToDevice));
  }
  for (unsigned i = 0; i < cfg->No; ++i) {
    ProdPlan &pp = d->pp[i];
    const unsigned N = pp.p.Nt;
    if (!pp.small) {
      std::vector<float2> f = twiddles(N, pp.N1, 1);
      RSD_CHECK(hipMalloc(&pp.d_twf, f.size() * sizeof(float2)));
      RSD_CHECK(hipMemcpy(pp.d_twf, f.data(), f.size() * sizeof(float2), hipMemcpyHostToDevice));
      if (pp.p.Nis > 1 && alloc_zero((void **)&pp.d_acc_int, pp.acc_len * sizeof(float))) return 1;
    }
    if (pp.bs) {
      std::vector<float2> ch, bk;
      bluestein_tables(N, pp.bs_m, ch, bk);
      RSD_CHECK(hipMalloc(&pp.d_chirp, ch.size() * sizeof(float2)));
      RSD_CHECK(hipMemcpy(pp.d_chirp, ch.data(), ch.size() * sizeof(float2), hipMemcpyHostToDevice));
      RSD_CHECK(hipMalloc(&pp.d_bsk, bk.size() * sizeof(float2)));
      RSD_CHECK(hipMemcpy(pp.d_bsk, bk.data(), bk.size() * sizeof(float2), hipMemcpyHostToDevice));
    }
    if (alloc_zero((void **)&pp.d_acc, pp.acc_len * sizeof(float))) return 1;
    if (cfg->ics && alloc_zero((void **)&pp.d_ics, pp.ics_len * sizeof(float))) return 1;
  }
  if (alloc_zero((void **)&d->d_part, d->part_len * sizeof(float))) return 1;
  if (alloc_zero((void **)&d->d_work, d->work_len * sizeof(float2))) return 1;
  if (d->partb_len && alloc_zero((void **)&d->d_partb, d->partb_len * sizeof(float))) return 1;
  RSD_CHECK(hipStreamCreateWithFlags(&d->own_stream, hipStreamNonBlocking));
  d->stream = d->own_stream;
  RSD_CHECK(hipEventCreateWithFlags(&d->ev_main, hipEventDisableTiming));
  for (int k = 0; k < 2; ++k) {
    RSD_CHECK(hipStreamCreateWithFlags(&d->lane[k], hipStreamNonBlocking));
    RSD_CHECK(hipEventCreateWithFlags(&d->ev_lane[k], hipEventDisableTiming));
  }
  // Lanes pay when a buffer is large enough for the overlap to outweigh the
  // extra stream synchronisation (two event waits per lane per call) and
  // when both lanes have work.
  bool fs = false, sm = false;
  for (unsigned i = 0; i < cfg->No; ++i) (d->pp[i].small ? sm : fs) = true;
  d->lanes = fs && sm && d->in_bytes >= kLaneMinBytes;
  return 0;
}

static rsd_ctx *new_ctx() {
  rsd_ctx *d = new rsd_ctx();
  memset(d->pp, 0, sizeof(d->pp));
  d->stream = d->own_stream = nullptr;
  d->lane[0] = d->lane[1] = nullptr;
  d->ev_main = d->ev_lane[0] = d->ev_lane[1] = nullptr;
  d->lanes = false;
  d->d_in = d->d_stage = nullptr;
  d->d_part = nullptr;
  d->d_work = d->d_tw = nullptr;
  d->d_partb = nullptr;
  d->partb_len = 0;
  d->hold_max = rsk::kMaxBatch;
  d->cus = 1;
  d->held_bytes = 0;
  d->d_w = nullptr;
  d->prof = false;
  d->mfma = 0;
  d->noalloc = false;
  return d;
}

int rsd_create(const rsd_config *cfg, rsd_ctx **out) {
  *out = nullptr;
  rsd_ctx *d = new_ctx();
  if (create_impl(cfg, d)) {
    rsd_destroy(d);
    return 1;
  }
  *out = d;
  return 0;
}

size_t rsd_plan_bytes(const rsd_config *cfg) {
  rsd_ctx *d = new_ctx();
  size_t b = plan_all(cfg, d) ? 0 : plan_bytes(d);
  delete d;
  return b;
}

int rsd_copy_block(rsd_ctx *d, const void *src, unsigned dblk) {
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipMemcpy2DAsync(d->d_in + dblk * d->chan_block_bytes, d->chan_stride, src,
                             d->chan_block_bytes, d->chan_block_bytes, d->cfg.Nc,
                             hipMemcpyHostToDevice, d->stream));
  RSD_CHECK(hipStreamSynchronize(d->stream));
  return 0;
}

int rsd_copy_block4(rsd_ctx *d, const void *src, unsigned dblk) {
  RSD_CHECK(hipSetDevice(d->dev));
  if (!d->d_stage) return 1;
  const unsigned long long cb4 = d->chan_block_bytes / 2;   // complex4 bytes per channel
  RSD_CHECK(hipMemcpyAsync(d->d_stage, src, cb4 * d->cfg.Nc, hipMemcpyHostToDevice, d->stream));
  if (rsk::launch_expand4(d->d_stage, d->d_in, d->cfg.Nc, cb4, d->chan_stride,
                          dblk * d->chan_block_bytes, d->stream))
    return 1;
  RSD_CHECK(hipStreamSynchronize(d->stream));
  return 0;
}

int rsd_zero_block(rsd_ctx *d, unsigned dblk) {
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipMemset2DAsync(d->d_in + dblk * d->chan_block_bytes, d->chan_stride, 0,
                             d->chan_block_bytes, d->cfg.Nc, d->stream));
  RSD_CHECK(hipStreamSynchronize(d->stream));
  return 0;
}

static int run_fourstep(rsd_ctx *d, int i, const uint8_t *const *ins, unsigned n, float isign,
                        float gam, float scale, int first, int last);
static int hold_ready(rsd_ctx *d, ProdPlan &pp);
static int run_fourstep_held(rsd_ctx *d, int i, const uint8_t *in, float isign, float gam,
                             float scale, int first, int last);

// Whether small product pp runs on the matrix cores under the current
// engine.  Level 2 runs every matrix-core kernel; level 1 only the unsplit
// one: the split kernel is never faster than the vector kernel (0.204
// against 0.192 ms per 256 MiB buffer of 64 channels; 2.6 times slower on
// 4 MiB buffers of 8 channels), which meets the tighter FP32 bar as well.
static bool small_on_mfma(const rsd_ctx *d, const ProdPlan &pp) {
  return pp.m_ok && (d->mfma == 2 || (d->mfma == 1 && pp.m_unsplit));
}

// Small product i over one buffer: its one kernel (transform, detection,
// integration) and, when a dump group is split over workgroups, the sum of
// their partial sums.  f8_out: the output of the 8-point product of a
// fused pair, which the unsplit matrix-core kernel then computes as well.
static int run_small(rsd_ctx *d, int i, const uint8_t *in, float isign, float gam, float scale,
                     int first, float *f8_out) {
  ProdPlan &pp = d->pp[i];
  const rsd_config &c = d->cfg;
  const hipStream_t s = prod_stream(d, pp);
  rsk::SmallArgs a;
  a.in = in;
  a.chan_stride = d->chan_stride;
  a.Nc = c.Nc;
  a.na_buf = pp.na_buf;
  a.chunk = pp.chunk;
  a.G = pp.G;
  a.Ng = pp.Ng;
  a.gpw = pp.gpw;
  a.spg = pp.spg;
  a.bpc = pp.bpc;
  a.isign = isign;
  a.gam = gam;
  a.scale = scale;
  a.direct = pp.direct;
  a.first = first;
  a.acc = pp.d_acc;
  a.part = d->d_part;
  a.tw = d->d_tw;
  a.nt = pp.p.Nt;
  a.chirp = pp.d_chirp;
  a.bsk = pp.d_bsk;
  a.f8_out = f8_out;
  // Both engines write the same (host) layout, so a product can change
  // engine between any two buffers.
  const bool mf = small_on_mfma(d, pp);
  if (f8_out && !(mf && d->mfma == 1)) return 1;
  if (mf) {
    a.chunk = pp.m_chunk;
    a.G = pp.m_G;
    a.gpw = 1;
    a.direct = pp.m_G == 1;
    const unsigned long long items = (unsigned long long)c.Nc * pp.Ng * pp.m_G;
    if (items > 0x7fffffffull) return 1;
    if (d->mfma == 1 && pp.m_unsplit) {
      ProfScope ps(d, i, f8_out ? "k_small_mfma_u8" : "k_small_mfma_u", pp.p.Nt, s);
      if (rsk::launch_small_mfma_u((int)pp.p.Nt, (int)c.Nbps, (int)c.Np, pp.p.npolout, a,
                                   (unsigned)items, s))
        return 1;
    } else {
      ProfScope ps(d, i, "k_small_mfma", pp.p.Nt, s);
      if (rsk::launch_small_mfma((int)pp.p.Nt, (int)c.Nbps, (int)c.Np, pp.p.npolout, a,
                                 (unsigned)items, s))
        return 1;
    }
  } else {
    const unsigned long long nb = (unsigned long long)pp.bpc * c.Nc;
    if (nb > 0x7fffffffull) return 1;
    ProfScope ps(d, i, "k_small", pp.p.Nt, s);
    if (rsk::launch_small((int)pp.bs_m, (int)c.Nbps, (int)c.Np, pp.p.npolout, a, (unsigned)nb,
                          s))
      return 1;
  }
  if (!a.direct) {
    rsk::ReduceArgs r;
    r.part = d->d_part;
    r.acc = pp.d_acc;
    r.Nc = c.Nc;
    r.Ng = pp.Ng;
    r.G = a.G;
    r.Q = (unsigned)pp.Q;
    r.N = pp.p.Nt;
    r.scale = scale;
    r.first = first;
    ProfScope ps(d, i, "k_reduce", pp.p.Nt, s);
    if (rsk::launch_reduce(r, s)) return 1;
  }
  return 0;
}

// Per-call constants: transform direction, the sign of the imaginary cross
// term (input_conjugated flips it, rawspec_gpu.cu:1120-1123, 1139-1142; the
// inverse-via-conjugation trick flips it once more) and the SNORM scale.
struct CallSigns {
  float isign, gam, scale;
  CallSigns(const rsd_config &c, int inverse)
      : isign(inverse ? -1.0f : 1.0f),
        gam((c.conj ? -1.0f : 1.0f) * (inverse ? -1.0f : 1.0f)),
        scale(1.0f / ((c.Nbps == 16 ? 32767.0f : 127.0f) * (c.Nbps == 16 ? 32767.0f : 127.0f))) {}
};

int rsd_run_product(rsd_ctx *d, int i, const void *d_in, int inverse, int first, int last) {
  RSD_CHECK(hipSetDevice(d->dev));
  ProdPlan &pp = d->pp[i];
  const uint8_t *in = d_in ? (const uint8_t *)d_in : d->d_in;
  const CallSigns k(d->cfg, inverse);
  if (pp.small) return run_small(d, i, in, k.isign, k.gam, k.scale, first, nullptr);
  if (first) pp.holding = hold_ready(d, pp);
  if (pp.holding) return run_fourstep_held(d, i, in, k.isign, k.gam, k.scale, first, last);
  return run_fourstep(d, i, &in, 1, k.isign, k.gam, k.scale, first, last);
}

int rsd_fused_pair(rsd_ctx *d, int *a, int *b) {
  if (!f8_on(d)) return 0;
  *a = d->f8_a;
  *b = d->f8_b;
  return 1;
}

int rsd_run_fused(rsd_ctx *d, const void *d_in, int inverse, int first_b) {
  RSD_CHECK(hipSetDevice(d->dev));
  if (!f8_on(d)) return 1;
  const uint8_t *in = d_in ? (const uint8_t *)d_in : d->d_in;
  const CallSigns k(d->cfg, inverse);
  return run_small(d, d->f8_b, in, k.isign, k.gam, k.scale, first_b, d->pp[d->f8_a].d_acc);
}

// Four-step product i over n consecutive buffers of one integration: launch
// pairs of up to fs_bmax buffers (batched: pass B sums their spectra before
// touching the running sum), each over channel chunks sized so the chunk's Y
// stays within its budget.
// Arguments shared by every pass A/B launch of product i; latches the
// integration's pass-B engine when it starts.
// Lowest engine level that runs the four-step pass A on the matrix cores
// where a kernel exists.  At level 1 the vector pass A is the faster one
// (0.300 against 0.360 ms per 256 MiB buffer of the 2^20 product, both
// storing 24-bit Y), so only level 2, the matrix cores everywhere, uses it.
#ifndef RSD_FSA_MFMA_LEVEL
#define RSD_FSA_MFMA_LEVEL 2
#endif

static void fs_args(rsd_ctx *d, ProdPlan &pp, rsk::FsArgs &a, float isign, float gam, float scale,
                    int first) {
  const rsd_config &c = d->cfg;
  const float full = c.Nbps == 16 ? 32767.0f : 127.0f;
  a.chan_stride = d->chan_stride;
  a.Nc = c.Nc;
  a.Nss = pp.p.Nss;
  a.N1 = pp.N1;
  a.N2 = pp.N2;
  a.isign = isign;
  a.gam = gam;
  a.scale = scale;
  a.work = d->d_work;
  a.acc = pp.d_acc;
  a.acc_int = pp.d_acc_int;
  a.tw = d->d_tw;
  a.twf = pp.d_twf;
  // The engine is chosen when an integration starts and kept until it
  // completes (the two pass-B kernels lay out the internal accumulator
  // differently).
  // Level 1 keeps pass B on the vector units: the matrix-core pass B must
  // split every operand (the tone of a four-step product sits coherently in
  // a row of the intermediate, so unsplit FP16 would miss the 1e-3 bar) and
  // is then slower than the vector pass B (0.358 against 0.256 ms per 256 MiB
  // buffer of the 64-channel 2^20 product); level 2 runs it.
  if (first) pp.fs_mfma = d->mfma == 2 && rsk::fs_b_mfma_ok((int)pp.N1, (int)c.Np);
  // Level 1 also keeps the intermediate in 24-bit floats (16 significant
  // bits, each value its own exponent): a quarter less of the traffic that
  // bounds both passes, for at most about 2e-5 of added error at one
  // spectrum per dump (numpy model of the 2^20 product with a strong tone;
  // it falls as the integration grows), fifty times under the 1e-3 bar.
  if (first) pp.fs_y24 = d->mfma == 1 && c.Np == 2 && !pp.fs_mfma;
  a.y24 = pp.fs_y24;
  // Pass A on the matrix cores where a kernel exists (int8, two pols,
  // N2 = 1024), from level RSD_FSA_MFMA_LEVEL up: split operands keep its
  // error at the FP32 level.
  if (first)
    pp.fs_amfma = d->mfma != 0 && d->mfma >= RSD_FSA_MFMA_LEVEL &&
                  rsk::fs_a_mfma_ok((int)pp.N1, (int)pp.N2, (int)c.Nbps, (int)c.Np);
  // both pass-B kernels read Y tiled by fs_tile_b(N1) = 4 consecutive k2
  // (for N1 = 1024, the matrix-core pass B's workgroup of 4 columns)
  a.tby = (unsigned)rsk::fs_tile_b((int)pp.N1);
  {
    // largest |Y| is N2 * full * sqrt(2); scale it to at most 2^15
    int e = 0;
    while (ldexp((double)pp.N2 * full * 1.4142135623730951, -e) > 32768.0) ++e;
    a.mscale = (float)ldexp(1.0, -e);
  }
  a.yb0 = 0;
  a.grid_a = d->cus * RSD_FSA_WG_PER_CU;
  a.ns = 1;                  // pass B's runs are chosen per launch
  a.sp_run = 0;
  a.partb = nullptr;
}

// Matrix-core pass-A workgroups per CU (all resident: two per CU at its
// register and LDS footprint), each looping over its share of the items.
#ifndef RSD_FSA_MFMA_WG_PER_CU
#define RSD_FSA_MFMA_WG_PER_CU 2
#endif

static int launch_pass_a(rsd_ctx *d, int i, rsk::FsArgs a, hipStream_t s) {
  ProdPlan &pp = d->pp[i];
  if (pp.fs_amfma) {
    a.grid_a = d->cus * RSD_FSA_MFMA_WG_PER_CU;
    ProfScope ps(d, i, "k_fs_a_mfma", pp.p.Nt, s);
    return rsk::launch_fs_a_mfma((int)d->cfg.Nbps, (int)d->cfg.Np, a, s);
  }
  ProfScope ps(d, i, "k_fs_a", pp.p.Nt, s);
  return rsk::launch_fs_a((int)d->cfg.Nbps, (int)d->cfg.Np, a, s);
}

static int launch_pass_b(rsd_ctx *d, int i, rsk::FsArgs a, hipStream_t s) {
  ProdPlan &pp = d->pp[i];
  a.ns = 1;
  a.sp_run = a.na_buf;
  a.partb = nullptr;
  if (pp.fs_mfma) {
    ProfScope ps(d, i, "k_fs_b_mfma", pp.p.Nt, s);
    return rsk::launch_fs_b_mfma((int)d->cfg.Np, pp.p.npolout, a, s);
  }
  // split the dump groups into runs where the launch is too narrow for the
  // GPU, as far as the partial-sum buffer allows
  unsigned ns = fsb_runs(pp, d->cus < kPlanCus ? d->cus : kPlanCus, d->cfg.Nc, a.na_buf);
  const size_t per_run = fsb_part_floats(pp, 2, a.nch) / 2;
  if (ns > 1 && per_run && (size_t)ns * per_run > d->partb_len) ns = (unsigned)(d->partb_len / per_run);
  if (ns > 1) {
    a.sp_run = (a.na_buf + ns - 1) / ns;
    a.ns = (a.na_buf + a.sp_run - 1) / a.sp_run;   // no empty run
    a.partb = d->d_partb;
  }
  {
    ProfScope ps(d, i, "k_fs_b", pp.p.Nt, s);
    if (int rc = rsk::launch_fs_b((int)d->cfg.Np, pp.p.npolout, a, s)) return rc;
  }
  if (a.ns < 2) return 0;
  ProfScope ps(d, i, "k_fs_b_reduce", pp.p.Nt, s);
  return rsk::launch_fs_b_reduce(pp.p.npolout, a, s);
}

// Pass B over the held buffers of product i, which complete the integration
// when `last`.
static int flush_held(rsd_ctx *d, int i, rsk::FsArgs &a, int last) {
  ProdPlan &pp = d->pp[i];
  if (pp.held == 0) return 0;
  a.work = pp.d_held;
  a.ynb = pp.hold_k;
  a.c0 = 0;
  a.nch = d->cfg.Nc;
  a.nbat = pp.held;
  a.na_buf = pp.held * pp.na_buf;
  a.Ng = pp.Ng;
  if (pp.held_first) a.mode = last ? rsk::FS_HOST : rsk::FS_FIRST;
  else a.mode = last ? rsk::FS_LAST : rsk::FS_MID;
  pp.held = 0;
  pp.held_first = 0;
  return launch_pass_b(d, i, a, prod_stream(d, pp));
}

// Whether product i holds first-pass output through the integration that
// starts now; allocates the product's slots on first use (not retried after
// a refusal).  Latches the number of slots used.
static int hold_ready(rsd_ctx *d, ProdPlan &pp) {
  const unsigned k = pp.hold_k < d->hold_max ? pp.hold_k : d->hold_max;
  pp.held = 0;
  if (k < 2 || pp.hold_failed) return 0;
  if (!pp.d_held && d->noalloc) return 0;   // holds only where the store exists
  if (!pp.d_held) {
    const size_t bytes = (size_t)pp.hold_k * d->cfg.Nc * pp.p.Nss * pp.p.Nt * d->cfg.Np *
                         sizeof(float2);
    if (d->held_bytes + bytes > kHoldBytes || hipMalloc(&pp.d_held, bytes) != hipSuccess) {
      (void)hipGetLastError();
      pp.d_held = nullptr;
      pp.hold_failed = 1;
      return 0;
    }
    d->held_bytes += bytes;
  }
  pp.hold_use = k;
  return 1;
}

// One buffer of product i while it holds first-pass output: pass A into the
// next slot, pass B over the slots when hold_use are full or the integration
// completes.
static int run_fourstep_held(rsd_ctx *d, int i, const uint8_t *in, float isign, float gam,
                             float scale, int first, int last) {
  ProdPlan &pp = d->pp[i];
  const hipStream_t s = prod_stream(d, pp);
  rsk::FsArgs a;
  fs_args(d, pp, a, isign, gam, scale, first);
  if (first) {
    pp.held = 0;
    pp.held_first = 1;
  }
  a.work = pp.d_held;
  a.ynb = pp.hold_k;
  a.yb0 = pp.held;
  a.c0 = 0;
  a.nch = d->cfg.Nc;
  a.nbat = 1;
  for (unsigned b = 0; b < (unsigned)rsk::kMaxBatch; ++b) a.in[b] = in;
  a.na_buf = pp.na_buf;
  a.Ng = pp.Ng;
  a.mode = rsk::FS_MID;      // pass A does not read it
  if (launch_pass_a(d, i, a, s)) return 1;
  ++pp.held;
  if (!last && pp.held < pp.hold_use) return 0;
  return flush_held(d, i, a, last);
}

static int run_fourstep(rsd_ctx *d, int i, const uint8_t *const *ins, unsigned n, float isign,
                        float gam, float scale, int first, int last) {
  ProdPlan &pp = d->pp[i];
  const rsd_config &c = d->cfg;
  const hipStream_t s = prod_stream(d, pp);
  rsk::FsArgs a;
  fs_args(d, pp, a, isign, gam, scale, first);
  // buffers held by earlier one-buffer calls of this integration go first
  if (first) pp.held = 0;
  if (pp.held && flush_held(d, i, a, 0)) return 1;
  a.work = d->d_work;
  for (unsigned j = 0; j < n;) {
    const unsigned kb = n - j < pp.fs_bmax ? n - j : pp.fs_bmax;
    const bool f = first && j == 0, l = last && j + kb == n;
    if (pp.p.Nis == 1 || (f && l)) a.mode = rsk::FS_HOST;
    else if (f) a.mode = rsk::FS_FIRST;
    else if (l) a.mode = rsk::FS_LAST;
    else a.mode = rsk::FS_MID;
    a.nbat = kb;
    a.ynb = kb;
    for (unsigned b = 0; b < (unsigned)rsk::kMaxBatch; ++b) a.in[b] = ins[j + (b < kb ? b : 0)];
    a.na_buf = kb * pp.na_buf;
    a.Ng = pp.Ng;
    const unsigned chc = fs_chunk_for(d, pp, kb);
    // pass A and pass B alternate per channel chunk
    for (unsigned c0 = 0; c0 < c.Nc; c0 += chc) {
      a.c0 = c0;
      a.nch = c.Nc - c0 < chc ? c.Nc - c0 : chc;
      if (launch_pass_a(d, i, a, s)) return 1;
      if (launch_pass_b(d, i, a, s)) return 1;
    }
    j += kb;
  }
  return 0;
}

int rsd_run_product_batch(rsd_ctx *d, int i, const void *const *d_in, unsigned n, int inverse,
                          int first, int last) {
  RSD_CHECK(hipSetDevice(d->dev));
  ProdPlan &pp = d->pp[i];
  if (n == 0) return 1;
  if (pp.small || pp.fs_bmax <= 1) {
    for (unsigned j = 0; j < n; ++j)
      if (rsd_run_product(d, i, d_in[j], inverse, first && j == 0, last && j == n - 1)) return 1;
    return 0;
  }
  const CallSigns k(d->cfg, inverse);
  // an integration started by a batch is batched throughout; one-buffer
  // calls continuing it run pass A/B per buffer
  if (first) pp.holding = 0;
  return run_fourstep(d, i, reinterpret_cast<const uint8_t *const *>(d_in), n, k.isign, k.gam,
                      k.scale, first, last);
}

// ---- autotuning of the launch plans -------------------------------------
// Grows the partial-sum and intermediate buffers to what product i's
// current plan needs (the device is idle while tuning).
static int fit_buffers(rsd_ctx *d, ProdPlan &pp, size_t work) {
  const size_t pn = part_need(d, pp);
  if (pn > d->part_len) {
    if (d->d_part) RSD_CHECK(hipFree(d->d_part));
    d->d_part = nullptr;
    if (alloc_zero((void **)&d->d_part, pn * sizeof(float))) return 1;
    d->part_len = pn;
  }
  if (work > d->work_len) {
    if (d->d_work) RSD_CHECK(hipFree(d->d_work));
    d->d_work = nullptr;
    if (alloc_zero((void **)&d->d_work, work * sizeof(float2))) return 1;
    d->work_len = work;
  }
  return 0;
}

namespace {
struct Tunables { unsigned target, maxspg, mtarget; size_t budget; };
}

static void set_tunables(ProdPlan &pp, const Tunables &t) {
  pp.t_target = t.target;
  pp.t_maxspg = t.maxspg;
  pp.t_mtarget = t.mtarget;
  pp.fs_budget = t.budget;
}

// Candidate tunables of product i, the defaults first.  Small products: the
// workgroups a launch aims for and the slots per dump group (on the
// matrix-core engine, the wavefronts a launch aims for); four-step products:
// the intermediate's budget per launch pair.
static std::vector<Tunables> tune_cands(const rsd_ctx *d, const ProdPlan &pp) {
  const Tunables dflt = {2048u, 16u, 2048u, kFsWorkBytes};
  std::vector<Tunables> c = {dflt};
  if (pp.small && small_on_mfma(d, pp)) {
    for (unsigned tg : {512u, 1024u, 4096u, 8192u}) c.push_back({dflt.target, dflt.maxspg, tg, dflt.budget});
  } else if (pp.small) {
    for (unsigned tg : {512u, 1024u, 2048u, 4096u, 8192u})
      for (unsigned sg : {2u, 4u, 8u, 16u, 32u})
        c.push_back({tg, sg, dflt.mtarget, dflt.budget});
  } else {
    for (size_t mb : {256, 512, 1024, 2048, 3072}) c.push_back({dflt.target, dflt.maxspg, dflt.mtarget, mb << 20});
  }
  return c;
}

// The launch plan the current tunables produce, for skipping duplicates.
static std::vector<unsigned long long> plan_key(const ProdPlan &pp) {
  if (pp.small) return {pp.chunk, pp.G, pp.spg, pp.gpw, pp.m_chunk};
  return {pp.fs_chunk, pp.fs_bmax};
}

int rsd_tune_count(rsd_ctx *d, int i) { return (int)tune_cands(d, d->pp[i]).size(); }

int rsd_tune_apply(rsd_ctx *d, int i, int k) {
  ProdPlan &pp = d->pp[i];
  const std::vector<Tunables> c = tune_cands(d, pp);
  if (k < 0 || k >= (int)c.size()) return 1;
  // a candidate whose plan equals an earlier candidate's is a duplicate
  std::vector<std::vector<unsigned long long>> seen;
  for (int j = 0; j < k; ++j) {
    set_tunables(pp, c[j]);
    (void)plan_product(d, pp);
    seen.push_back(plan_key(pp));
  }
  set_tunables(pp, c[k]);
  const size_t work = plan_product(d, pp);
  if (std::find(seen.begin(), seen.end(), plan_key(pp)) != seen.end()) return 2;
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipStreamSynchronize(d->stream));
  return fit_buffers(d, pp, work);
}

int rsd_noalloc_begin(rsd_ctx *d) {
  if (d->noalloc) return 1;
  void *base = (d->dev >= 0 && d->dev < kScratchDevices) ? g_scratch_ptr[d->dev] : nullptr;
  if (!base) {
    fprintf(stderr, "librawspec: no allocation-free scratch on device %d\n", d->dev);
    fflush(stderr);
    return 1;
  }
  // the small products' partial sums, pass B's partial sums (the same
  // capacity as the context's own, so pass B splits its groups the same
  // way), then the four-step intermediate
  size_t part = 0;
  for (unsigned i = 0; i < d->cfg.No; ++i) {
    const size_t pn = part_need(d, d->pp[i]);
    if (pn > part) part = pn;
  }
  const size_t part_bytes = (part * sizeof(float) + 255) / 256 * 256;
  const size_t partb_bytes = (d->partb_len * sizeof(float) + 255) / 256 * 256;
  if (part_bytes + partb_bytes > kScratchBytes) {
    fprintf(stderr,
            "librawspec: allocation-free processing needs %zu MiB of partial sums; the scratch "
            "holds %zu MiB\n", (part_bytes + partb_bytes + (1u << 20) - 1) >> 20, kScratchBytes >> 20);
    fflush(stderr);
    return 1;
  }
  const size_t work_bytes = kScratchBytes - part_bytes - partb_bytes;
  for (unsigned i = 0; i < d->cfg.No; ++i) {
    const ProdPlan &pp = d->pp[i];
    if (pp.small) continue;
    const size_t per_chan = (size_t)pp.p.Nss * pp.p.Nt * d->cfg.Np * sizeof(float2);
    if (per_chan > work_bytes) {
      fprintf(stderr,
              "librawspec: allocation-free processing needs %zu MiB of intermediate for one "
              "channel of product %u; the scratch holds %zu MiB beside %zu MiB of partial sums\n",
              per_chan >> 20, i, kScratchBytes >> 20, (part_bytes + partb_bytes) >> 20);
      fflush(stderr);
      return 1;
    }
  }
  d->sv_part = d->d_part;
  d->sv_part_len = d->part_len;
  d->sv_work = d->d_work;
  d->sv_work_len = d->work_len;
  d->sv_partb = d->d_partb;
  d->sv_partb_len = d->partb_len;
  d->d_partb = d->partb_len ? reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(base) +
                                                        part_bytes)
                            : nullptr;
  d->d_part = reinterpret_cast<float *>(base);
  d->part_len = part;
  d->d_work = reinterpret_cast<float2 *>(reinterpret_cast<unsigned char *>(base) + part_bytes +
                                         partb_bytes);
  d->work_len = work_bytes / sizeof(float2);
  for (unsigned i = 0; i < d->cfg.No; ++i) {
    ProdPlan &pp = d->pp[i];
    d->sv_budget[i] = pp.fs_budget;
    if (!pp.small && pp.fs_budget > work_bytes) pp.fs_budget = work_bytes;
  }
  d->noalloc = true;
  return 0;
}

void rsd_noalloc_end(rsd_ctx *d) {
  if (!d->noalloc) return;
  d->d_part = d->sv_part;
  d->part_len = d->sv_part_len;
  d->d_work = d->sv_work;
  d->work_len = d->sv_work_len;
  d->d_partb = d->sv_partb;
  d->partb_len = d->sv_partb_len;
  for (unsigned i = 0; i < d->cfg.No; ++i) d->pp[i].fs_budget = d->sv_budget[i];
  d->noalloc = false;
}

int rsd_tune_fit(rsd_ctx *d) {
  // what the products' current plans need, at most what is allocated
  size_t part = 0, work = 0;
  for (unsigned i = 0; i < d->cfg.No; ++i) {
    ProdPlan &pp = d->pp[i];
    const size_t w = plan_product(d, pp);
    if (w > work) work = w;
    const size_t pn = part_need(d, pp);
    if (pn > part) part = pn;
  }
  if (part >= d->part_len && work >= d->work_len) return 0;
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipStreamSynchronize(d->stream));
  if (part < d->part_len) {
    if (d->d_part) RSD_CHECK(hipFree(d->d_part));
    d->d_part = nullptr;
    d->part_len = 0;
    if (alloc_zero((void **)&d->d_part, part * sizeof(float))) return 1;
    d->part_len = part;
  }
  if (work < d->work_len) {
    if (d->d_work) RSD_CHECK(hipFree(d->d_work));
    d->d_work = nullptr;
    d->work_len = 0;
    if (alloc_zero((void **)&d->d_work, work * sizeof(float2))) return 1;
    d->work_len = work;
  }
  return 0;
}

int rsd_tune_describe(rsd_ctx *d, int i, char *buf, size_t len) {
  const ProdPlan &pp = d->pp[i];
  if (pp.small)
    return snprintf(buf, len, "small target %u spg %u mtarget %u (chunk %u, %u per group, %u groups per workgroup)",
                    pp.t_target, pp.t_maxspg, pp.t_mtarget, pp.chunk, pp.spg, pp.gpw);
  return snprintf(buf, len, "fourstep budget %zu MiB (chunk %u channels, batch %u)", pp.fs_budget >> 20,
                  pp.fs_chunk, pp.fs_bmax);
}

int rsd_lanes_begin(rsd_ctx *d) {
  if (!use_lanes(d)) return 0;
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipEventRecord(d->ev_main, d->stream));
  for (int k = 0; k < 2; ++k) RSD_CHECK(hipStreamWaitEvent(d->lane[k], d->ev_main, 0));
  return 0;
}

int rsd_lane_end(rsd_ctx *d, int i) {
  if (!use_lanes(d)) return 0;
  RSD_CHECK(hipSetDevice(d->dev));
  for (int k = 0; k < 2; ++k) {
    if (i >= 0 && k != (d->pp[i].small ? 1 : 0)) continue;
    RSD_CHECK(hipEventRecord(d->ev_lane[k], d->lane[k]));
    RSD_CHECK(hipStreamWaitEvent(d->stream, d->ev_lane[k], 0));
  }
  return 0;
}

int rsd_lane_resume(rsd_ctx *d, int i) {
  if (!use_lanes(d)) return 0;
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipEventRecord(d->ev_main, d->stream));
  RSD_CHECK(hipStreamWaitEvent(d->lane[d->pp[i].small ? 1 : 0], d->ev_main, 0));
  return 0;
}

int rsd_run_ics(rsd_ctx *d, int i) {
  RSD_CHECK(hipSetDevice(d->dev));
  ProdPlan &pp = d->pp[i];
  if (!pp.d_ics) return 1;
  rsk::IcsArgs a;
  a.acc = pp.d_acc;
  a.ics = pp.d_ics;
  a.w = d->d_w;
  a.Nc = d->cfg.Nc;
  a.Nant = d->cfg.Nant;
  a.Ncpa = d->cfg.Nc / d->cfg.Nant;
  a.Q = (unsigned)pp.Q;
  a.N = pp.p.Nt;
  a.Nds = pp.p.Nds;
  ProfScope ps(d, i, "k_ics", pp.p.Nt);
  return rsk::launch_ics(a, d->stream);
}

int rsd_dump_to_host(rsd_ctx *d, int i, float *h_pwr, size_t pwr_bytes, float *h_ics,
                     size_t ics_bytes) {
  RSD_CHECK(hipSetDevice(d->dev));
  ProdPlan &pp = d->pp[i];
  if (pwr_bytes > pp.acc_len * sizeof(float)) return 1;
  RSD_CHECK(hipMemcpyAsync(h_pwr, pp.d_acc, pwr_bytes, hipMemcpyDeviceToHost, d->stream));
  if (h_ics && pp.d_ics) {
    if (ics_bytes > pp.ics_len * sizeof(float)) return 1;
    RSD_CHECK(hipMemcpyAsync(h_ics, pp.d_ics, ics_bytes, hipMemcpyDeviceToHost, d->stream));
  }
  return 0;
}

int rsd_host_fn(rsd_ctx *d, void (*fn)(void *), void *arg) {
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipLaunchHostFunc(d->stream, fn, arg));
  return 0;
}

int rsd_query(rsd_ctx *d) {
  if (hipSetDevice(d->dev) != hipSuccess) return -1;
  hipError_t e = hipStreamQuery(d->stream);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return -1;
}

int rsd_sync(rsd_ctx *d) {
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipStreamSynchronize(d->stream));
  return 0;
}

void *rsd_input_ptr(rsd_ctx *d) { return d->d_in; }
size_t rsd_input_bytes(rsd_ctx *d) { return d->in_bytes; }
float *rsd_pwr_ptr(rsd_ctx *d, int i) { return d->pp[i].d_acc; }
float *rsd_ics_ptr(rsd_ctx *d, int i) { return d->pp[i].d_ics; }
void *rsd_stream(rsd_ctx *d) { return (void *)d->stream; }

int rsd_set_stream(rsd_ctx *d, void *stream) {
  hipStream_t s = stream ? (hipStream_t)stream : d->own_stream;
  if (s == d->stream) return 0;
  RSD_CHECK(hipSetDevice(d->dev));
  // work already queued on the old stream precedes everything queued on the
  // new one (the integration buffers carry over)
  RSD_CHECK(hipEventRecord(d->ev_main, d->stream));
  RSD_CHECK(hipStreamWaitEvent(s, d->ev_main, 0));
  d->stream = s;
  return 0;
}

int rsd_profile(rsd_ctx *d, int enable) {
  d->prof = enable != 0;
  return 0;
}

int rsd_set_mfma(rsd_ctx *d, int level) {
  if (level < 0 || level > 2) return 1;
  d->mfma = level;
  return 0;
}

int rsd_get_mfma(rsd_ctx *d) { return d->mfma; }

int rsd_set_hold(rsd_ctx *d, unsigned max_buffers) {
  d->hold_max = max_buffers;
  return 0;
}

unsigned rsd_get_hold(rsd_ctx *d) { return d->hold_max; }

int rsd_set_lanes(rsd_ctx *d, int enable) {
  if (enable && !(d->lane[0] && d->lane[1])) return 1;
  d->lanes = enable != 0;
  return 0;
}

int rsd_get_lanes(rsd_ctx *d) { return d->lanes ? 1 : 0; }

int rsd_profile_report(rsd_ctx *d, char *buf, size_t len) {
  RSD_CHECK(hipSetDevice(d->dev));
  RSD_CHECK(hipStreamSynchronize(d->stream));
  std::map<std::string, std::pair<long, double>> agg;
  for (auto &r : d->recs) {
    float ms = 0.0f;
    RSD_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
    auto &e = agg[r.name];
    e.first += 1;
    e.second += ms;
    d->pool.push_back(r.a);
    d->pool.push_back(r.b);
  }
  d->recs.clear();
  size_t off = 0;
  if (len) buf[0] = 0;
  for (auto &kv : agg) {
    int n = snprintf(buf + off, len > off ? len - off : 0, "%s %ld %.6f\n", kv.first.c_str(),
                     kv.second.first, kv.second.second);
    if (n < 0) return -1;
    off += (size_t)n;
    if (off >= len) return -1;
  }
  return (int)off;
}

}  // extern "C"
