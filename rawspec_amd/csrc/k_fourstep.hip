// k_fourstep.hip -- spectrometer kernels for FFT lengths 8192 .. 2^24.
//
// Bailey's four-step decomposition N = N1*N2 (n = n1 + N1*n2,
// k = k2 + N2*k1):
//   pass A  Y[n1][k2] = W_N^(n1*k2) * sum_n2 x[n1 + N1*n2] W_N2^(n2*k2)
//   pass B  X[k2 + N2*k1] = sum_n1 Y[n1][k2] W_N1^(n1*k1)
// Pass A fuses the int8/int16 unpack into its first radix stage; pass B fuses
// detection, integration of the dump group's spectra and the FFT-shift
// (k -> k + N/2 becomes k1 -> k1 + N1/2).  Together they replace the
// reference's multi-pass cuFFT plan with load/store callbacks
// (rawspec_gpu.cu:133-267, 1182-1195, 1555-1569) plus its accumulate kernel
// and shift copies (:297-318, :1647-1697).
//
// Memory plan: the host launches A and B per chunk of channels whose Y fits
// the work buffer (the whole product for the usual shapes: launches over
// fewer channels leave the device underfilled, which costs more than the
// Infinity Cache saves).  Y is tiled [k2 / TB][n1][k2 % TB][pol] so that a pass-B wave reads
// 1 KB contiguous per load and a pass-A wave writes 256-byte runs.  For
// integrations spanning several buffers pass B keeps the running sum in an
// internal accumulator laid out one contiguous block per workgroup (fully
// coalesced read-modify-write) and writes the host layout only when the
// integration completes.  Workgroup ids are remapped so each XCD processes a
// contiguous range of tiles and shares cache lines in its L2.
#include "fft_device.h"
#include "kernels.h"

namespace rsk {

// Threads per workgroup for a team transform of n points: RSK_FS_WG or one
// team, whichever is larger.  More teams per workgroup means longer
// contiguous runs in pass A's input gather and pass B's output.
#ifndef RSK_FS_WG
#define RSK_FS_WG 256
#endif
constexpr int kLdsBytes = 160 * 1024;
template <int NP>
constexpr int wg_team_np(int n) {
  // as many teams as RSK_FS_WG asks for, as far as their LDS regions fit
  const int t = n / elems_fs(n);
  const int per_team = (RSK_SEQ_POL ? 1 : NP) * (RSK_LDS_PAD > 0 ? n + (n >> RSK_LDS_PAD) : n) +
                       RSK_TEAM_PAD;
  const int tw = (RSK_LDS_TW && n > 1) ? n : 1;
  int teams = RSK_FS_WG / t > 0 ? RSK_FS_WG / t : 1;
  while (teams > 1 && (teams * per_team + tw) * 8 > kLdsBytes) teams /= 2;
  return teams * t;
}
constexpr int wg_team(int n) { return wg_team_np<2>(n); }
constexpr int tile_of(int n) { return wg_team(n) / (n / elems_fs(n)); }

// Offset within one spectrum's Y of the value (n1, k2 = j + m * T) in the
// layout [k2 / TBY][n1][k2 % TBY]; linear in m when T is a multiple of TBY.
template <int T, int TBY, int N1C>
__device__ __forceinline__ unsigned y_off(unsigned j, unsigned n1, int m) {
  if constexpr (T % TBY == 0) {
    return ((j / TBY) * N1C + n1) * TBY + (j % TBY) + (unsigned)m * (T * N1C);
  } else {
    const unsigned k2 = j + m * T;
    return ((k2 / TBY) * N1C + n1) * TBY + (k2 % TBY);
  }
}

// Pass A for N = N1 * N2 with N1 = R1 * N2 (R1 = 1 or 2).  The four-step
// twiddle W_N^e, e = n1 * k2 < N, is split as e = h * N1 + r:
// W_N^(h * N1) = W_N2^h is the sub-transform's own table (already in LDS)
// and W_N^r, r < N1 <= 4096, is a second LDS table filled from the first
// N1 entries of the FP64-generated fine table; so each twiddle costs two LDS
// reads and one complex product (two correctly rounded factors, as with the
// global two-level table) instead of two scattered global gathers.
template <int N2, int R1, int NBPS, int NP, bool Y24>
__global__ __launch_bounds__(wg_team(N2)) void k_fs_a(FsArgs a) {
  static_assert(!Y24 || NP == 2, "24-bit intermediate: two pols");
  constexpr int E = elems_fs(N2);
  constexpr int T = N2 / E;
  constexpr int TA = tile_of(N2);      // n1 values per workgroup
  constexpr int N1C = R1 * N2;
  constexpr int L1 = ilog2(N1C);
  static_assert(RSK_LDS_TW, "pass A reads the sub-transform table from LDS");
  static_assert((TA * team_stride<N2, NP>() + TwTab<N2>::LEN + N1C) * 8 <= kLdsBytes,
                "pass A LDS budget");
  using S = Sample<NBPS, NP>;
  using W = typename S::word;
  __shared__ float2 lds[TA * team_stride<N2, NP>()];
  __shared__ float2 twl[TwTab<N2>::LEN];
  __shared__ float2 twr[N1C];
  for (int m = threadIdx.x; m < N1C; m += wg_team(N2)) twr[m] = a.twf[m];
  const float2 *twp = tw_setup<N2, wg_team(N2)>(twl, a.tw);   // also waits for twr

  const unsigned N1 = N1C;
  const unsigned long long N = (unsigned long long)N1 * N2;
  const unsigned tiles = N1 / TA;
  const unsigned nsp = a.Nss * a.ynb;  // spectra per channel in Y
  const unsigned items = tiles * a.Nss * a.nbat * a.nch;
  const int tid = threadIdx.x;
  int sg = tid % TA;                   // n1 slot: fastest across lanes
  int j = tid / TA;
  constexpr unsigned TBY = tile_of(N1C);  // k2 tile of Y: pass B's workgroup tile

  // Item l: (channel cl, buffer b, spectrum sp, tile).  The workgroup takes
  // items l0, l0 + G, l0 + 2G, ... (G workgroups), loading the next item's
  // samples before transforming the current one, so the loads overlap the
  // transform; workgroup ids are remapped so each XCD takes a contiguous
  // range of every round of G items.
  auto load = [&](unsigned l, W (&raw)[E]) {
    const unsigned tile = l % tiles;
    const unsigned cs = l / tiles;
    const unsigned sp = cs % a.Nss;    // spectrum within its buffer
    const unsigned cb = cs / a.Nss;
    const unsigned b = cb % a.nbat;    // buffer within the batch
    const unsigned cl = cb / a.nbat;   // channel within the chunk
    // a base common to the workgroup and 32-bit element offsets per lane
    // (below N <= 2^24), so each load adds one constant to one offset
    const W *src = reinterpret_cast<const W *>(a.in[b] + (unsigned long long)(a.c0 + cl) * a.chan_stride) +
                   (unsigned long long)sp * N + tile * TA;
    const unsigned o = (unsigned)sg + (unsigned)j * N1;
#pragma unroll
    for (int m = 0; m < E; ++m) raw[m] = src[o + (unsigned)(m * T) * N1];
  };
  unsigned l = xcd_remap(blockIdx.x, gridDim.x);
  W raw[E];
  load(l, raw);
  for (;;) {
    // j and sg pass through an empty asm each item, so the compiler
    // recomputes the item's addresses instead of keeping 16 loads' and 16
    // stores' offsets live across the loop (which cost a wave per SIMD)
    asm volatile("" : "+v"(j), "+v"(sg));
    const unsigned ln = l + gridDim.x;   // uniform across the workgroup
    W rawn[E];
    if (ln < items) load(ln, rawn);
    float2 v[NP][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float2 o[NP];
      S::cvt(raw[m], o, a.isign);
#pragma unroll
      for (int p = 0; p < NP; ++p) v[p][m] = o[p];
    }
    team_fft<N2, E, NP, TwTab<N2>::TWN>(v, lds + sg * team_stride<N2, NP>(), j, twp);

    // Y[cl][(yb0 + b) * Nss + sp][k2 / TBY][n1][k2 % TBY][pol]; offsets
    // within one spectrum (N <= 2^24 values) fit 32 bits
    const unsigned tile = l % tiles;
    const unsigned cs = l / tiles;
    const unsigned sp = cs % a.Nss;
    const unsigned cb = cs / a.Nss;
    const unsigned b = cb % a.nbat;
    const unsigned cl = cb / a.nbat;
    const unsigned n1 = tile * TA + sg;
    const unsigned long long spec = (unsigned long long)cl * nsp + (a.yb0 + b) * a.Nss + sp;
    // spectrum `spec` starts at N * NP float2 (24-bit: N * 12 bytes) into Y
    float2 *dst = Y24 ? reinterpret_cast<float2 *>(reinterpret_cast<unsigned int *>(a.work) + spec * N * 3)
                      : a.work + spec * N * NP;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned k2 = j + m * T;
      const unsigned e = n1 * k2;          // < N <= 2^24
      const float2 w = cmul(twl[e >> L1], twr[e & (N1C - 1)]);
      const unsigned off = y_off<T, TBY, N1C>(j, n1, m);
      if constexpr (NP == 2) {
        const float2 y0 = cmul(v[0][m], w), y1 = cmul(v[1][m], w);
        const float4 yv = make_float4(y0.x, y0.y, y1.x, y1.y);
        if constexpr (Y24) store_y24(reinterpret_cast<unsigned int *>(dst) + 3u * off, yv);   // 3 off < 2^26
        else store_y(reinterpret_cast<float4 *>(dst) + off, yv);
      } else {
        store_y(dst + off, cmul(v[0][m], w));
      }
    }
    if (ln >= items) break;
    l = ln;
#pragma unroll
    for (int m = 0; m < E; ++m) raw[m] = rawn[m];
  }
}

// Waves per SIMD pass B's register budget allows (2: at most 256 registers
// per thread).
#ifndef RSK_FSB_WAVES
#define RSK_FSB_WAVES 2
#endif
template <int N1, int NP, int PMODE, bool Y24>
__global__ __launch_bounds__(wg_team(N1)) __attribute__((amdgpu_waves_per_eu(RSK_FSB_WAVES)))
void k_fs_b(FsArgs a) {
  static_assert(!Y24 || NP == 2, "24-bit intermediate: two pols");
  constexpr int E = elems_fs(N1);
  constexpr int T = N1 / E;
  constexpr int TB = tile_of(N1);      // k2 values per workgroup
  constexpr int Q = nq<PMODE>();
  __shared__ float2 lds[TB * team_stride<N1, NP>()];
  __shared__ float2 twl[TwTab<N1>::LEN];
  const float2 *twp = tw_setup<N1, wg_team(N1)>(twl, a.tw);

  const unsigned N2 = a.N2;
  const unsigned long long N = (unsigned long long)N1 * N2;
  const unsigned tiles = N2 / TB;
  const unsigned l = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned tile = l % tiles;
  const unsigned rg = l / tiles;
  const unsigned run = rg % a.ns;
  const unsigned cg = rg / a.ns;
  const unsigned g = cg % a.Ng;
  const unsigned cl = cg / a.Ng;
  const unsigned c = a.c0 + cl;
  const int tid = threadIdx.x;
  const int sg = tid % TB;             // k2 slot: fastest across lanes
  const int j = tid / TB;
  const unsigned k2 = tile * TB + sg;

  float acc[E][Q];
#pragma unroll
  for (int m = 0; m < E; ++m)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[m][q] = 0.0f;

  const unsigned sp0 = g * a.na_buf + run * a.sp_run;
  const unsigned sp1 = min(sp0 + a.sp_run, (g + 1) * a.na_buf);
  for (unsigned sp = sp0; sp < sp1; ++sp) {
    // this workgroup's tile: N1 rows of TB k2 values, contiguous
    const unsigned long long e0 = ((unsigned long long)cl * a.Nss * a.ynb + sp) * N +
                                  (unsigned long long)tile * N1 * TB;   // first value of the tile
    const float2 *y = a.work + e0 * NP;
    const unsigned int *y3 = reinterpret_cast<const unsigned int *>(a.work) + e0 * 3;
    float2 v[NP][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned long long ev = (unsigned long long)(j + m * T) * TB + sg;
      if constexpr (NP == 2) {
        float4 t;
        if constexpr (Y24) t = load_y24(y3 + 3 * ev);
        else t = load_y(reinterpret_cast<const float4 *>(y + ev * NP));
        v[0][m] = make_float2(t.x, t.y);
        v[1][m] = make_float2(t.z, t.w);
      } else {
        v[0][m] = load_y(y + ev * NP);
      }
    }
    team_fft<N1, E, NP, TwTab<N1>::TWN>(v, lds + sg * team_stride<N1, NP>(), j, twp);
#pragma unroll
    for (int m = 0; m < E; ++m) {
      detect_raw_np<PMODE, NP>(acc[m], v[0][m], v[NP - 1][m]);
    }
  }
#pragma unroll
  for (int m = 0; m < E; ++m) detect_finish<PMODE>(acc[m], a.gam);

  if (a.ns > 1) {
    // this run's sums, unscaled, in the internal accumulator's block layout
    float *pb = a.partb + ((((unsigned long long)run * a.Ng + g) * a.nch + cl) * tiles + tile) *
                              Q * N1 * TB + (unsigned long long)j * TB + sg;
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) store_acc(pb + ((unsigned long long)q * N1 + m * T) * TB, acc[m][q]);
    return;
  }

  // internal accumulator block of this workgroup: [Q][N1][TB]; thread
  // element (q, m) at ib + (q*N1 + m*T)*TB.  The mode is uniform, so each
  // branch is straight-line code whose loads are all in flight together.
  float *ib = a.acc_int + ((((unsigned long long)g * a.Nc + c) * tiles + tile) * Q) * N1 * TB +
              (unsigned long long)j * TB + sg;
  float *ob = a.acc + ((unsigned long long)g * Q * a.Nc + c) * N + k2;
#pragma unroll
  for (int m = 0; m < E; ++m)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[m][q] *= a.scale;
  if (a.mode == FS_MID || a.mode == FS_LAST) {
    float prev[E][Q];
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) prev[m][q] = load_acc(ib + ((unsigned long long)q * N1 + m * T) * TB);
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) acc[m][q] += prev[m][q];
  }
  if (a.mode == FS_FIRST || a.mode == FS_MID) {
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) store_acc(ib + ((unsigned long long)q * N1 + m * T) * TB, acc[m][q]);
  } else {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const unsigned k1 = j + m * T;
      const unsigned long long jo = (unsigned long long)N2 * ((k1 + N1 / 2) & (N1 - 1));
#pragma unroll
      for (int q = 0; q < Q; ++q) ob[(unsigned long long)q * a.Nc * N + jo] = acc[m][q];
    }
  }
}

int fs_tile_a(int n2) { return (int)(wg_team(n2) / (n2 / elems_fs(n2))); }
int fs_tile_b(int n1) { return (int)(wg_team(n1) / (n1 / elems_fs(n1))); }

template <int N2, int R1, int NBPS>
static int go_a_r(int np, const FsArgs &a, unsigned nb, hipStream_t s) {
  constexpr int WGT = wg_team(N2);
  if (np == 2 && a.y24) hipLaunchKernelGGL((k_fs_a<N2, R1, NBPS, 2, true>), dim3(nb), dim3(WGT), 0, s, a);
  else if (np == 2) hipLaunchKernelGGL((k_fs_a<N2, R1, NBPS, 2, false>), dim3(nb), dim3(WGT), 0, s, a);
  else hipLaunchKernelGGL((k_fs_a<N2, R1, NBPS, 1, false>), dim3(nb), dim3(WGT), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <int N2, int NBPS>
static int go_a(int np, const FsArgs &a, unsigned nb, hipStream_t s) {
  if (a.N1 == (unsigned)N2) return go_a_r<N2, 1, NBPS>(np, a, nb, s);
  if constexpr (2 * N2 <= kTwN) {
    if (a.N1 == 2u * N2) return go_a_r<N2, 2, NBPS>(np, a, nb, s);
  }
  return 2;
}

template <int NBPS>
static int go_a_n(int np, const FsArgs &a, unsigned nb, hipStream_t s) {
  switch (a.N2) {
    case 64: return go_a<64, NBPS>(np, a, nb, s);
    case 128: return go_a<128, NBPS>(np, a, nb, s);
    case 256: return go_a<256, NBPS>(np, a, nb, s);
    case 512: return go_a<512, NBPS>(np, a, nb, s);
    case 1024: return go_a<1024, NBPS>(np, a, nb, s);
    case 2048: return go_a<2048, NBPS>(np, a, nb, s);
    case 4096: return go_a<4096, NBPS>(np, a, nb, s);
  }
  return 2;
}

int launch_fs_a(int nbps, int np, const FsArgs &a, hipStream_t s) {
  if (a.y24 && np != 2) return 2;
  const unsigned tiles = a.N1 / fs_tile_a((int)a.N2);
  const unsigned long long items = (unsigned long long)tiles * a.Nss * a.nbat * a.nch;
  const unsigned long long nb = a.grid_a && a.grid_a < items ? a.grid_a : items;
  // Y's k2 tile: pass B's workgroup tile
  if (nb == 0 || items > 0x7fffffffull || a.tby != (unsigned)fs_tile_b((int)a.N1) || a.nbat == 0 ||
      a.nbat > (unsigned)kMaxBatch || a.yb0 + a.nbat > a.ynb)
    return 2;
  if (nbps == 8) return go_a_n<8>(np, a, (unsigned)nb, s);
  if (nbps == 16) return go_a_n<16>(np, a, (unsigned)nb, s);
  return 2;
}

template <int N1, bool Y24>
static void go_b2(int pmode, const FsArgs &a, unsigned nb, hipStream_t s) {
  constexpr int WGT = wg_team(N1);
  if (pmode == 1) hipLaunchKernelGGL((k_fs_b<N1, 2, 1, Y24>), dim3(nb), dim3(WGT), 0, s, a);
  else if (pmode == 4) hipLaunchKernelGGL((k_fs_b<N1, 2, 4, Y24>), dim3(nb), dim3(WGT), 0, s, a);
  else hipLaunchKernelGGL((k_fs_b<N1, 2, -4, Y24>), dim3(nb), dim3(WGT), 0, s, a);
}

template <int N1>
static int go_b(int np, int pmode, const FsArgs &a, unsigned nb, hipStream_t s) {
  constexpr int WGT = wg_team(N1);
  if (np == 1) {
    hipLaunchKernelGGL((k_fs_b<N1, 1, 1, false>), dim3(nb), dim3(WGT), 0, s, a);
  } else if (pmode == 1 || pmode == 4 || pmode == -4) {
    if (a.y24) go_b2<N1, true>(pmode, a, nb, s);
    else go_b2<N1, false>(pmode, a, nb, s);
  } else {
    return 2;
  }
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The runs must cover every spectrum of a group, and a split launch needs
// its partial-sum buffer.
static bool runs_ok(const FsArgs &a) {
  return a.ns >= 1 && (unsigned long long)a.ns * a.sp_run >= a.na_buf &&
         (a.ns == 1 || a.partb != nullptr);
}

int launch_fs_b(int np, int pmode, const FsArgs &a, hipStream_t s) {
  if (a.y24 && np != 2) return 2;
  const unsigned tiles = a.N2 / fs_tile_b((int)a.N1);
  const unsigned long long nb = (unsigned long long)tiles * a.ns * a.Ng * a.nch;
  if (nb == 0 || nb > 0x7fffffffull || a.nbat == 0 || a.nbat > a.ynb || !runs_ok(a)) return 2;
  switch (a.N1) {
    case 128: return go_b<128>(np, pmode, a, (unsigned)nb, s);
    case 256: return go_b<256>(np, pmode, a, (unsigned)nb, s);
    case 512: return go_b<512>(np, pmode, a, (unsigned)nb, s);
    case 1024: return go_b<1024>(np, pmode, a, (unsigned)nb, s);
    case 2048: return go_b<2048>(np, pmode, a, (unsigned)nb, s);
    case 4096: return go_b<4096>(np, pmode, a, (unsigned)nb, s);
  }
  return 2;
}

// One thread per value of the launch's sums: the ns partial sums in run
// order, then the finish k_fs_b applies to its own sums (same scale, same
// internal-accumulator block, same shifted output position).
__global__ __launch_bounds__(256) void k_fs_b_reduce(FsArgs a, unsigned Q, unsigned TB) {
  const unsigned N1 = a.N1, N2 = a.N2, tiles = N2 / TB;
  const unsigned long long N = (unsigned long long)N1 * N2;
  const unsigned long long total = (unsigned long long)a.Ng * a.nch * tiles * Q * N1 * TB;
  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (unsigned long long)gridDim.x * blockDim.x) {
    // idx runs over [Ng][nch][tiles][Q][N1][TB], one run's block of partb
    const unsigned sg = (unsigned)(idx % TB);
    unsigned long long r = idx / TB;
    const unsigned k1 = (unsigned)(r % N1);
    r /= N1;
    const unsigned q = (unsigned)(r % Q);
    r /= Q;
    const unsigned tile = (unsigned)(r % tiles);
    r /= tiles;
    const unsigned cl = (unsigned)(r % a.nch);
    const unsigned g = (unsigned)(r / a.nch);
    float sum = 0.0f;
    for (unsigned run = 0; run < a.ns; ++run) sum += load_acc(a.partb + run * total + idx);
    float v = sum * a.scale;
    const unsigned c = a.c0 + cl;
    float *ie = a.acc_int + ((((unsigned long long)g * a.Nc + c) * tiles + tile) * Q + q) * N1 * TB +
                (unsigned long long)k1 * TB + sg;
    if (a.mode == FS_MID || a.mode == FS_LAST) v += load_acc(ie);
    if (a.mode == FS_FIRST || a.mode == FS_MID) {
      store_acc(ie, v);
    } else {
      const unsigned k2 = tile * TB + sg;
      a.acc[((unsigned long long)g * Q * a.Nc + c) * N + k2 + (unsigned long long)q * a.Nc * N +
            (unsigned long long)N2 * ((k1 + N1 / 2) & (N1 - 1))] = v;
    }
  }
}

int launch_fs_b_reduce(int pmode, const FsArgs &a, hipStream_t s) {
  if (a.ns < 2 || !runs_ok(a)) return 2;
  const unsigned Q = pmode == 1 ? 1u : 4u;
  const unsigned TB = (unsigned)fs_tile_b((int)a.N1);
  const unsigned long long total = (unsigned long long)a.Ng * a.nch * (a.N2 / TB) * Q * a.N1 * TB;
  if (total == 0) return 2;
  unsigned long long nb = (total + 255) / 256;
  if (nb > 16384) nb = 16384;
  hipLaunchKernelGGL(k_fs_b_reduce, dim3((unsigned)nb), dim3(256), 0, s, a, Q, TB);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace rsk
