// k_mfma.hip -- spectrometer kernels whose transforms run on the gfx950
// matrix cores (v_mfma_f32_32x32x16_f16 with hi/lo split operands; see
// mfma_dft.h for the decomposition, lane layout and precision).
//
// k_small_mfma: the fused kernel for N = 1024 (below).
//
// k_fs_b_mfma: pass B of the four-step path for N1 = 1024 and two pols, the
// same contract as k_fs_b (kernels.h: FsArgs, modes FS_HOST .. FS_LAST):
// length-1024 DFTs over n1 for every k2, detection, integration of the dump
// group's spectra, FFT-shifted host layout or the internal accumulator.
// One wavefront owns one k2 column of one channel and dump group, so the
// transform needs no LDS exchange and no barrier; a workgroup is four
// wavefronts on four consecutive k2 and shares LDS only to write the host
// layout in 16-byte runs.  Pass A writes Y untiled for it,
// Y[cl][s][k2][n1][pol], so each wavefront's loads are contiguous.
#include "fft_device.h"
#include "kernels.h"
#include "mfma_dft.h"

namespace rsk {

constexpr int kMfmaWaves = 4;                      // k2 columns per workgroup
constexpr int kOutRow = 4 * 1024 + 32;             // floats per q in the staging area

// Raw detection sums (detect_raw) of one k2 column and spectrum (both pols
// transformed, then detected into acc, which is overwritten when SET, else
// added to).
template <int PMODE, bool SET>
__device__ __forceinline__ void fs_b_column(const float4 *y, float s1,
                                            const Dft1024K &K, float (&acc)[16][nq<PMODE>()]) {
  const int lane = threadIdx.x & 63;
  const int col = lane & 31, h = lane >> 5;
  float4 v[4][4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int i = 0; i < 4; ++i) v[kb][i] = y[kMfmaWaves * (col + 32 * (8 * kb + 4 * h + i))];
  f16v xr[2], xi[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    float2 x[4][4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        x[kb][i] = p == 0 ? make_float2(v[kb][i].x * s1, v[kb][i].y * s1)
                          : make_float2(v[kb][i].z * s1, v[kb][i].w * s1);
    dft1024<true>(x, K, xr[p], xi[p]);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    if constexpr (SET) {
#pragma unroll
      for (int q = 0; q < nq<PMODE>(); ++q) acc[r][q] = 0.0f;
    }
    detect_raw<PMODE>(acc[r], xr[0][r], xi[0][r], xr[1][r], xi[1][r]);
  }
}

// Wavefronts per SIMD the loop variant of k_fs_b_mfma is built for (the
// running sum is live across its transforms).  At two (256 VGPRs) the
// four-component modes spill 76 bytes per lane, for which the runtime keeps
// scratch memory for every wave the device can hold; at one the spare
// registers hold them, at the same speed (0.348 ms per 256 MiB buffer of
// the 64-channel 2^20 product either way).
#ifndef RSK_FSB_LOOP_WAVES
#define RSK_FSB_LOOP_WAVES 1
#endif

// ONE: the dump group is a single spectrum (na_buf == 1, e.g. one 2^20
// spectrum per channel and buffer), so the running sum is not live during
// the transforms; otherwise a loop over the group's spectra.
template <int PMODE, bool ONE>
__global__ __launch_bounds__(64 * kMfmaWaves, ONE ? 2 : RSK_FSB_LOOP_WAVES) void k_fs_b_mfma(FsArgs a) {
  constexpr int N1 = 1024;
  constexpr int Q = nq<PMODE>();
  // LDS: the DFT constants while transforming, then the output staging area
  constexpr int kStBytes = Q * kOutRow * 4;
  constexpr int kLds = kStBytes > (int)sizeof(Dft1024K) ? kStBytes : (int)sizeof(Dft1024K);
  __shared__ __attribute__((aligned(16))) unsigned char lds_raw[kLds];
  Dft1024K &K = *reinterpret_cast<Dft1024K *>(lds_raw);
  float *st = reinterpret_cast<float *>(lds_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, h = lane >> 5;
  const unsigned N2 = a.N2;
  const unsigned long long N = (unsigned long long)N1 * N2;
  const unsigned tiles = N2 / kMfmaWaves;
  const unsigned l = xcd_remap(blockIdx.x, gridDim.x);
  const unsigned tile = l % tiles;
  const unsigned cg = l / tiles;
  const unsigned g = cg % a.Ng;
  const unsigned cl = cg / a.Ng;
  const unsigned c = a.c0 + cl;
  const unsigned k2 = tile * kMfmaWaves + wv;

  dft1024_consts(K, a.tw, 64 * kMfmaWaves);
  __syncthreads();

  float acc[16][Q];
  const float s1 = a.mscale;
  // Y tiled [k2 / 4][n1][k2 % 4][pol]: the workgroup's 4 columns are one
  // tile, interleaved, so this wavefront's column is every 4th float4
  const float4 *y0 = reinterpret_cast<const float4 *>(a.work) +
                     ((unsigned long long)cl * a.Nss * a.ynb + g * a.na_buf) * N +
                     (unsigned long long)tile * N1 * kMfmaWaves + wv;
  if constexpr (ONE) {
    fs_b_column<PMODE, true>(y0, s1, K, acc);
  } else {
#pragma unroll
    for (int r = 0; r < 16; ++r)
#pragma unroll
      for (int q = 0; q < Q; ++q) acc[r][q] = 0.0f;
    for (unsigned sp = 0; sp < a.na_buf; ++sp) fs_b_column<PMODE, false>(y0 + sp * N, s1, K, acc);
  }

  // raw sums -> components; undo the operand scaling (X = 32 * result /
  // s1) and normalise
  const float sc = a.scale * (32.0f / s1) * (32.0f / s1);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    detect_finish<PMODE>(acc[r], a.gam);
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[r][q] *= sc;
  }

  // internal accumulator (integrations over several buffers): one
  // contiguous [Q][16][64] block per wavefront
  if (a.mode != FS_HOST) {
    float *ib = a.acc_int + (((unsigned long long)g * a.Nc + c) * N2 + k2) * (Q * 16 * 64) + lane;
    if (a.mode == FS_MID || a.mode == FS_LAST) {
      float prev[16][Q];
#pragma unroll
      for (int r = 0; r < 16; ++r)
#pragma unroll
        for (int q = 0; q < Q; ++q) prev[r][q] = load_acc(ib + (q * 16 + r) * 64);
#pragma unroll
      for (int r = 0; r < 16; ++r)
#pragma unroll
        for (int q = 0; q < Q; ++q) acc[r][q] += prev[r][q];
    }
    if (a.mode == FS_FIRST || a.mode == FS_MID) {
#pragma unroll
      for (int r = 0; r < 16; ++r)
#pragma unroll
        for (int q = 0; q < Q; ++q) store_acc(ib + (q * 16 + r) * 64, acc[r][q]);
      return;   // uniform across the workgroup: no wavefront reaches the barrier below
    }
  }
  // host layout: bin k1 = c_r + 32 col goes to shifted row (k1 + 512) % 1024;
  // stage [q][row][4 k2] in LDS (one pad float per 32 rows), then write 16
  // bytes (the workgroup's 4 consecutive k2) per row
  __syncthreads();   // every wavefront is done with the constants
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int k1 = 8 * (r >> 2) + 4 * h + (r & 3) + 32 * col;
    const int row = (k1 + N1 / 2) & (N1 - 1);
#pragma unroll
    for (int q = 0; q < Q; ++q) st[q * kOutRow + 4 * row + (row >> 5) + wv] = acc[r][q];
  }
  __syncthreads();
  float *ob = a.acc + ((unsigned long long)g * Q * a.Nc + c) * N + (unsigned long long)tile * kMfmaWaves;
  for (int t = threadIdx.x; t < Q * N1; t += 64 * kMfmaWaves) {
    const int q = t / N1, row = t - q * N1;
    const float *s = st + q * kOutRow + 4 * row + (row >> 5);
    reinterpret_cast<float4 *>(ob + (unsigned long long)q * a.Nc * N + (unsigned long long)N2 * row)[0] =
        make_float4(s[0], s[1], s[2], s[3]);
  }
}

// k_small_mfma: the fused small-N kernel (k_small's contract, SmallArgs)
// for N = 1024 with two pols: unpack -> 1024-point DFT on the matrix cores
// -> detection -> integration in registers -> FFT-shifted output.  One
// wavefront integrates one chunk of consecutive spectra of one coarse
// channel (item = (channel, group, chunk)); four independent wavefronts per
// workgroup share the DFT constants in LDS, and the same LDS then stages
// each wavefront's sums so the output is written in 256-byte runs.  Integer
// samples are exact in FP16 up to 2048, so int8 input skips the lo operand
// of the first stage.
constexpr int kSmRow = 1024 + 32;                  // staged floats per (wave, q)

template <int NBPS, int PMODE>
__global__ __launch_bounds__(64 * kMfmaWaves, (NBPS == 16 && PMODE != 1) ? 1 : 2) void k_small_mfma(SmallArgs a, unsigned items) {
  constexpr int N = 1024;
  constexpr int Q = nq<PMODE>();
  constexpr bool LO = NBPS == 16;
  using S = Sample<NBPS, 2>;
  using W = typename S::word;
  constexpr int kStBytes = kMfmaWaves * Q * kSmRow * 4;
  constexpr int kLds = kStBytes > (int)sizeof(Dft1024K) ? kStBytes : (int)sizeof(Dft1024K);
  __shared__ __attribute__((aligned(16))) unsigned char lds_raw[kLds];
  Dft1024K &K = *reinterpret_cast<Dft1024K *>(lds_raw);
  float *st = reinterpret_cast<float *>(lds_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, h = lane >> 5;
  const unsigned item = blockIdx.x * kMfmaWaves + wv;
  const bool valid = item < items;
  // item -> (channel, group, chunk) as k_small's (c, gb, chunk) with gpw = 1
  const unsigned it = valid ? item : 0u;
  const unsigned c = it / (a.Ng * a.G);
  const unsigned rest = it - c * (a.Ng * a.G);
  const unsigned g = rest / a.G;
  const unsigned chunk = rest - g * a.G;

  dft1024_consts(K, a.tw, 64 * kMfmaWaves);
  __syncthreads();

  float acc[16][Q];
#pragma unroll
  for (int r = 0; r < 16; ++r)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[r][q] = 0.0f;
  const W *src = reinterpret_cast<const W *>(a.in + (unsigned long long)c * a.chan_stride) +
                 ((unsigned long long)g * a.na_buf + (unsigned long long)chunk * a.chunk) * N;
  const unsigned nspec = valid ? a.chunk : 0u;
  for (unsigned s = 0; s < nspec; ++s) {
    W raw[4][4];
#pragma unroll
    for (int kb = 0; kb < 4; ++kb)
#pragma unroll
      for (int i = 0; i < 4; ++i) raw[kb][i] = src[(unsigned long long)s * N + col + 32 * (8 * kb + 4 * h + i)];
    f16v xr[2], xi[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      float2 x[4][4];
#pragma unroll
      for (int kb = 0; kb < 4; ++kb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          float2 o[2];
          S::cvt(raw[kb][i], o, a.isign);
          x[kb][i] = o[p];
        }
      dft1024<LO>(x, K, xr[p], xi[p]);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) detect_raw<PMODE>(acc[r], xr[0][r], xi[0][r], xr[1][r], xi[1][r]);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) detect_finish<PMODE>(acc[r], a.gam);

  // the transforms return X / 32: sums are 1024 x too small
  __syncthreads();   // every wavefront is done with the constants
  float *sw = st + wv * (Q * kSmRow);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int k = 8 * (r >> 2) + 4 * h + (r & 3) + 32 * col;
    const int jo = (k + N / 2) & (N - 1);            // FFT-shifted fine channel
#pragma unroll
    for (int q = 0; q < Q; ++q) sw[q * kSmRow + jo + (jo >> 5)] = acc[r][q] * 1024.0f;
  }
  __syncthreads();
  if (!valid) return;
  for (int q = 0; q < Q; ++q) {
    float *dst;
    if (a.direct) dst = a.acc + (((unsigned long long)g * Q + q) * a.Nc + c) * N;
    else dst = a.part + ((((unsigned long long)c * a.Ng + g) * a.G + chunk) * Q + q) * N;
#pragma unroll 4
    for (int jo = lane; jo < N; jo += 64) {
      float v = sw[q * kSmRow + jo + (jo >> 5)];
      if (a.direct) {
        v *= a.scale;
        if (!a.first) v += dst[jo];
      }
      dst[jo] = v;
    }
  }
}

// ---- an 8-point product fused into the unsplit 1024-point kernel ----------
// Forward 8-point DFT in place, natural order: radix 2, decimation in
// frequency, W8 = exp(-2 pi i / 8).
__device__ __forceinline__ void dft8(float2 (&x)[8]) {
  constexpr float r = 0.70710678118654752f;
  const float2 a0 = cadd(x[0], x[4]), a1 = cadd(x[1], x[5]);
  const float2 a2 = cadd(x[2], x[6]), a3 = cadd(x[3], x[7]);
  const float2 b0 = csub(x[0], x[4]), d1 = csub(x[1], x[5]);
  const float2 d2 = csub(x[2], x[6]), d3 = csub(x[3], x[7]);
  // times W8^1 = (1 - i) r, W8^2 = -i, W8^3 = (-1 - i) r
  const float2 b1 = make_float2((d1.x + d1.y) * r, (d1.y - d1.x) * r);
  const float2 b2 = make_float2(d2.y, -d2.x);
  const float2 b3 = make_float2((d3.y - d3.x) * r, -(d3.x + d3.y) * r);
  // the two 4-point halves: even outputs from a, odd outputs from b
  const float2 c0 = cadd(a0, a2), c1 = cadd(a1, a3), c2 = csub(a0, a2), e3 = csub(a1, a3);
  const float2 c3 = make_float2(e3.y, -e3.x);
  const float2 g0 = cadd(b0, b2), g1 = cadd(b1, b3), g2 = csub(b0, b2), f3 = csub(b1, b3);
  const float2 g3 = make_float2(f3.y, -f3.x);
  x[0] = cadd(c0, c1);
  x[4] = csub(c0, c1);
  x[2] = cadd(c2, c3);
  x[6] = csub(c2, c3);
  x[1] = cadd(g0, g1);
  x[5] = csub(g0, g1);
  x[3] = cadd(g2, g3);
  x[7] = csub(g2, g3);
}

template <int CTRL, int BANKS>
__device__ __forceinline__ float dpp_f(float old, float src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, src),
                                                               CTRL, 0xf, BANKS, false));
}
template <int CTRL>
__device__ __forceinline__ float dpp_q(float src) {   // quad_perm: every lane written
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, src),
                                                               CTRL, 0xf, 0xf, false));
}

// One step of a cross-lane reduce-scatter between lanes l and l ^ D: x and
// y are values k and k + n/2 of a lane's n partial sums; afterwards x holds,
// summed over both lanes, value k where lane bit D is clear and value
// k + n/2 where it is set.
template <int D>
__device__ __forceinline__ void halve(float &x, float y, int lane) {
  if constexpr (D == 32 || D == 16) {
    // gfx950 lane swaps: the upper 32 lanes (D = 32) or the odd 16-lane rows
    // (D = 16) of the first operand trade places with the lower lanes / even
    // rows of the second
    const auto r = D == 32 ? __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, x),
                                                              __builtin_bit_cast(unsigned, y), false, false)
                           : __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, x),
                                                              __builtin_bit_cast(unsigned, y), false, false);
    // the second result passes through an empty asm: with both results
    // bit-cast to float, this compiler's optimiser replaces the second by
    // the first (fadd of one value with itself)
    unsigned int r1 = r[1];
    asm volatile("" : "+v"(r1));
    x = __builtin_bit_cast(float, r[0]) + __builtin_bit_cast(float, r1);
  } else if constexpr (D == 8 || D == 4) {
    // row shifts; each writes only the lanes (4-lane banks) it serves
    constexpr int lo = D == 8 ? 0x3 : 0x5;
    const float m1 = dpp_f<0x100 + D, lo>(y, x);          // row_shl:D
    const float m2 = dpp_f<0x110 + D, 0xf & ~lo>(x, y);   // row_shr:D
    x = m1 + m2;
  } else {
    static_assert(D == 2 || D == 1, "lane distance");
    const bool up = (lane & D) != 0;
    const float keep = up ? y : x, send = up ? x : y;
    x = keep + dpp_q<D == 2 ? 0x4e : 0xb1>(send);        // quad_perm lane ^ D
  }
}

// Sum of V partial sums v[0..V) over the 64 lanes of the wavefront: value o
// ends in lanes (o << (6 - log2 V)) .. (o + 1 << (6 - log2 V)) - 1.
template <int V>
__device__ __forceinline__ float wave_reduce(float (&v)[V], int lane) {
  static_assert(V == 32 || V == 8, "32 or 8 partial sums");
#pragma unroll
  for (int k = 0; k < V / 2; ++k) halve<32>(v[k], v[k + V / 2], lane);
#pragma unroll
  for (int k = 0; k < V / 4; ++k) halve<16>(v[k], v[k + V / 4], lane);
#pragma unroll
  for (int k = 0; k < V / 8; ++k) halve<8>(v[k], v[k + V / 8], lane);
  float x = v[0];
  if constexpr (V == 32) {
    halve<4>(x, v[2], lane);
    float y = v[1];
    halve<4>(y, v[3], lane);
    halve<2>(x, y, lane);
    return x + dpp_q<0xb1>(x);                            // lanes 2m, 2m + 1
  } else {
    x += dpp_f<0x141, 0xf>(0.0f, x);                      // row_half_mirror: l, 7 - l
    x += dpp_q<0xb1>(x);
    return x + dpp_q<0x4e>(x);                            // 8-lane groups
  }
}

// Staging of one spectrum's samples for the fused 8-point product: sample n
// (its two pols' FP16 pairs) at uint2 n + 2 (n >> 4), so the 16-byte reads
// of eight consecutive lanes, each lane one spectrum of 8 samples, fall in
// different banks.
constexpr int kF8Tile = 1024 + 2 * 64;                 // uint2 per wavefront
__device__ __forceinline__ int f8_slot(int n) { return n + 2 * (n >> 4); }

// k_small_mfma_u: k_small_mfma's contract for int8 samples with unsplit
// FP16 operands (mfma_dft.h, dft1024u): the samples become FP16 pairs by a
// byte permute (no FP32 conversion), the inverse transform uses conjugated
// constants (no sign flips per sample), detection accumulates raw sums with
// fused multiply-adds and forms the components once per output.  The DFT
// matrix (32 registers) stays in registers for the whole chunk, the
// twiddles are read from LDS per transform; the samples are loaded two
// spectra ahead of the transforms.
//
// F8: the same wavefronts also compute an 8-point product whose dumps
// integrate 128 spectra, i.e. one dump per 1024-sample spectrum of this
// product, from the samples already loaded: each spectrum's FP16 pairs go
// through a per-wavefront LDS tile, every lane reads two 8-sample spectra
// back, transforms both pols in FP32 (exact inputs, the FP32 vector
// engine's precision) and detects; the wavefront's 32 (or 8) partial sums
// are reduced across lanes and written, FFT-shifted, to a.f8_out.  The
// transform is always forward: for the inverse, bin k is forward bin -k and
// the cross term's sign is folded into gam as for the 1024-point product.
// The 8-point work needs registers, so with F8 the DFT matrix is re-read
// from LDS per spectrum and the samples are loaded one spectrum ahead (the
// longer iteration covers the load latency).
template <int PMODE, bool F8>
__global__ __launch_bounds__(64 * kMfmaWaves, 2) void k_small_mfma_u(SmallArgs a, unsigned items) {
  constexpr int N = 1024;
  constexpr int Q = nq<PMODE>();
  constexpr int kStBytes = kMfmaWaves * Q * kSmRow * 4;
  constexpr int kLoopBytes = (int)sizeof(Dft1024U) + (F8 ? kMfmaWaves * kF8Tile * 8 : 0);
  constexpr int kLds = kStBytes > kLoopBytes ? kStBytes : kLoopBytes;
  __shared__ __attribute__((aligned(16))) unsigned char lds_raw[kLds];
  Dft1024U &K = *reinterpret_cast<Dft1024U *>(lds_raw);
  float *st = reinterpret_cast<float *>(lds_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  uint2 *t8 = reinterpret_cast<uint2 *>(lds_raw + sizeof(Dft1024U)) + wv * kF8Tile;
  const int col = lane & 31, h = lane >> 5;
  uint2 *const t8w = t8 + f8_slot(col + 128 * h);   // this lane's first tile store
  const unsigned item = blockIdx.x * kMfmaWaves + wv;
  const bool valid = item < items;
  const unsigned it = valid ? item : 0u;
  const unsigned c = it / (a.Ng * a.G);
  const unsigned rest = it - c * (a.Ng * a.G);
  const unsigned g = rest / a.G;
  const unsigned chunk = rest - g * a.G;

  // isign = -1 asks for the inverse transform: conjugated constants, and the
  // cross term's sign is the input-conjugation sign alone
  dft1024u_consts(K, a.tw, 64 * kMfmaWaves, a.isign);
  __syncthreads();
  // the DFT matrix: held for the whole chunk, or (F8, whose 8-point work
  // needs the registers) read from LDS per spectrum
  h8v bre[4], bim[4];
  if constexpr (!F8) {
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      bre[kb] = K.b[0][kb][lane];
      bim[kb] = K.b[1][kb][lane];
    }
  }
  const float gam = a.gam * a.isign;

  float acc[16][Q];
#pragma unroll
  for (int r = 0; r < 16; ++r)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[r][q] = 0.0f;
  const unsigned int *src = reinterpret_cast<const unsigned int *>(a.in + (unsigned long long)c * a.chan_stride) +
                            ((unsigned long long)g * a.na_buf + (unsigned long long)chunk * a.chunk) * N +
                            col + 32 * 4 * h;
  const unsigned nspec = valid ? a.chunk : 0u;
  // samples of spectrum s in raw, of s + 1 in nxt (loads two spectra ahead
  // of the transforms: one spectrum's compute does not cover HBM latency;
  // F8, with the 8-point work besides, loads one spectrum ahead into raw)
  constexpr unsigned AHEAD = F8 ? 1 : 2;
  unsigned int raw[16], nxt[16];
  if (nspec) {
#pragma unroll
    for (int e = 0; e < 16; ++e) raw[e] = src[32 * (8 * (e >> 2) + (e & 3))];
  }
  if (!F8 && nspec > 1) {
#pragma unroll
    for (int e = 0; e < 16; ++e) nxt[e] = src[N + 32 * (8 * (e >> 2) + (e & 3))];
  }
  for (unsigned s = 0; s < nspec; ++s) {
    // FP16 (re, im) pairs of both pols: element (kb, i) is sample
    // col + 32 (8 kb + 4 h + i)
    h8v ah[2][4];
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const unsigned int t = raw[e] ^ 0x80808080u;
      const unsigned int u0 = i8_pair_f16<0>(t), u1 = i8_pair_f16<1>(t);
      const h2v p0 = __builtin_bit_cast(h2v, u0);
      const h2v p1 = __builtin_bit_cast(h2v, u1);
      ah[0][e >> 2][2 * (e & 3)] = p0[0];
      ah[0][e >> 2][2 * (e & 3) + 1] = p0[1];
      ah[1][e >> 2][2 * (e & 3)] = p1[0];
      ah[1][e >> 2][2 * (e & 3) + 1] = p1[1];
      // f8_slot(col + 128 h + 32 K) = f8_slot(col + 128 h) + 36 K (32 K is a
      // multiple of 16), so every store is one base plus a constant
      if constexpr (F8) t8w[36 * (8 * (e >> 2) + (e & 3))] = make_uint2(u0, u1);
    }
    if constexpr (!F8) {
#pragma unroll
      for (int e = 0; e < 16; ++e) raw[e] = nxt[e];
    }
    if (s + AHEAD < nspec) {
      const unsigned int *nx = src + (unsigned long long)(s + AHEAD) * N;
#pragma unroll
      for (int e = 0; e < 16; ++e) (F8 ? raw : nxt)[e] = nx[32 * (8 * (e >> 2) + (e & 3))];
    }
    // the twiddles are read from LDS per transform (an opaque zero in the
    // index keeps the compiler from holding all 32 registers of them)
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    if constexpr (F8) {
#pragma unroll
      for (int kb = 0; kb < 4; ++kb) {
        bre[kb] = K.b[0][kb][lane + z];
        bim[kb] = K.b[1][kb][lane + z];
      }
    }
    // one polarisation after the other (scheduling barriers): interleaving
    // the two transforms would double their live registers
    f16v xr0, xi0, xr1, xi1;
    __builtin_amdgcn_sched_barrier(0);
    dft1024u(ah[0], bre, bim, K.tw, lane + z, xr0, xi0);
    __builtin_amdgcn_sched_barrier(0);
    dft1024u(ah[1], bre, bim, K.tw, lane + z, xr1, xi1);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < 16; ++r) detect_raw<PMODE>(acc[r], xr0[r], xi0[r], xr1[r], xi1[r]);
    if constexpr (F8) {
      // the tile is this wavefront's own: its LDS operations complete in
      // order, so a fence for the compiler is all the exchange needs
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      float a8[8][Q];
#pragma unroll
      for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int q = 0; q < Q; ++q) a8[b][q] = 0.0f;
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int m = lane + 64 * t;                      // 8-point spectrum of this one
        const uint4 *p4 = reinterpret_cast<const uint4 *>(t8 + f8_slot(8 * m));
        uint4 w[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) w[k] = p4[k];
        float2 x0[8], x1[8];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const h2v s00 = __builtin_bit_cast(h2v, w[k].x), s01 = __builtin_bit_cast(h2v, w[k].y);
          const h2v s10 = __builtin_bit_cast(h2v, w[k].z), s11 = __builtin_bit_cast(h2v, w[k].w);
          x0[2 * k] = make_float2((float)s00[0], (float)s00[1]);
          x1[2 * k] = make_float2((float)s01[0], (float)s01[1]);
          x0[2 * k + 1] = make_float2((float)s10[0], (float)s10[1]);
          x1[2 * k + 1] = make_float2((float)s11[0], (float)s11[1]);
        }
        dft8(x0);
        dft8(x1);
#pragma unroll
        for (int b = 0; b < 8; ++b) detect_raw<PMODE>(a8[b], x0[b].x, x0[b].y, x1[b].x, x1[b].y);
        __builtin_amdgcn_sched_barrier(0);             // one spectrum's samples live at a time
      }
      constexpr int V = 8 * Q;
      float v[V];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        detect_finish<PMODE>(a8[b], gam);
#pragma unroll
        for (int q = 0; q < Q; ++q) v[q * 8 + b] = a8[b][q];
      }
      const float sum = wave_reduce<V>(v, lane);
      constexpr int SH = V == 32 ? 1 : 3;
      if ((lane & ((1 << SH) - 1)) == 0) {
        const int o = lane >> SH, q = o >> 3, b = o & 7;
        const int k = a.isign > 0.0f ? b : (8 - b) & 7;
        const unsigned long long d = (unsigned long long)g * a.na_buf + (unsigned long long)chunk * a.chunk + s;
        a.f8_out[((d * Q + q) * a.Nc + c) * 8 + ((k + 4) & 7)] = sum * a.scale;
      }
      // the next spectrum's stores to the tile follow these loads
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }

  __syncthreads();   // every wavefront is done with the constants
  float *sw = st + wv * (Q * kSmRow);
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int k = 8 * (r >> 2) + 4 * h + (r & 3) + 32 * col;
    const int jo = (k + N / 2) & (N - 1);            // FFT-shifted fine channel
    detect_finish<PMODE>(acc[r], gam);
#pragma unroll
    for (int q = 0; q < Q; ++q) sw[q * kSmRow + jo + (jo >> 5)] = acc[r][q];
  }
  __syncthreads();
  if (!valid) return;
  for (int q = 0; q < Q; ++q) {
    float *dst;
    if (a.direct) dst = a.acc + (((unsigned long long)g * Q + q) * a.Nc + c) * N;
    else dst = a.part + ((((unsigned long long)c * a.Ng + g) * a.G + chunk) * Q + q) * N;
#pragma unroll 4
    for (int jo = lane; jo < N; jo += 64) {
      float v = sw[q * kSmRow + jo + (jo >> 5)];
      if (a.direct) {
        v *= a.scale;
        if (!a.first) v += dst[jo];
      }
      dst[jo] = v;
    }
  }
}

template <int PMODE, bool F8>
static int go_small_mfma_u(const SmallArgs &a, unsigned items, hipStream_t s) {
  const unsigned nb = (items + kMfmaWaves - 1) / kMfmaWaves;
  hipLaunchKernelGGL((k_small_mfma_u<PMODE, F8>), dim3(nb), dim3(64 * kMfmaWaves), 0, s, a, items);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_small_mfma_u(int n, int nbps, int np, int pmode, const SmallArgs &a, unsigned items,
                        hipStream_t s) {
  if (!small_mfma_ok(n, np) || nbps != 8 || items == 0 || a.gpw != 1) return 2;
  if (a.f8_out) {
    switch (pmode) {
      case 1: return go_small_mfma_u<1, true>(a, items, s);
      case 4: return go_small_mfma_u<4, true>(a, items, s);
      case -4: return go_small_mfma_u<-4, true>(a, items, s);
    }
    return 2;
  }
  switch (pmode) {
    case 1: return go_small_mfma_u<1, false>(a, items, s);
    case 4: return go_small_mfma_u<4, false>(a, items, s);
    case -4: return go_small_mfma_u<-4, false>(a, items, s);
  }
  return 2;
}

template <int NBPS, int PMODE>
static int go_small_mfma(const SmallArgs &a, unsigned items, hipStream_t s) {
  const unsigned nb = (items + kMfmaWaves - 1) / kMfmaWaves;
  hipLaunchKernelGGL((k_small_mfma<NBPS, PMODE>), dim3(nb), dim3(64 * kMfmaWaves), 0, s, a, items);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int small_mfma_ok(int n, int np) { return n == 1024 && np == 2; }

int launch_small_mfma(int n, int nbps, int np, int pmode, const SmallArgs &a, unsigned items,
                      hipStream_t s) {
  if (!small_mfma_ok(n, np) || items == 0 || a.gpw != 1) return 2;
  if (nbps == 8) {
    switch (pmode) {
      case 1: return go_small_mfma<8, 1>(a, items, s);
      case 4: return go_small_mfma<8, 4>(a, items, s);
      case -4: return go_small_mfma<8, -4>(a, items, s);
    }
  } else if (nbps == 16) {
    switch (pmode) {
      case 1: return go_small_mfma<16, 1>(a, items, s);
      case 4: return go_small_mfma<16, 4>(a, items, s);
      case -4: return go_small_mfma<16, -4>(a, items, s);
    }
  }
  return 2;
}

template <int PMODE>
static int go_b_mfma(const FsArgs &a, unsigned nb, hipStream_t s) {
  if (a.na_buf == 1)
    hipLaunchKernelGGL((k_fs_b_mfma<PMODE, true>), dim3(nb), dim3(64 * kMfmaWaves), 0, s, a);
  else
    hipLaunchKernelGGL((k_fs_b_mfma<PMODE, false>), dim3(nb), dim3(64 * kMfmaWaves), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

// k_fs_a_mfma: pass A of the four-step path on the matrix cores for
// N2 = 1024 (N1 = R1 * 1024; launched for R1 = 1), two pols, int8 samples.
// k_fs_a's contract and items (a tile of 4 consecutive n1 per workgroup and
// spectrum, workgroups looping over the items); the tile (16 bytes per n2
// row) is loaded cooperatively into LDS, the next item's into registers
// meanwhile.  One wavefront transforms one n1 column: the 1024-point
// DFT over n2 as split-FP16 matrix products (dft1024: int8 input exact, the
// DFT matrix hi + lo in both stages, the stage-2 input hi + lo -- the error
// of an FP32 transform), then the four-step twiddle W_N^(n1 k2) from the same
// two LDS tables as k_fs_a, and Y in pass B's tiled layout [k2 / 4][n1][k2 %
// 4], as two-pol FP32 values or 24-bit floats (Y24), staged through LDS so
// that the workgroup's four columns leave as runs of 16 entries.
template <int R1, bool Y24>
__global__ __launch_bounds__(64 * kMfmaWaves, 2) void k_fs_a_mfma(FsArgs a) {
  constexpr int N2 = 1024, N1C = R1 * N2, L1 = ilog2(N1C), TA = kMfmaWaves, WG = 64 * kMfmaWaves;
  constexpr int RPT = N2 / WG;         // tile rows each thread loads
  static_assert(RPT == 4, "the tile prefetch holds four rows per thread in r0..r3");
  using S = Sample<8, 2>;
  __shared__ Dft1024K K;
  __shared__ float2 twl[N2];           // W_N2^h, h < N2
  __shared__ float2 twr[N1C];          // W_N^r, r < N1
  // the item's samples: [n2][4 consecutive n1] words
  __shared__ __attribute__((aligned(16))) unsigned int tin[N2 * TA];
  // output staging of one quarter of an item: [wavefront][group][4 entries]
  __shared__ __attribute__((aligned(16))) unsigned int st[kMfmaWaves * 64 * 4 * (Y24 ? 3 : 4)];
  for (int m = threadIdx.x; m < N2; m += WG) twl[m] = a.tw[m * (kTwN / N2)];
  for (int m = threadIdx.x; m < N1C; m += WG) twr[m] = a.twf[m];
  dft1024_consts(K, a.tw, WG);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, h = lane >> 5;
  const unsigned N1 = N1C;
  const unsigned long long N = (unsigned long long)N1 * N2;
  const unsigned tiles = N1 / TA;
  const unsigned nsp = a.Nss * a.ynb;
  const unsigned items = tiles * a.Nss * a.nbat * a.nch;

  // the item's tile, 16 bytes (4 n1 of one n2 row) per thread and row:
  // thread t loads rows t + WG q, q < RPT
  auto tile_src = [&](unsigned l) -> const unsigned int * {
    const unsigned tile = l % tiles;
    const unsigned cs = l / tiles;
    const unsigned sp = cs % a.Nss;
    const unsigned cb = cs / a.Nss;
    const unsigned b = cb % a.nbat;
    const unsigned cl = cb / a.nbat;
    return reinterpret_cast<const unsigned int *>(a.in[b] + (unsigned long long)(a.c0 + cl) * a.chan_stride) +
           (unsigned long long)sp * N + tile * TA + (unsigned long long)threadIdx.x * N1;
  };
  unsigned l = xcd_remap(blockIdx.x, gridDim.x);
  uint4 r0, r1, r2, r3;
  {
    const unsigned int *p = tile_src(l);
    r0 = *reinterpret_cast<const uint4 *>(p);
    r1 = *reinterpret_cast<const uint4 *>(p + (unsigned long long)WG * N1);
    r2 = *reinterpret_cast<const uint4 *>(p + (unsigned long long)2 * WG * N1);
    r3 = *reinterpret_cast<const uint4 *>(p + (unsigned long long)3 * WG * N1);
  }
  for (;;) {
    // the tile into LDS (the previous item's readers are past the last
    // barrier of its store loop)
    reinterpret_cast<uint4 *>(tin)[threadIdx.x] = r0;
    reinterpret_cast<uint4 *>(tin)[threadIdx.x + WG] = r1;
    reinterpret_cast<uint4 *>(tin)[threadIdx.x + 2 * WG] = r2;
    reinterpret_cast<uint4 *>(tin)[threadIdx.x + 3 * WG] = r3;
    __syncthreads();                     // also orders the constants before first use
    const unsigned ln = l + gridDim.x;   // uniform across the workgroup
    if (ln < items) {
      const unsigned int *p = tile_src(ln);
      r0 = *reinterpret_cast<const uint4 *>(p);
      r1 = *reinterpret_cast<const uint4 *>(p + (unsigned long long)WG * N1);
      r2 = *reinterpret_cast<const uint4 *>(p + (unsigned long long)2 * WG * N1);
      r3 = *reinterpret_cast<const uint4 *>(p + (unsigned long long)3 * WG * N1);
    }
    f16v xr[2], xi[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      float2 x[4][4];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        float2 o[2];
        S::cvt(tin[(col + 32 * (8 * (e >> 2) + 4 * h + (e & 3))) * TA + wv], o, a.isign);
        x[e >> 2][e & 3] = o[p];
      }
      dft1024<false>(x, K, xr[p], xi[p]);   // X / 32
    }
    const unsigned tile = l % tiles;
    const unsigned cs = l / tiles;
    const unsigned sp = cs % a.Nss;
    const unsigned cb = cs / a.Nss;
    const unsigned b = cb % a.nbat;
    const unsigned cl = cb / a.nbat;
    const unsigned n1 = tile * TA + wv;
    const unsigned long long spec = (unsigned long long)cl * nsp + (a.yb0 + b) * a.Nss + sp;
    // Y entries ((k2 / 4) N1 + n1) * 4 + k2 % 4.  A lane holds k2 = 8 rq +
    // 4 h + i + 32 col (i < 4) of its column for rq < 4; per rq the four
    // wavefronts' values (4 consecutive n1) go through LDS so that 4 threads
    // write one 16-entry run (k2 / 4 fixed, n1 = tile * 4 .. + 3, k2 % 4 =
    // 0 .. 3): the run's 16-byte chunks c = w + 4 q go to thread w at store
    // q, so each store instruction covers 64 contiguous bytes of every run
    // (non-temporal stores with gaps reach HBM as separate partial writes).
    constexpr int EW = Y24 ? 3 : 4;      // words per entry
    // Four-step twiddle W_N^(n1 k2), k2 = (8 rq + 4 h + i) + 32 col, as
    // W_N^(n1 (8 rq + 4 h + i)) (two distinct table entries per instruction:
    // a broadcast) times this lane's W_N^(32 n1 col), looked up once per
    // item.  Looking up W_N^(n1 k2) directly puts the 32 lanes of a
    // half-wave 32 n1 entries apart in the tables: one LDS bank, a 32-way
    // conflict per lookup.
    float2 wc;
    {
      const unsigned e2 = 32u * n1 * (unsigned)col;   // < 2^20 * R1
      wc = cmul(twl[e2 >> L1], twr[e2 & (N1C - 1)]);
      wc.x *= 32.0f;                     // undo the transform's 1/32
      wc.y *= 32.0f;
    }
#pragma unroll
    for (int rq = 0; rq < 4; ++rq) {
      unsigned int *sw = st + ((wv * 64 + 2 * col + h) * 4) * EW;   // group g = 2 col + h
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * rq + i;
        const unsigned e1 = n1 * (unsigned)(8 * rq + 4 * h + i);  // < 32 N1
        const float2 w = cmul(cmul(twl[e1 >> L1], twr[e1 & (N1C - 1)]), wc);
        const float2 y0 = cmul(make_float2(xr[0][r], xi[0][r]), w);
        const float2 y1 = cmul(make_float2(xr[1][r], xi[1][r]), w);
        if constexpr (Y24) {
          const unsigned int b0 = round24(y0.x), b1 = round24(y0.y), b2 = round24(y1.x),
                             b3 = round24(y1.y);
          sw[3 * i] = __builtin_amdgcn_perm(b1, b0, 0x05030201u);
          sw[3 * i + 1] = __builtin_amdgcn_perm(b2, b1, 0x06050302u);
          sw[3 * i + 2] = __builtin_amdgcn_perm(b3, b2, 0x07060503u);
        } else {
          reinterpret_cast<float4 *>(sw)[i] = make_float4(y0.x, y0.y, y1.x, y1.y);
        }
      }
      __syncthreads();
      {
        const int t = threadIdx.x, g = t >> 2, w = t & 3;          // group, thread within the group
        const unsigned k2q = 2 * rq + (g & 1) + 8 * (g >> 1);       // k2 / 4
        // the run's first entry (n1 = tile * 4, k2 % 4 = 0); it is 4 slots
        // (n1) of EW chunks
        const unsigned long long ent = spec * N + ((unsigned long long)k2q * N1C + tile * TA) * 4;
        unsigned int *dst = reinterpret_cast<unsigned int *>(a.work) + ent * EW;
#pragma unroll
        for (int q = 0; q < EW; ++q) {
          const int c = w + 4 * q, slot = c / EW, part = c % EW;
          const f32x4 v = *reinterpret_cast<const f32x4 *>(st + ((slot * 64 + g) * 4) * EW + 4 * part);
          __builtin_nontemporal_store(v, reinterpret_cast<f32x4 *>(dst + 4 * c));
        }
      }
      __syncthreads();
    }
    if (ln >= items) break;
    l = ln;
  }
}

// N1 = 1024 only: pass B's k2 tile is then 4, the run this kernel writes
// (N1 = 2048 reads 2-wide tiles)
int fs_a_mfma_ok(int n1, int n2, int nbps, int np) {
  return n2 == 1024 && n1 == 1024 && nbps == 8 && np == 2;
}

int launch_fs_a_mfma(int nbps, int np, const FsArgs &a, hipStream_t s) {
  if (!fs_a_mfma_ok((int)a.N1, (int)a.N2, nbps, np) || a.tby != (unsigned)kMfmaWaves ||
      a.nbat == 0 || a.nbat > (unsigned)kMaxBatch || a.yb0 + a.nbat > a.ynb)
    return 2;
  const unsigned long long items = (unsigned long long)(a.N1 / kMfmaWaves) * a.Nss * a.nbat * a.nch;
  const unsigned long long nb = a.grid_a && a.grid_a < items ? a.grid_a : items;
  if (nb == 0 || items > 0x7fffffffull) return 2;
  const dim3 g((unsigned)nb), t(64 * kMfmaWaves);
  if (a.y24) hipLaunchKernelGGL((k_fs_a_mfma<1, true>), g, t, 0, s, a);
  else hipLaunchKernelGGL((k_fs_a_mfma<1, false>), g, t, 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int fs_b_mfma_ok(int n1, int np) { return n1 == 1024 && np == 2; }

int launch_fs_b_mfma(int np, int pmode, const FsArgs &a, hipStream_t s) {
  // one workgroup per (k2 tile, group, channel) integrates the whole group
  if (!fs_b_mfma_ok((int)a.N1, np) || a.y24 || a.N2 % kMfmaWaves != 0 || !(a.mscale > 0.0f) ||
      a.nbat == 0 || a.nbat > a.ynb || a.tby != (unsigned)kMfmaWaves || a.ns != 1)
    return 2;
  const unsigned long long nb = (unsigned long long)(a.N2 / kMfmaWaves) * a.Ng * a.nch;
  if (nb == 0 || nb > 0x7fffffffull) return 2;
  switch (pmode) {
    case 1: return go_b_mfma<1>(a, (unsigned)nb, s);
    case 4: return go_b_mfma<4>(a, (unsigned)nb, s);
    case -4: return go_b_mfma<-4>(a, (unsigned)nb, s);
  }
  return 2;
}

}  // namespace rsk
