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

// k_small_mfma_u: k_small_mfma's contract for int8 samples with unsplit
// FP16 operands (mfma_dft.h, dft1024u): the samples become FP16 pairs by a
// byte permute (no FP32 conversion), the inverse transform uses conjugated
// constants (no sign flips per sample), detection accumulates raw sums with
// fused multiply-adds and forms the components once per output.  The DFT
// matrix (32 registers) stays in registers for the whole chunk, the
// twiddles are read from LDS per transform; the samples are loaded two
// spectra ahead of the transforms.
template <int PMODE>
__global__ __launch_bounds__(64 * kMfmaWaves, 2) void k_small_mfma_u(SmallArgs a, unsigned items) {
  constexpr int N = 1024;
  constexpr int Q = nq<PMODE>();
  constexpr int kStBytes = kMfmaWaves * Q * kSmRow * 4;
  constexpr int kLds = kStBytes > (int)sizeof(Dft1024U) ? kStBytes : (int)sizeof(Dft1024U);
  __shared__ __attribute__((aligned(16))) unsigned char lds_raw[kLds];
  Dft1024U &K = *reinterpret_cast<Dft1024U *>(lds_raw);
  float *st = reinterpret_cast<float *>(lds_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int col = lane & 31, h = lane >> 5;
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
  h8v bre[4], bim[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    bre[kb] = K.b[0][kb][lane];
    bim[kb] = K.b[1][kb][lane];
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
  // of the transforms: one spectrum's compute does not cover HBM latency)
  unsigned int raw[16], nxt[16];
  if (nspec) {
#pragma unroll
    for (int e = 0; e < 16; ++e) raw[e] = src[32 * (8 * (e >> 2) + (e & 3))];
  }
  if (nspec > 1) {
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
      const h2v p0 = __builtin_bit_cast(h2v, i8_pair_f16<0>(t));
      const h2v p1 = __builtin_bit_cast(h2v, i8_pair_f16<1>(t));
      ah[0][e >> 2][2 * (e & 3)] = p0[0];
      ah[0][e >> 2][2 * (e & 3) + 1] = p0[1];
      ah[1][e >> 2][2 * (e & 3)] = p1[0];
      ah[1][e >> 2][2 * (e & 3) + 1] = p1[1];
    }
#pragma unroll
    for (int e = 0; e < 16; ++e) raw[e] = nxt[e];
    if (s + 2 < nspec) {
      const unsigned int *nx = src + (unsigned long long)(s + 2) * N;
#pragma unroll
      for (int e = 0; e < 16; ++e) nxt[e] = nx[32 * (8 * (e >> 2) + (e & 3))];
    }
    // the twiddles are read from LDS per transform (an opaque zero in the
    // index keeps the compiler from holding all 32 registers of them)
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
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

template <int PMODE>
static int go_small_mfma_u(const SmallArgs &a, unsigned items, hipStream_t s) {
  const unsigned nb = (items + kMfmaWaves - 1) / kMfmaWaves;
  hipLaunchKernelGGL((k_small_mfma_u<PMODE>), dim3(nb), dim3(64 * kMfmaWaves), 0, s, a, items);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_small_mfma_u(int n, int nbps, int np, int pmode, const SmallArgs &a, unsigned items,
                        hipStream_t s) {
  if (!small_mfma_ok(n, np) || nbps != 8 || items == 0 || a.gpw != 1) return 2;
  switch (pmode) {
    case 1: return go_small_mfma_u<1>(a, items, s);
    case 4: return go_small_mfma_u<4>(a, items, s);
    case -4: return go_small_mfma_u<-4>(a, items, s);
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
