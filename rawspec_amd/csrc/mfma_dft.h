// mfma_dft.h -- 1024-point DFTs on the gfx950 matrix cores, one wavefront
// per transform, no LDS and no barriers.
//
// N = 1024 = 32 x 32 with n = a + 32 b and k = c + 32 d:
//   stage 1   Z[a][c] = sum_b x[a + 32 b] W_32^(b c)          (GEMM, K = b)
//   twiddle   Z[a][c] *= W_1024^(a c)
//   stage 2   X[c + 32 d] = sum_a Z[a][c] W_32^(a d)          (GEMM, K = a)
// Each stage is a complex 32 x 32 x 32 product done as real
// v_mfma_f32_32x32x16_f16 instructions: complex values are interleaved
// (re, im) along K, and the 32-point DFT matrix is laid out as two real
// 64 x 32 operands (real part and imaginary part of the output).
//
// Lane layout (l = lane, h = l / 32), verified on gfx950:
//   A operand  row l % 32, K elements 8 h .. 8 h + 7 of the K-block
//   B operand  col l % 32, same K elements
//   C result   col l % 32, register r holds row 8 (r / 4) + 4 h + r % 4
// So stage 1's result (col c = l % 32, rows a in registers) is already
// stage 2's A operand (row c = l % 32, K = a): register r of the result is
// element a = 8 (r / 4) + 4 h + r % 4 of K-block r / 4.  Input: lane l holds
// samples n = l % 32 + 32 (8 kb + 4 h + i), kb, i < 4.  Output: register r
// of lane l holds bin k = 8 (r / 4) + 4 h + r % 4 + 32 (l % 32).
//
// Precision: every FP16 operand is a hi + lo pair (hi = fp16(x),
// lo = fp16(x - hi)), the DFT matrix included, and each product is
// hi*hi + hi*lo + lo*hi accumulated in FP32 (the lo*lo term is below FP32
// rounding).  That keeps about 22 significant bits per operand, so the
// transform's error is at the level of an FP32 FFT.  Integer inputs that
// FP16 holds exactly (|x| <= 2048) skip their lo part.  Inputs are scaled by
// powers of two so that no FP16 operand can overflow; the caller undoes the
// scaling on the detected power.
#pragma once

#include <hip/hip_runtime.h>

namespace rsk {

typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Element e of (H, L) = hi/lo FP16 split of x (e compile-time after unrolling).
__device__ __forceinline__ void set_hl(h8v &H, h8v &L, int e, float x) {
  const _Float16 hi = (_Float16)x;
  H[e] = hi;
  L[e] = (_Float16)(x - (float)hi);
}

// Per-lane constants, kept in LDS and shared by the workgroup's wavefronts:
// B operands of the 32-point DFT matrix (hi and lo parts of the real-output
// and imaginary-output halves, 4 K-blocks) and the twiddles W_1024^(a c) of
// each lane's 16 stage-1 results.  Indexing [..][lane] makes every read
// one contiguous 1 KB (B) or 512 B (twiddle) wave access.
struct Dft1024K {
  h8v b[2][4][2][64];     // [re/im output][K-block][hi/lo][lane]
  float2 tw[16][64];      // [stage-1 result register][lane]
};

// Fills the constants (all threads of a workgroup of wg threads take part;
// the caller synchronises).  tw4096: W_4096^m, m < 4096 (FP64-generated).
__device__ __forceinline__ void dft1024_consts(Dft1024K &k, const float2 *tw4096, int wg) {
  for (int e = threadIdx.x; e < 2 * 4 * 64; e += wg) {
    const int lane = e & 63, kb = (e >> 6) & 3, t = e >> 8;
    const int col = lane & 31, h = lane >> 5;
    h8v H, L;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int b = 8 * kb + 4 * h + i;                 // K index (complex)
      const float2 w = tw4096[((b * col) & 31) * 128];  // W_32^(b col)
      // real output: re(x) w.x - im(x) w.y; imaginary output: re(x) w.y + im(x) w.x
      set_hl(H, L, 2 * i, t == 0 ? w.x : w.y);
      set_hl(H, L, 2 * i + 1, t == 0 ? -w.y : w.x);
    }
    k.b[t][kb][0][lane] = H;
    k.b[t][kb][1][lane] = L;
  }
  for (int e = threadIdx.x; e < 16 * 64; e += wg) {
    const int lane = e & 63, r = e >> 6;
    const int col = lane & 31, h = lane >> 5;
    const int a = 8 * (r >> 2) + 4 * h + (r & 3);
    k.tw[r][lane] = tw4096[((a * col) & 1023) * 4];     // W_1024^(a col)
  }
}

// One complex 32 x 32 x 32 stage: (cr, ci) = A x F32, A given as hi/lo
// FP16 operands per K-block.  LO = false when A is exact in FP16.
template <bool LO>
__device__ __forceinline__ void dft32_stage(const h8v (&ah)[4], const h8v (&al)[4],
                                            const Dft1024K &k, int lane, f16v &cr, f16v &ci) {
#pragma unroll
  for (int r = 0; r < 16; ++r) { cr[r] = 0.0f; ci[r] = 0.0f; }
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    const h8v brh = k.b[0][kb][0][lane], brl = k.b[0][kb][1][lane];
    const h8v bih = k.b[1][kb][0][lane], bil = k.b[1][kb][1][lane];
    if constexpr (LO) {
      cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[kb], brh, cr, 0, 0, 0);
      ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[kb], bih, ci, 0, 0, 0);
    }
    cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[kb], brl, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[kb], bil, ci, 0, 0, 0);
    cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[kb], brh, cr, 0, 0, 0);
    ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[kb], bih, ci, 0, 0, 0);
  }
}

// Full 1024-point DFT of one wavefront-held transform.  x[kb][i]: the
// lane's input samples (layout above), already scaled so |re|, |im| <=
// 32768.  Returns X = DFT(x) / 32 in (xr, xi) (register r = bin as above).
template <bool LO>
__device__ __forceinline__ void dft1024(const float2 (&x)[4][4], const Dft1024K &k, f16v &xr,
                                        f16v &xi) {
  // The constants are re-read from LDS for every transform: an opaque zero
  // in the index keeps the compiler from hoisting them out of the caller's
  // loop into 96 registers per lane, which would halve the occupancy.
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  const int lane = (threadIdx.x & 63) + z;
  h8v ah[4], al[4];
#pragma unroll
  for (int kb = 0; kb < 4; ++kb)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if constexpr (LO) {
        set_hl(ah[kb], al[kb], 2 * i, x[kb][i].x);
        set_hl(ah[kb], al[kb], 2 * i + 1, x[kb][i].y);
      } else {
        ah[kb][2 * i] = (_Float16)x[kb][i].x;
        ah[kb][2 * i + 1] = (_Float16)x[kb][i].y;
      }
    }
  f16v zr, zi;
  dft32_stage<LO>(ah, al, k, lane, zr, zi);
  // twiddle, scale by 1/32 (|Z| <= 2^20 -> <= 2^15), split for stage 2
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const float2 t = k.tw[r][lane];
    const float wr = (zr[r] * t.x - zi[r] * t.y) * 0.03125f;
    const float wi = (zr[r] * t.y + zi[r] * t.x) * 0.03125f;
    set_hl(ah[r >> 2], al[r >> 2], 2 * (r & 3), wr);
    set_hl(ah[r >> 2], al[r >> 2], 2 * (r & 3) + 1, wi);
  }
  dft32_stage<true>(ah, al, k, lane, xr, xi);
}

// ---- unsplit variant ------------------------------------------------------
// The same two-stage DFT with every operand a single FP16 value: the 32-point
// DFT matrix rounded once, int8 samples (exact in FP16) as they are, and the
// twiddled stage-1 result rounded once to FP16 before stage 2 (FP32
// accumulation in both stages).  A third of the matrix instructions of the
// split form and no split arithmetic; each transform carries about 2^-12
// relative rounding error, random from spectrum to spectrum except for the
// matrix's own rounding, so an integration of K spectra has a worst bin error
// of roughly 6e-3 / sqrt(K) over a floor near 1.5e-4 (measured on a numpy
// emulation of these roundings with int8 noise and a strong tone).  Inputs are
// used unscaled: |stage-1 result| <= 32 * 127 * sqrt(2) < 2^13.
typedef _Float16 h2v __attribute__((ext_vector_type(2)));

// Constants of the unsplit transform: B operands of the 32-point DFT matrix
// (real-output and imaginary-output halves, 4 K-blocks) and the twiddles
// W_1024^(a c) of each lane's 16 stage-1 results.  sg = -1 conjugates every
// constant, which makes the transform the inverse DFT (unnormalised).
struct Dft1024U {
  h8v b[2][4][64];        // [re/im output][K-block][lane]
  float2 tw[16][64];      // [stage-1 result register][lane]
};

__device__ __forceinline__ void dft1024u_consts(Dft1024U &k, const float2 *tw4096, int wg,
                                                float sg) {
  for (int e = threadIdx.x; e < 2 * 4 * 64; e += wg) {
    const int lane = e & 63, kb = (e >> 6) & 3, t = e >> 8;
    const int col = lane & 31, h = lane >> 5;
    h8v H;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int b = 8 * kb + 4 * h + i;
      const float2 w = tw4096[((b * col) & 31) * 128];
      const float wx = w.x, wy = sg * w.y;
      H[2 * i] = (_Float16)(t == 0 ? wx : wy);
      H[2 * i + 1] = (_Float16)(t == 0 ? -wy : wx);
    }
    k.b[t][kb][lane] = H;
  }
  for (int e = threadIdx.x; e < 16 * 64; e += wg) {
    const int lane = e & 63, r = e >> 6;
    const int col = lane & 31, h = lane >> 5;
    const int a = 8 * (r >> 2) + 4 * h + (r & 3);
    const float2 w = tw4096[((a * col) & 1023) * 4];
    k.tw[r][lane] = make_float2(w.x, sg * w.y);
  }
}

// One unsplit 1024-point DFT.  ah: the lane's input samples as FP16 (re, im)
// pairs (layout as dft1024); bre/bim: the B operands; tw: the twiddle table
// (Dft1024U::tw, read at [r][lane]).
// Returns X = DFT(x) in (xr, xi), register r = bin 8 (r / 4) + 4 h + r % 4 +
// 32 (l % 32).
__device__ __forceinline__ void dft1024u(const h8v (&ah)[4], const h8v (&bre)[4],
                                         const h8v (&bim)[4], const float2 (*tw)[64], int lane,
                                         f16v &xr, f16v &xi) {
  f16v zr, zi;
#pragma unroll
  for (int r = 0; r < 16; ++r) { zr[r] = 0.0f; zi[r] = 0.0f; }
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    zr = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[kb], bre[kb], zr, 0, 0, 0);
    zi = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[kb], bim[kb], zi, 0, 0, 0);
  }
  h8v a2[4];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const float2 t = tw[r][lane];
    const float wr = zr[r] * t.x - zi[r] * t.y;
    const float wi = zr[r] * t.y + zi[r] * t.x;
    a2[r >> 2][2 * (r & 3)] = (_Float16)wr;
    a2[r >> 2][2 * (r & 3) + 1] = (_Float16)wi;
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) { xr[r] = 0.0f; xi[r] = 0.0f; }
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) {
    xr = __builtin_amdgcn_mfma_f32_32x32x16_f16(a2[kb], bre[kb], xr, 0, 0, 0);
    xi = __builtin_amdgcn_mfma_f32_32x32x16_f16(a2[kb], bim[kb], xi, 0, 0, 0);
  }
}

// int8 sample word (pol 0 re, im, pol 1 re, im) -> the (re, im) FP16 pair of
// pol p, exact, with the SNORM clamp of -128 to -127: the bytes, offset by
// 128, become the low bytes of FP16 values 1024 + u (exponent byte 0x64,
// unit spacing), and 1152 is subtracted.
template <int P>
__device__ __forceinline__ unsigned int i8_pair_f16(unsigned int t) {   // t = word ^ 0x80808080
  const unsigned int sel = P == 0 ? 0x04010400u : 0x04030402u;
  h2v v = __builtin_bit_cast(h2v, __builtin_amdgcn_perm(0x64646464u, t, sel));
  v = v - (h2v){(_Float16)1152.0f, (_Float16)1152.0f};
  v = __builtin_elementwise_max(v, (h2v){(_Float16)-127.0f, (_Float16)-127.0f});
  return __builtin_bit_cast(unsigned int, v);
}

}  // namespace rsk
