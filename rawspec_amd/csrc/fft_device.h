// fft_device.h -- device-side building blocks for the gfx950 spectrometer
// kernels: radix-2/4/8/16 butterflies, an LDS-exchanged Stockham FFT executed
// by a "team" of threads, sample unpacking and polarisation detection.
//
// Conventions used by every kernel:
//  * All transforms are computed with the forward kernel exp(-2 pi i k m / N).
//    The inverse transform rawspec asks for (fft_dir <= 0) is obtained as
//    conj(FFT(conj(x))): the unpacker negates the imaginary input and the
//    detector negates the imaginary cross term (powers are unchanged).
//  * Samples are transformed as raw integers (after the SNORM clamp of the
//    most negative code); the 1/127^2 or 1/32767^2 normalisation of the
//    texture read (rawspec_gpu.cu:845-863) is applied once to the detected
//    sums.
//  * A team of T threads holds an N = T*E point transform; thread j owns
//    elements j + m*T (m < E) on entry and frequency bins j + m*T on exit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace rsk {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int ilog2(int n) { return n <= 1 ? 0 : 1 + ilog2(n / 2); }

// Multiplies a by W_R^k = exp(-2 pi i k / R), R <= 16, k compile-time after
// unrolling.  Exact special cases for multiples of pi/4.
template <int R>
__device__ __forceinline__ float2 twc(float2 a, int k) {
  constexpr float R2 = 0.70710678118654752440f;   // cos(pi/4)
  constexpr float C8 = 0.92387953251128675613f;   // cos(pi/8)
  constexpr float S8 = 0.38268343236508977173f;   // sin(pi/8)
  const int e = (k * (16 / R)) & 15;              // angle in units of 2 pi / 16
  switch (e) {
    case 0: return a;
    case 4: return make_float2(a.y, -a.x);
    case 8: return make_float2(-a.x, -a.y);
    case 12: return make_float2(-a.y, a.x);
    case 2: return make_float2((a.x + a.y) * R2, (a.y - a.x) * R2);
    case 6: return make_float2((a.y - a.x) * R2, -(a.x + a.y) * R2);
    case 10: return make_float2((-a.x - a.y) * R2, (a.x - a.y) * R2);
    case 14: return make_float2((a.x - a.y) * R2, (a.x + a.y) * R2);
    default: {
      // general angle: exp(-i t) * (x + i y) = (x c + y s) + i (y c - x s)
      float c, s;
      switch (e) {
        case 1: c = C8; s = S8; break;
        case 3: c = S8; s = C8; break;
        case 5: c = -S8; s = C8; break;
        case 7: c = -C8; s = S8; break;
        case 9: c = -C8; s = -S8; break;
        case 11: c = -S8; s = -C8; break;
        case 13: c = S8; s = -C8; break;
        default: c = C8; s = -S8; break;   // 15
      }
      return make_float2(a.x * c + a.y * s, a.y * c - a.x * s);
    }
  }
}

// In-place forward DFT of R points held in registers (radix-2 recursion,
// fully unrolled).
template <int R>
struct Dft {
  __device__ __forceinline__ static void run(float2 (&a)[R]) {
    float2 e[R / 2], o[R / 2];
#pragma unroll
    for (int k = 0; k < R / 2; ++k) { e[k] = a[2 * k]; o[k] = a[2 * k + 1]; }
    Dft<R / 2>::run(e);
    Dft<R / 2>::run(o);
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      float2 t = twc<R>(o[k], k);
      a[k] = cadd(e[k], t);
      a[k + R / 2] = csub(e[k], t);
    }
  }
};
template <>
struct Dft<1> {
  __device__ __forceinline__ static void run(float2 (&)[1]) {}
};
template <>
struct Dft<2> {
  __device__ __forceinline__ static void run(float2 (&a)[2]) {
    float2 t = a[1];
    a[1] = csub(a[0], t);
    a[0] = cadd(a[0], t);
  }
};

// LDS exchange layout: element `pos` of polarisation p of a team lives at
// p*lds_len<N>() + lds_pos(pos).  With RSK_LDS_PAD = s > 0 one float2 of
// padding follows every 2^s elements, which spreads the strided Stockham
// writes over the banks.
#ifndef RSK_LDS_PAD
#define RSK_LDS_PAD 5
#endif
__device__ __forceinline__ int lds_pos(int pos) {
  if constexpr (RSK_LDS_PAD > 0) return pos + (pos >> RSK_LDS_PAD);
  else return pos;
}
template <int N>
constexpr int lds_len() { return RSK_LDS_PAD > 0 ? N + (N >> RSK_LDS_PAD) : N; }

// Distance in float2 between the LDS regions of consecutive teams of a
// workgroup; RSK_TEAM_PAD extra float2 keep teams that share a wave off the
// same banks.
#ifndef RSK_TEAM_PAD
#define RSK_TEAM_PAD 8
#endif
// RSK_SEQ_POL = 1: the polarisations go through the LDS exchange one after
// the other (half the LDS per team, twice the barriers).
#ifndef RSK_SEQ_POL
#define RSK_SEQ_POL 1
#endif
template <int N, int NP>
constexpr int team_stride() { return (RSK_SEQ_POL ? 1 : NP) * lds_len<N>() + RSK_TEAM_PAD; }

// Synchronises a team between the write and the read half of an LDS
// exchange.  A team whose threads are consecutive lanes of one wavefront
// (WAVE) needs no workgroup barrier: a wavefront's LDS instructions complete
// in issue order, so only the compiler must keep the accesses on their side
// of this point.  Otherwise every thread of the workgroup waits here.
template <bool WAVE>
__device__ __forceinline__ void team_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// One Stockham stage of radix R on an N = T*E point team transform, for NP
// interleaved polarisations.  Butterfly b = j + u*T (u < E/R) reads elements
// b + r*N/R, which are the thread's own slots u + r*E/R; it writes element r
// to (b / NS)*NS*R + (b % NS) + r*NS.  The last stage (NS*R == N) writes back
// into the same register slots, so no exchange follows it.  tw holds
// W_TWN^m for m < TWN (TWN a multiple of N).
template <int N, int E, int NP, int R, int NS, bool LAST, int TWN, bool WAVE>
__device__ __forceinline__ void stage(float2 (&v)[NP][E], float2 *lds, int j,
                                      const float2 *__restrict__ tw) {
  constexpr int T = N / E;
  constexpr int NB = E / R;
  constexpr int LL = lds_len<N>();
  if constexpr (RSK_SEQ_POL && !LAST && NP > 1) {
    // one polarisation at a time through a single LDS region
#pragma unroll
    for (int p = 0; p < NP; ++p) {
#pragma unroll
      for (int u = 0; u < NB; ++u) {
        const int b = j + u * T;
        const int k = b & (NS - 1);
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = v[p][u + r * NB];
        if constexpr (NS > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) a[r] = cmul(a[r], tw[(r * k) * (TWN / (NS * R))]);
        }
        Dft<R>::run(a);
        const int base = (b / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lds_pos(base + r * NS)] = a[r];
      }
      team_sync<WAVE>();
#pragma unroll
      for (int m = 0; m < E; ++m) v[p][m] = lds[lds_pos(j + m * T)];
      team_sync<WAVE>();
    }
    return;
  }
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int b = j + u * T;
    const int k = b & (NS - 1);
    float2 w[R];
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) w[r] = tw[(r * k) * (TWN / (NS * R))];
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
      float2 a[R];
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = v[p][u + r * NB];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) a[r] = cmul(a[r], w[r]);
      }
      Dft<R>::run(a);
      if constexpr (LAST) {
#pragma unroll
        for (int r = 0; r < R; ++r) v[p][u + r * NB] = a[r];
      } else {
        const int base = (b / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[p * LL + lds_pos(base + r * NS)] = a[r];
      }
    }
  }
  if constexpr (!LAST) {
    team_sync<WAVE>();
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int m = 0; m < E; ++m) v[p][m] = lds[p * LL + lds_pos(j + m * T)];
    team_sync<WAVE>();
  }
}

template <int N, int E, int NP, int NS, int TWN, bool WAVE>
struct Stages {
  __device__ __forceinline__ static void run(float2 (&v)[NP][E], float2 *lds, int j,
                                             const float2 *__restrict__ tw) {
    if constexpr (NS * E == N) {
      stage<N, E, NP, E, NS, true, TWN, WAVE>(v, lds, j, tw);
    } else {
      stage<N, E, NP, E, NS, false, TWN, WAVE>(v, lds, j, tw);
      Stages<N, E, NP, NS * E, TWN, WAVE>::run(v, lds, j, tw);
    }
  }
};

// Forward FFT of N points (power of two, E | N, N <= 4096) for NP
// polarisations held by a team of N/E threads.  lds: NP*lds_len<N>() float2
// private to the team.  tw: W_TWN^m table (global or LDS).  WAVE: the team is
// N/E consecutive lanes of one wavefront (N/E divides 64), so its exchanges
// need no workgroup barrier; otherwise every thread of the workgroup must call
// it (barriers inside).
template <int N, int E, int NP, int TWN = kTwN, bool WAVE = false>
__device__ __forceinline__ void team_fft(float2 (&v)[NP][E], float2 *lds, int j,
                                         const float2 *__restrict__ tw) {
  static_assert(!WAVE || (64 % (N / E) == 0), "a wave-synchronised team must fit in a wavefront");
  if constexpr (N > 1) {
    constexpr int R0 = 1 << (ilog2(N) % ilog2(E));
    if constexpr (R0 > 1) {
      if constexpr (R0 == N) {
        stage<N, E, NP, R0, 1, true, TWN, WAVE>(v, lds, j, tw);
      } else {
        stage<N, E, NP, R0, 1, false, TWN, WAVE>(v, lds, j, tw);
        Stages<N, E, NP, R0, TWN, WAVE>::run(v, lds, j, tw);
      }
    } else {
      Stages<N, E, NP, 1, TWN, WAVE>::run(v, lds, j, tw);
    }
  }
}

// Twiddle source of a team FFT of N points.  With RSK_LDS_TW the workgroup
// copies W_N^m (m < N) out of the master W_4096 table into LDS once, and the
// stages read it from there; otherwise they read the master table in global
// memory.
#ifndef RSK_LDS_TW
#define RSK_LDS_TW 1
#endif
template <int N>
struct TwTab {
  static constexpr int LEN = (RSK_LDS_TW && N > 1) ? N : 1;
  static constexpr int TWN = RSK_LDS_TW ? (N > 1 ? N : 1) : kTwN;
};
template <int N, int WG>
__device__ __forceinline__ const float2 *tw_setup(float2 *twl, const float2 *tw) {
  if constexpr (RSK_LDS_TW && N > 1) {
    for (int m = threadIdx.x; m < N; m += WG) twl[m] = tw[m * (kTwN / N)];
    __syncthreads();
    return twl;
  } else {
    return tw;
  }
}

// Elements per thread for a team transform of N points (compile-time
// tunables; 16 is the largest radix the butterflies implement).
#ifndef RSK_E_MID
#define RSK_E_MID 8
#endif
#ifndef RSK_E_BIG
#define RSK_E_BIG 16
#endif
constexpr int elems_for(int n) {
  return n <= 16 ? n : (n <= 128 ? RSK_E_MID : RSK_E_BIG);
}
// Elements per thread of the four-step passes' sub-transforms.
#ifndef RSK_E_FS
#define RSK_E_FS 16
#endif
constexpr int elems_fs(int n) { return n <= 128 ? 8 : RSK_E_FS; }

// ---- sample unpacking ----------------------------------------------------
// Raw sample of NP pols with NBPS-bit components -> NP complex floats
// (integer valued), most negative code clamped to -(2^(NBPS-1)-1) as the
// SNORM read does, imaginary part negated when inv != 0.
template <int NBPS, int NP>
struct Sample;

template <>
struct Sample<8, 2> {
  using word = uint32_t;
  __device__ __forceinline__ static void cvt(word w, float2 (&o)[2], float isign) {
    int a = (int)(int8_t)(w & 0xff), b = (int)(int8_t)((w >> 8) & 0xff);
    int c = (int)(int8_t)((w >> 16) & 0xff), d = (int)(int8_t)(w >> 24);
    a = max(a, -127); b = max(b, -127); c = max(c, -127); d = max(d, -127);
    o[0] = make_float2((float)a, isign * (float)b);
    o[1] = make_float2((float)c, isign * (float)d);
  }
  __device__ __forceinline__ static void cvt_offset(word w, float2 (&o)[2], float isign);
};
// int8 samples offset by 128 (bytes XOR 0x80, so each converts with one
// unsigned-byte conversion), clamped at 1 (the SNORM clamp of -128); the
// offset 128 (1 + i isign) is removed from bin 0 after the transform.  Exact
// for single-stage transforms (N <= 16): on integers every butterfly
// difference cancels the offset exactly, so only bin 0 carries it.
__device__ __forceinline__ float u8f(unsigned int t, int byte) {
  return fmaxf((float)((t >> (8 * byte)) & 0xffu), 1.0f);
}
template <>
struct Sample<8, 1> {
  using word = uint16_t;
  __device__ __forceinline__ static void cvt(word w, float2 (&o)[1], float isign) {
    int a = (int)(int8_t)(w & 0xff), b = (int)(int8_t)(w >> 8);
    a = max(a, -127); b = max(b, -127);
    o[0] = make_float2((float)a, isign * (float)b);
  }
  __device__ __forceinline__ static void cvt_offset(word w, float2 (&o)[1], float isign) {
    const unsigned int t = (unsigned int)w ^ 0x8080u;
    o[0] = make_float2(u8f(t, 0), isign * u8f(t, 1));
  }
};
__device__ __forceinline__ void Sample<8, 2>::cvt_offset(word w, float2 (&o)[2], float isign) {
  const unsigned int t = w ^ 0x80808080u;
  o[0] = make_float2(u8f(t, 0), isign * u8f(t, 1));
  o[1] = make_float2(u8f(t, 2), isign * u8f(t, 3));
}
template <>
struct Sample<16, 2> {
  using word = uint2;
  __device__ __forceinline__ static void cvt(word w, float2 (&o)[2], float isign) {
    int a = (int)(int16_t)(w.x & 0xffff), b = (int)(int16_t)(w.x >> 16);
    int c = (int)(int16_t)(w.y & 0xffff), d = (int)(int16_t)(w.y >> 16);
    a = max(a, -32767); b = max(b, -32767); c = max(c, -32767); d = max(d, -32767);
    o[0] = make_float2((float)a, isign * (float)b);
    o[1] = make_float2((float)c, isign * (float)d);
  }
};
template <>
struct Sample<16, 1> {
  using word = uint32_t;
  __device__ __forceinline__ static void cvt(word w, float2 (&o)[1], float isign) {
    int a = (int)(int16_t)(w & 0xffff), b = (int)(int16_t)(w >> 16);
    a = max(a, -32767); b = max(b, -32767);
    o[0] = make_float2((float)a, isign * (float)b);
  }
};

// ---- detection -------------------------------------------------------------
// PMODE 1: total power (sum over pols); 4: XX, YY, Re XY*, Im XY*;
// -4: I, Q, U, V.  gam = +-1 is the sign applied to the imaginary cross term
// (input_conjugated and the inverse-via-conjugation trick combined).
template <int PMODE>
constexpr int nq() { return PMODE == 1 ? 1 : 4; }

// Detection into raw sums, fused multiply-adds only: PMODE 1 accumulates the
// total power; 4 and -4 accumulate XX, YY, Re X0 X1* and Im X0 X1* (no sign
// applied), and detect_finish turns the sums into the product's components
// (-4: I = XX + YY, Q = XX - YY; both: the sign gam on the imaginary cross
// term).  The finishing step is linear, so it may be applied to any sum of
// raw sums.
template <int PMODE>
__device__ __forceinline__ void detect_raw(float (&acc)[nq<PMODE>()], float xr0, float xi0,
                                           float xr1, float xi1) {
  if constexpr (PMODE == 1) {
    acc[0] = fmaf(xi1, xi1, fmaf(xr1, xr1, fmaf(xi0, xi0, fmaf(xr0, xr0, acc[0]))));
  } else {
    acc[0] = fmaf(xi0, xi0, fmaf(xr0, xr0, acc[0]));
    acc[1] = fmaf(xi1, xi1, fmaf(xr1, xr1, acc[1]));
    acc[2] = fmaf(xi0, xi1, fmaf(xr0, xr1, acc[2]));
    acc[3] = fmaf(-xr0, xi1, fmaf(xi0, xr1, acc[3]));
  }
}
// The same for NP polarisations held as complex values (NP = 1: power).
template <int PMODE, int NP>
__device__ __forceinline__ void detect_raw_np(float (&acc)[nq<PMODE>()], const float2 &x0,
                                              const float2 &x1) {
  if constexpr (NP == 2) {
    detect_raw<PMODE>(acc, x0.x, x0.y, x1.x, x1.y);
  } else {
    static_assert(PMODE == 1, "one polarisation: total power only");
    acc[0] = fmaf(x0.y, x0.y, fmaf(x0.x, x0.x, acc[0]));
  }
}
template <int PMODE>
__device__ __forceinline__ void detect_finish(float (&v)[nq<PMODE>()], float gam) {
  if constexpr (PMODE != 1) {
    if constexpr (PMODE == -4) {
      const float xx = v[0], yy = v[1];
      v[0] = xx + yy;
      v[1] = xx - yy;
    }
    v[3] *= gam;
  }
}

// Four-step intermediate Y: written once by pass A and read once by the
// vector pass B a whole launch later, when far more than the caches hold
// has passed through them.  Bit 0 of RSK_NT_Y marks pass B's loads of Y
// non-temporal (streaming: the lines do not displace cached data), bit 1
// pass A's stores; both measured faster.  The matrix-core pass B reads Y
// with plain loads: its waves share lines, which streaming loads would
// fetch again.
#ifndef RSK_NT_Y
#define RSK_NT_Y 3
#endif
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void store_y(float4 *p, float4 v) {
  if constexpr (RSK_NT_Y & 2) {
    const f32x4 t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<f32x4 *>(p));
  } else {
    *p = v;
  }
}
__device__ __forceinline__ void store_y(float2 *p, float2 v) {
  if constexpr (RSK_NT_Y & 2) {
    const f32x2 t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<f32x2 *>(p));
  } else {
    *p = v;
  }
}
__device__ __forceinline__ float4 load_y(const float4 *p) {
  if constexpr (RSK_NT_Y & 1) {
    const f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
    return make_float4(t.x, t.y, t.z, t.w);
  } else {
    return *p;
  }
}
__device__ __forceinline__ float2 load_y(const float2 *p) {
  if constexpr (RSK_NT_Y & 1) {
    const f32x2 t = __builtin_nontemporal_load(reinterpret_cast<const f32x2 *>(p));
    return make_float2(t.x, t.y);
  } else {
    return *p;
  }
}

// 24-bit intermediate (engine level 1, two pols): each float of Y is rounded
// to nearest at mantissa bit 8 (ties away from zero) and its upper three
// bytes kept, so a two-pol complex value takes 12 bytes instead of 16 (every
// value keeps its own exponent; 16 significant bits).  Adding half a unit of
// the kept bits to the sign-magnitude pattern rounds the magnitude; a carry
// out of the mantissa steps the exponent, as rounding up to the next binade
// needs.  One add per value (ties to even took a bit extract and a
// three-operand add); packing and unpacking are byte permutes.
__device__ __forceinline__ unsigned int round24(float x) {
  return __float_as_uint(x) + 0x80u;   // the upper three bytes hold the rounded value
}
// One 12-byte access per value (global_{load,store}_dwordx3 at 4-byte
// alignment): lanes holding consecutive values then cover a contiguous run
// with each instruction.  Three dword accesses per value would leave 8-byte
// gaps in every instruction's footprint, and non-temporal partial sectors
// are written to HBM separately.
typedef unsigned int u32x3 __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ u32x3 pack_y24(float4 v) {
  const unsigned int b0 = round24(v.x), b1 = round24(v.y), b2 = round24(v.z), b3 = round24(v.w);
  u32x3 w;
  w.x = __builtin_amdgcn_perm(b1, b0, 0x05030201u);
  w.y = __builtin_amdgcn_perm(b2, b1, 0x06050302u);
  w.z = __builtin_amdgcn_perm(b3, b2, 0x07060503u);
  return w;
}
__device__ __forceinline__ void store_y24(unsigned int *p, float4 v) {
  __builtin_nontemporal_store(pack_y24(v), reinterpret_cast<u32x3 *>(p));
}
__device__ __forceinline__ float4 load_y24(const unsigned int *p) {
  const u32x3 w = __builtin_nontemporal_load(reinterpret_cast<const u32x3 *>(p));
  return make_float4(__uint_as_float(w.x << 8), __uint_as_float(__builtin_amdgcn_perm(w.y, w.x, 0x0504030cu)),
                     __uint_as_float(__builtin_amdgcn_perm(w.z, w.y, 0x0403020cu)),
                     __uint_as_float(w.z & 0xffffff00u));
}

// The running sum of an integration in its internal layout: each block is
// read and written once per pass-B launch by the one workgroup (or
// wavefront) that owns it, so RSK_NT_ACC = 1 streams it past the caches
// (measured: one-buffer calls 1% faster, batches unchanged).
#ifndef RSK_NT_ACC
#define RSK_NT_ACC 1
#endif
__device__ __forceinline__ float load_acc(const float *p) {
  if constexpr (RSK_NT_ACC) return __builtin_nontemporal_load(p);
  else return *p;
}
__device__ __forceinline__ void store_acc(float *p, float v) {
  if constexpr (RSK_NT_ACC) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Bijective blockIdx -> logical tile map that hands each of the 8 XCDs a
// contiguous range of tiles (hardware dispatches workgroups round-robin over
// XCDs), so neighbouring tiles share one L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned bid, unsigned nb) {
  const unsigned q = nb >> 3, r = nb & 7, x = bid & 7, i = bid >> 3;
  return x * q + min(x, r) + i;
}

}  // namespace rsk
