// k_small.hip -- fused spectrometer kernel for FFT lengths N <= 4096.
//
// Replaces, for one output product and one device input buffer, the chain
// cuFFT load callback (rawspec_gpu.cu:133-148) -> cuFFT C2C (:1182-1195,
// :1555-1569) -> store callbacks (:153-267) -> accumulate (:297-318) ->
// FFT-shift copies (:1647-1697) of the reference.  One workgroup integrates
// a chunk of consecutive spectra of one coarse channel:
//   * each spectrum's samples (both pols) are read once with coalesced loads
//     (thread j of a team reads samples j + m*T, consecutive across lanes),
//   * converted from int8/int16 and transformed in registers + LDS,
//   * detected (power, full pol or Stokes) and summed in registers,
//   * summed across the workgroup's concurrent spectra through LDS, in a
//     fixed order so results are reproducible run to run,
//   * written once, FFT-shifted, either straight into the integration buffer
//     or as a partial sum reduced by k_reduce.
#include "fft_device.h"
#include "kernels.h"

namespace rsk {

// Threads per workgroup: one wave for the register-only lengths, otherwise
// 256 or one team, whichever is larger.
constexpr int wg_for(int n) {
  return n <= 16 ? 64 : (n / elems_for(n) > 256 ? n / elems_for(n) : 256);
}

// LDS of one workgroup (float2): the FFT exchange area, reused for the sums
// of half the slots (a group of one slot writes from registers), and the
// twiddle table.
template <int N, int NP, int PMODE>
constexpr int small_lds_f2() {
  constexpr int E = elems_for(N), T = N / E, SLOTS = wg_for(N) / T, Q = nq<PMODE>();
  constexpr int FFT_F2 = (N > E) ? SLOTS * team_stride<N, NP>() : 1;
  constexpr int RED_F2 = SLOTS >= 2 ? ((SLOTS / 2) * Q * N + 1) / 2 : 1;
  return FFT_F2 > RED_F2 ? FFT_F2 : RED_F2;
}
// at most 256 registers per thread, so two waves share each SIMD.
//
// BS: a transform length NT = a.nt that is not a power of two, by
// Bluestein's identity with the padded length N >= 2 NT - 1 (a power of
// two):  X[k] = c[k] * sum_n (x[n] c[n]) conj(c[k - n]),  c[n] =
// exp(-i pi n^2 / NT), the sum a cyclic convolution of length N.  The
// samples times c (zero past NT) go through the team FFT, times a.bsk =
// FFT_N(the conjugate chirp, wrapped) / N, then through the team FFT again
// conjugated (an inverse FFT), and conjugated times c; bins k >= NT are
// zero and dropped.
template <int N, int NBPS, int NP, int PMODE, bool BS = false>
__global__ __launch_bounds__(wg_for(N)) __attribute__((amdgpu_waves_per_eu(BS ? 1 : 2)))
void k_small(SmallArgs a) {
  constexpr int E = elems_for(N);
  constexpr int T = N / E;
  constexpr int WG = wg_for(N);
  constexpr int SLOTS = WG / T;
  constexpr int Q = nq<PMODE>();
  constexpr int LDS_F2 = small_lds_f2<N, NP, PMODE>();
  using S = Sample<NBPS, NP>;
  using W = typename S::word;
  __shared__ float2 lds[LDS_F2];
  __shared__ float2 twl[TwTab<N>::LEN];

  // Workgroup -> (channel, first group, chunk).  gpw groups per workgroup
  // (each with spg = SLOTS/gpw slots) when groups are short; otherwise one
  // group (gpw = 1) split over G chunks of `chunk` spectra.
  const unsigned per_c = a.bpc;
  const unsigned c = blockIdx.x / per_c;
  const unsigned rest = blockIdx.x - c * per_c;
  const unsigned gb = rest / a.G;
  const unsigned chunk = rest - gb * a.G;
  const unsigned g0 = gb * a.gpw;
  const int tid = threadIdx.x;
  const int slot = tid / T;
  const int j = tid - slot * T;
  const int gi = slot / a.spg;
  const int si = slot - gi * a.spg;
  const unsigned g = g0 + gi;
  const bool gvalid = g < a.Ng;

  const float2 *twp = tw_setup<N, WG>(twl, a.tw);
  const W *src = reinterpret_cast<const W *>(a.in + (unsigned long long)c * a.chan_stride);
  // transform length (samples per spectrum, bins per output spectrum)
  const unsigned NT = BS ? a.nt : (unsigned)N;
  const unsigned s0 = (gvalid ? g : 0u) * a.na_buf + chunk * a.chunk;

  float acc[E][Q];
#pragma unroll
  for (int m = 0; m < E; ++m)
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[m][q] = 0.0f;

  const int iters = (int)((a.chunk + a.spg - 1) / a.spg);
  for (int it = 0; it < iters; ++it) {
    const unsigned sl = (unsigned)(si + it * a.spg);
    const bool valid = gvalid && sl < a.chunk;
    const unsigned long long sbase = (unsigned long long)(s0 + (valid ? sl : 0u)) * NT;
    float2 v[NP][E];
    W raw[E];
    if constexpr (BS) {
      // samples past NT are the padding (and may lie past the buffer)
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const unsigned n = (unsigned)(j + m * T);
        raw[m] = n < NT ? src[sbase + n] : W{};
      }
    } else if constexpr (T == 1 && (E * sizeof(W)) % 16 == 0) {
      // one thread per transform: the spectrum is contiguous, 16-byte loads
      const uint4 *s4 = reinterpret_cast<const uint4 *>(src + sbase);
      uint4 *r4 = reinterpret_cast<uint4 *>(raw);
#pragma unroll
      for (int m = 0; m < (int)(E * sizeof(W) / 16); ++m) r4[m] = s4[m];
    } else {
#pragma unroll
      for (int m = 0; m < E; ++m) raw[m] = src[sbase + j + m * T];
    }
    // one thread per int8 transform: samples offset by 128, removed from bin
    // 0 (the thread's element 0) after the single-stage transform
    constexpr bool kOffset = T == 1 && NBPS == 8 && !BS;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float2 o[NP];
      if constexpr (kOffset) S::cvt_offset(raw[m], o, a.isign);
      else S::cvt(raw[m], o, a.isign);
#pragma unroll
      for (int p = 0; p < NP; ++p) v[p][m] = o[p];
    }
    // Bluestein tables are read per spectrum (L1 and L2 hold them): an
    // opaque zero in the index keeps the compiler from holding all 3 E of
    // their values in registers across the spectrum loop, which spills
    unsigned zt = 0;
    if constexpr (BS) asm volatile("v_mov_b32 %0, 0" : "=v"(zt));
    if constexpr (BS) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const unsigned n = (unsigned)(j + m * T);
        const float2 cn = n < NT ? a.chirp[n + zt] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int p = 0; p < NP; ++p) v[p][m] = cmul(v[p][m], cn);
      }
    }
    // a team is T consecutive threads: within one wavefront when T <= 64
    team_fft<N, E, NP, TwTab<N>::TWN, (T <= 64)>(v, lds + slot * team_stride<N, NP>(), j, twp);
    if constexpr (BS) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const float2 bk = a.bsk[j + m * T + zt];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          const float2 z = cmul(v[p][m], bk);
          v[p][m] = make_float2(z.x, -z.y);
        }
      }
      // the second transform reads the twiddle table through an opaque
      // offset, so its loads are not merged with the first's (which would
      // keep those values live in between)
      unsigned zw;
      asm volatile("v_mov_b32 %0, 0" : "=v"(zw));
      team_fft<N, E, NP, TwTab<N>::TWN, (T <= 64)>(v, lds + slot * team_stride<N, NP>(), j, twp + zw);
#pragma unroll
      for (int m = 0; m < E; ++m) {
        const unsigned k = (unsigned)(j + m * T);
        const float2 ck = k < NT ? a.chirp[k + zt] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int p = 0; p < NP; ++p) v[p][m] = cmul(make_float2(v[p][m].x, -v[p][m].y), ck);
      }
    }
    if constexpr (kOffset) {
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        v[p][0].x -= 128.0f * N;
        v[p][0].y -= a.isign * (128.0f * N);
      }
    }
    if (valid) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        detect_raw_np<PMODE, NP>(acc[m], v[0][m], v[NP - 1][m]);
      }
    }
  }
  // raw sums -> the product's components (linear, so the slot sums below
  // add finished values)
#pragma unroll
  for (int m = 0; m < E; ++m) detect_finish<PMODE>(acc[m], a.gam);

  __syncthreads();   // the FFT exchange area is reused
  float *base = a.direct ? a.acc : a.part;
  const bool accum = a.direct && !a.first;
  if (a.spg == 1) {
    // Every slot is a group of its own: its sums go straight from registers
    // (consecutive lanes hold consecutive bins).
    const unsigned gg = g0 + slot;
    if (gg >= a.Ng) return;
    float *bg = a.direct ? base + (((unsigned long long)gg * Q) * a.Nc + c) * NT
                         : base + ((((unsigned long long)c * a.Ng + gg) * a.G + chunk) * Q) * NT;
    const unsigned long long qstride = a.direct ? (unsigned long long)a.Nc * NT : NT;
    const float sc = a.direct ? a.scale : 1.0f;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      float *bq = bg + q * qstride;
      float val[E];
      unsigned pos[E];                    // output position of bin j + m*T
#pragma unroll
      for (int m = 0; m < E; ++m) {
        val[m] = acc[m][q] * sc;
        const unsigned k = (unsigned)(j + m * T);
        pos[m] = BS ? (k + NT / 2) % NT : (k + N / 2) & (N - 1);
      }
      if (accum) {
#pragma unroll
        for (int m = 0; m < E; ++m)
          if (!BS || (unsigned)(j + m * T) < NT) val[m] += bq[pos[m]];
      }
#pragma unroll
      for (int m = 0; m < E; ++m)
        if (!BS || (unsigned)(j + m * T) < NT) bq[pos[m]] = val[m];
    }
    return;
  }
  // Groups of spg >= 2 slots: the upper half of each group's slots parks its
  // sums in LDS ([group][slot - half][q][bin]), the lower half adds its own
  // to them, and the write-out loop adds the half slots of one group in
  // increasing order, so the order of the sums is fixed.
  float *red = reinterpret_cast<float *>(lds);
  const int half = (int)a.spg >> 1;
  if (si >= half) {
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) red[((gi * half + si - half) * Q + q) * N + j + m * T] = acc[m][q];
  }
  __syncthreads();
  if (si < half) {
#pragma unroll
    for (int m = 0; m < E; ++m)
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        float *r = red + ((gi * half + si) * Q + q) * N + j + m * T;
        *r = acc[m][q] + *r;
      }
  }
  __syncthreads();

  // Write-out in batches of OB outputs per thread: sums first, then (when
  // accumulating) the batch's previous values are loaded together before the
  // stores, so read-modify-writes do not serialise on memory latency.
  constexpr int OB = 4;
  const int nout = a.gpw * Q * (int)NT;
  for (int i0 = tid; i0 < nout; i0 += OB * WG) {
    float val[OB];
    unsigned long long dst[OB];
#pragma unroll
    for (int u = 0; u < OB; ++u) {
      const int idx = i0 + u * WG;
      const int go = idx / (Q * (int)NT);
      const int r2 = idx - go * (Q * (int)NT);
      const int q = r2 / (int)NT;
      const int jo = r2 - q * (int)NT;             // output (shifted) fine channel
      // FFT bin stored there: (k + NT/2) mod NT = jo
      const int k = BS ? (int)((jo + NT - NT / 2) % NT) : (jo + N / 2) & (N - 1);
      const unsigned gg = g0 + go;
      dst[u] = ~0ull;
      val[u] = 0.0f;
      if (idx < nout && gg < a.Ng) {
        const float *rp = red + ((go * half) * Q + q) * N + k;
        float sum = rp[0];
        for (int s2 = 1; s2 < half; ++s2) sum += rp[s2 * Q * N];
        if (a.direct) {
          dst[u] = (((unsigned long long)gg * Q + q) * a.Nc + c) * NT + jo;
          val[u] = sum * a.scale;
        } else {
          dst[u] = ((((unsigned long long)c * a.Ng + gg) * a.G + chunk) * Q + q) * NT + jo;
          val[u] = sum;
        }
      }
    }
    if (accum) {
#pragma unroll
      for (int u = 0; u < OB; ++u)
        if (dst[u] != ~0ull) val[u] += base[dst[u]];
    }
#pragma unroll
    for (int u = 0; u < OB; ++u)
      if (dst[u] != ~0ull) base[dst[u]] = val[u];
  }
}

int small_wg(int n) { return wg_for(n); }
int small_slots(int n) { return small_wg(n) / (n / elems_for(n) > 0 ? n / elems_for(n) : 1); }

template <int N, int NBPS, int NP, int PMODE>
static int go(const SmallArgs &a, unsigned nb, hipStream_t s) {
  if constexpr (N >= 8) {
    if (a.nt != (unsigned)N) {
      hipLaunchKernelGGL((k_small<N, NBPS, NP, PMODE, true>), dim3(nb), dim3(wg_for(N)), 0, s, a);
      return hipGetLastError() == hipSuccess ? 0 : 1;
    }
  }
  hipLaunchKernelGGL((k_small<N, NBPS, NP, PMODE>), dim3(nb), dim3(wg_for(N)), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

template <int N, int NBPS>
static int go_mode(int np, int pmode, const SmallArgs &a, unsigned nb, hipStream_t s) {
  if (np == 1) return go<N, NBPS, 1, 1>(a, nb, s);
  switch (pmode) {
    case 1: return go<N, NBPS, 2, 1>(a, nb, s);
    case 4: return go<N, NBPS, 2, 4>(a, nb, s);
    case -4: return go<N, NBPS, 2, -4>(a, nb, s);
  }
  return 2;
}

template <int NBPS>
static int go_n(int n, int np, int pmode, const SmallArgs &a, unsigned nb, hipStream_t s) {
  switch (n) {
    case 1: return go_mode<1, NBPS>(np, pmode, a, nb, s);
    case 2: return go_mode<2, NBPS>(np, pmode, a, nb, s);
    case 4: return go_mode<4, NBPS>(np, pmode, a, nb, s);
    case 8: return go_mode<8, NBPS>(np, pmode, a, nb, s);
    case 16: return go_mode<16, NBPS>(np, pmode, a, nb, s);
    case 32: return go_mode<32, NBPS>(np, pmode, a, nb, s);
    case 64: return go_mode<64, NBPS>(np, pmode, a, nb, s);
    case 128: return go_mode<128, NBPS>(np, pmode, a, nb, s);
    case 256: return go_mode<256, NBPS>(np, pmode, a, nb, s);
    case 512: return go_mode<512, NBPS>(np, pmode, a, nb, s);
    case 1024: return go_mode<1024, NBPS>(np, pmode, a, nb, s);
    case 2048: return go_mode<2048, NBPS>(np, pmode, a, nb, s);
    case 4096: return go_mode<4096, NBPS>(np, pmode, a, nb, s);
  }
  return 2;
}

int launch_small(int n, int nbps, int np, int pmode, const SmallArgs &a, unsigned nblocks,
                 hipStream_t s) {
  if (nblocks == 0) return 0;
  // a Bluestein product: 2 <= nt, 2 nt - 1 <= n, with both tables
  if (a.nt != (unsigned)n &&
      (n < 8 || a.nt < 2 || 2ull * a.nt - 1 > (unsigned)n || !a.chirp || !a.bsk))
    return 2;
  if (nbps == 8) return go_n<8>(n, np, pmode, a, nblocks, s);
  if (nbps == 16) return go_n<16>(n, np, pmode, a, nblocks, s);
  return 2;
}

}  // namespace rsk
