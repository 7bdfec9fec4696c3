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

constexpr int wg_for(int n) { return n <= 16 ? 64 : 256; }

template <int N, int NBPS, int NP, int PMODE>
__global__ __launch_bounds__(wg_for(N)) void k_small(SmallArgs a) {
  constexpr int E = elems_for(N);
  constexpr int T = N / E;
  constexpr int WG = wg_for(N);
  constexpr int SLOTS = WG / T;
  constexpr int Q = nq<PMODE>();
  constexpr int FFT_F2 = (N > E) ? SLOTS * NP * N : 1;
  constexpr int RED_F2 = (SLOTS * Q * N + 1) / 2;
  constexpr int LDS_F2 = FFT_F2 > RED_F2 ? FFT_F2 : RED_F2;
  using S = Sample<NBPS, NP>;
  using W = typename S::word;
  __shared__ float2 lds[LDS_F2];

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

  const W *src = reinterpret_cast<const W *>(a.in + (unsigned long long)c * a.chan_stride);
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
    const unsigned long long sbase = (unsigned long long)(s0 + (valid ? sl : 0u)) * N;
    float2 v[NP][E];
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const W w = src[sbase + j + m * T];
      float2 o[NP];
      S::cvt(w, o, a.isign);
#pragma unroll
      for (int p = 0; p < NP; ++p) v[p][m] = o[p];
    }
    team_fft<N, E, NP>(v, lds + slot * NP * N, j, a.tw);
    if (valid) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        float2 x[NP];
#pragma unroll
        for (int p = 0; p < NP; ++p) x[p] = v[p][m];
        detect_add<PMODE, NP>(acc[m], x, a.gam);
      }
    }
  }

  // Every slot parks its sums in LDS ([slot][q][bin]); the write-out loop
  // adds the slots of one group in increasing slot order.
  __syncthreads();   // the FFT exchange area is reused
  float *red = reinterpret_cast<float *>(lds);
#pragma unroll
  for (int m = 0; m < E; ++m)
#pragma unroll
    for (int q = 0; q < Q; ++q) red[(slot * Q + q) * N + j + m * T] = acc[m][q];
  __syncthreads();

  const int nout = a.gpw * Q * N;
  for (int idx = tid; idx < nout; idx += WG) {
    const int go = idx / (Q * N);
    const int r2 = idx - go * (Q * N);
    const int q = r2 / N;
    const int jo = r2 - q * N;                   // output (shifted) fine channel
    const int k = (jo + N / 2) & (N - 1);        // FFT bin stored there
    const unsigned gg = g0 + go;
    if (gg >= a.Ng) continue;
    const float *rp = red + ((go * a.spg) * Q + q) * N + k;
    float sum = rp[0];
    for (int s2 = 1; s2 < a.spg; ++s2) sum += rp[s2 * Q * N];
    if (a.direct) {
      float *o = a.acc + (((unsigned long long)gg * Q + q) * a.Nc + c) * N + jo;
      const float val = sum * a.scale;
      *o = a.first ? val : (*o + val);
    } else {
      a.part[((((unsigned long long)c * a.Ng + gg) * a.G + chunk) * Q + q) * N + jo] = sum;
    }
  }
}

int small_wg(int n) { return n <= 16 ? 64 : 256; }
int small_slots(int n) { return small_wg(n) / (n / elems_for(n) > 0 ? n / elems_for(n) : 1); }

template <int N, int NBPS, int NP, int PMODE>
static int go(const SmallArgs &a, unsigned nb, hipStream_t s) {
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
  if (nbps == 8) return go_n<8>(n, np, pmode, a, nblocks, s);
  if (nbps == 16) return go_n<16>(n, np, pmode, a, nblocks, s);
  return 2;
}

}  // namespace rsk
