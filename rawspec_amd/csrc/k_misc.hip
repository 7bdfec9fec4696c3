// k_misc.hip -- the small elementwise kernels around the spectrometer core.
//
//  k_reduce   sums a dump group's partial integrations in a fixed order and
//             writes (or adds to) the integration buffer.  It replaces the
//             reference's serial accumulate kernel (rawspec_gpu.cu:297-318)
//             for groups split across workgroups.
//  k_ics      weighted incoherent sum across antennas of the integrated,
//             FFT-shifted spectra (rawspec_gpu.cu:321-340, weights :1029-1056),
//             written straight into the layout of h_icsbuf.
//  k_expand4  complex4 -> complex8 expansion of one block into its slot of the
//             device input buffer (rawspec_gpu.cu:342-374, 1395-1437).
#include "kernels.h"

namespace rsk {

__global__ __launch_bounds__(256) void k_reduce(ReduceArgs a) {
  const unsigned long long total = (unsigned long long)a.Nc * a.Ng * a.Q * a.N;
  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned jo = (unsigned)(idx % a.N);
    unsigned long long r = idx / a.N;
    const unsigned q = (unsigned)(r % a.Q);
    r /= a.Q;
    const unsigned g = (unsigned)(r % a.Ng);
    const unsigned c = (unsigned)(r / a.Ng);
    const float *p = a.part + ((((unsigned long long)c * a.Ng + g) * a.G) * a.Q + q) * a.N + jo;
    const unsigned long long step = (unsigned long long)a.Q * a.N;
    float sum = 0.0f;
    for (unsigned ch = 0; ch < a.G; ++ch) sum += p[ch * step];
    float *o = a.acc + (((unsigned long long)g * a.Q + q) * a.Nc + c) * a.N + jo;
    const float val = sum * a.scale;
    *o = a.first ? val : (*o + val);
  }
}

__global__ __launch_bounds__(256) void k_ics(IcsArgs a) {
  const unsigned long long total = (unsigned long long)a.Nds * a.Q * a.Ncpa * a.N;
  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned jo = (unsigned)(idx % a.N);
    unsigned long long r = idx / a.N;
    const unsigned cc = (unsigned)(r % a.Ncpa);
    const unsigned long long dq = r / a.Ncpa;     // d*Q + q
    const float *src = a.acc + (dq * a.Nc + cc) * a.N + jo;
    const unsigned long long astep = (unsigned long long)a.Ncpa * a.N;
    float sum = 0.0f;
    for (unsigned ant = 0; ant < a.Nant; ++ant) sum += a.w[ant] * src[ant * astep];
    a.ics[idx] = sum;
  }
}

// One thread per complex4 byte: high nibble -> real, low nibble -> imaginary,
// both sign-extended to int8.
__global__ __launch_bounds__(256) void k_expand4(const uint8_t *__restrict__ stage,
                                                 uint8_t *__restrict__ dst,
                                                 unsigned long long chan_bytes4,
                                                 unsigned long long total,
                                                 unsigned long long dst_chan_stride,
                                                 unsigned long long dst_block_off) {
  for (unsigned long long idx = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       idx < total; idx += (unsigned long long)gridDim.x * blockDim.x) {
    const unsigned long long c = idx / chan_bytes4;
    const unsigned long long x = idx - c * chan_bytes4;
    const uint8_t b = stage[idx];
    const int8_t re = (int8_t)(b & 0xf0) >> 4;
    const int8_t im = (int8_t)(uint8_t)(b << 4) >> 4;
    uint8_t *o = dst + c * dst_chan_stride + dst_block_off + 2 * x;
    o[0] = (uint8_t)re;
    o[1] = (uint8_t)im;
  }
}

static unsigned grid_for(unsigned long long total) {
  unsigned long long nb = (total + 255) / 256;
  if (nb > 65536ull * 16) nb = 65536ull * 16;
  return nb ? (unsigned)nb : 1u;
}

int launch_reduce(const ReduceArgs &a, hipStream_t s) {
  const unsigned long long total = (unsigned long long)a.Nc * a.Ng * a.Q * a.N;
  hipLaunchKernelGGL(k_reduce, dim3(grid_for(total)), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_ics(const IcsArgs &a, hipStream_t s) {
  const unsigned long long total = (unsigned long long)a.Nds * a.Q * a.Ncpa * a.N;
  hipLaunchKernelGGL(k_ics, dim3(grid_for(total)), dim3(256), 0, s, a);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

int launch_expand4(const uint8_t *stage, uint8_t *dst, unsigned Nc,
                   unsigned long long chan_bytes4, unsigned long long dst_chan_stride,
                   unsigned long long dst_block_off, hipStream_t s) {
  const unsigned long long total = (unsigned long long)Nc * chan_bytes4;
  hipLaunchKernelGGL(k_expand4, dim3(grid_for(total)), dim3(256), 0, s, stage, dst, chan_bytes4,
                     total, dst_chan_stride, dst_block_off);
  return hipGetLastError() == hipSuccess ? 0 : 1;
}

}  // namespace rsk
