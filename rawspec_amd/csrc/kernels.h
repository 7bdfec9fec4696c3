// kernels.h -- launch interface of the spectrometer kernels (internal to
// librawspec; the host C layer reaches them only through rsd.h).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rsk {

constexpr int kTwLog = 12;            // master twiddle table: W_4096^m
constexpr int kTwN = 1 << kTwLog;     // also the largest single-pass FFT

// Fused small-N path (N <= 4096): unpack -> FFT -> detect -> integrate a
// chunk of `chunk` spectra.  Workgroup b covers channel b / bpc and either
// groups [gpw*(b % bpc), +gpw) whole, or chunk (b % bpc) % G of group
// (b % bpc) / G.
struct SmallArgs {
  const uint8_t *in;              // device input buffer [Nc][Nb*Ntpb][Np][2]
  unsigned long long chan_stride; // bytes per channel of the input buffer
  unsigned Nc;
  unsigned na_buf;                // spectra per dump group in this buffer
  unsigned chunk;                 // spectra per workgroup
  unsigned G;                     // chunks per group (gpw == 1 when G > 1)
  unsigned Ng;                    // groups per channel
  unsigned gpw;                   // groups per workgroup
  unsigned spg;                   // concurrent slots per group (SLOTS / gpw)
  unsigned bpc;                   // workgroups per channel: ceil(Ng/gpw)*G
  float isign;                    // -1: negate imaginary input (inverse FFT)
  float gam;                      // sign of the imaginary cross term
  float scale;                    // SNORM normalisation of the detected sums
  int direct;                     // 1: write the integration buffer directly
  int first;                      // direct: overwrite (1) or accumulate (0)
  float *acc;                     // [Nds][Q][Nc][N], FFT-shifted
  float *part;                    // [Nc][Ng][G][Q][N] unscaled, FFT-shifted
  const float2 *tw;               // W_4096^m, m < 4096
  // transform length: the kernel's N, or for a product whose length is not
  // a power of two (Bluestein, N >= 2 nt - 1) nt with its tables
  unsigned nt;
  const float2 *chirp;            // c[n] = exp(-i pi n^2 / nt), n < nt
  const float2 *bsk;              // FFT_N of the wrapped conjugate chirp, over N
  // k_small_mfma_u with a fused 8-point product: that product's output
  // [Nds][Q][Nc][8], one dump per 1024-sample spectrum (nullptr: none)
  float *f8_out;
};

// Four-step path (N = N1*N2 > 4096), launched per chunk of channels
// [c0, c0 + nch) so the intermediate stays cache-resident.
// Pass A: length-N2 FFTs over stride-N1 input, twiddle W_N^(n1*k2), store Y
// tiled for pass B: Y[cl][s][k2 / TB][n1][k2 % TB][pol].
// Pass B: length-N1 FFTs per k2, detection, integration of the group's
// spectra; output per `mode`:
//   FS_HOST   write the FFT-shifted host layout (integration complete)
//   FS_FIRST  write the internal accumulator (first buffer of Nis > 1)
//   FS_MID    add into the internal accumulator
//   FS_LAST   internal + this buffer -> host layout
// Internal accumulator: [Ng][Nc][N2/TB][Q][N1][TB] (one contiguous block
// per workgroup).
enum { FS_HOST = 0, FS_FIRST = 1, FS_MID = 2, FS_LAST = 3 };

// Buffers one pass A/B launch pair can take (batched processing of
// consecutive buffers of one integration).
constexpr int kMaxBatch = 8;

struct FsArgs {
  const uint8_t *in[kMaxBatch];   // input buffer of each batched buffer
  unsigned nbat;                  // buffers in this launch (1 .. kMaxBatch)
  unsigned long long chan_stride;
  // Nss: spectra per channel and buffer; na_buf, Ng: pass B's dump groups
  // over the launch's nbat * Nss spectra per channel
  unsigned Nc, Nss, na_buf, Ng;
  unsigned c0, nch;               // channel chunk
  unsigned N1, N2;
  float isign, gam, scale;
  int mode;
  float2 *work;                   // Y for the chunk
  float *acc;                     // host layout [Nds][Q][Nc][N]
  float *acc_int;                 // internal accumulator (Nis > 1)
  const float2 *tw;               // W_4096^m
  const float2 *twf;              // W_N^r, r < N1 (N1 <= 4096)
  // Y holds ynb buffers' spectra per channel ([cl][slot * Nss + sp]); pass A
  // writes this launch's nbat buffers to slots yb0 .., pass B reads slots
  // 0 .. nbat - 1
  unsigned yb0, ynb;
  unsigned grid_a;                // pass-A workgroups, each looping over items (0: one per item)
  unsigned tby;                   // k2 tile of Y: fs_tile_b(N1), both pass-B kernels
  float mscale;                   // k_fs_b_mfma: power-of-two input scale (fs_mfma_scale)
  int y24;                        // two pols: Y holds 12-byte (24-bit float) values
  // vector pass B: each dump group's na_buf spectra run in ns runs of at
  // most sp_run spectra, one workgroup per (run, k2 tile, group, channel);
  // with ns > 1 the runs' sums go to partb, laid out [ns][Ng][nch][tiles]
  // [Q][N1][tby], and k_fs_b_reduce adds them and finishes the launch
  unsigned ns, sp_run;
  float *partb;
};

struct ReduceArgs {
  const float *part;
  float *acc;
  unsigned Nc, Ng, G, Q, N;
  float scale;
  int first;
};

struct IcsArgs {
  const float *acc;
  float *ics;
  const float *w;
  unsigned Nc, Nant, Ncpa, Q, N, Nds;
};

// Query functions used by the host-side planner.
int small_wg(int n);            // threads per workgroup of the small kernel
int small_slots(int n);         // spectra processed concurrently per workgroup
int fs_tile_a(int n2);          // n1 values per pass-A workgroup
int fs_tile_b(int n1);          // k2 values per pass-B workgroup

// Launchers; return 0 on success, non-zero for an unsupported shape or a
// launch error.
int launch_small(int n, int nbps, int np, int pmode, const SmallArgs &a,
                 unsigned nblocks, hipStream_t s);
int launch_fs_a(int nbps, int np, const FsArgs &a, hipStream_t s);
int launch_fs_b(int np, int pmode, const FsArgs &a, hipStream_t s);
// The ns runs' sums of a split pass B, in run order, then pass B's own
// finish (scale, internal accumulator, shifted output) for a.mode.
int launch_fs_b_reduce(int pmode, const FsArgs &a, hipStream_t s);
// Matrix-core fused small kernel (N = 1024, two pols): one wavefront per
// (channel, group, chunk) item, items = Nc * Ng * G, gpw must be 1.
int small_mfma_ok(int n, int np);
int launch_small_mfma(int n, int nbps, int np, int pmode, const SmallArgs &a, unsigned items,
                      hipStream_t s);
// The same with unsplit FP16 operands (int8 samples only): a third of the
// matrix instructions, error of order 6e-3 / sqrt(spectra integrated) over a
// floor near 1.5e-4, for products that integrate enough spectra.
int launch_small_mfma_u(int n, int nbps, int np, int pmode, const SmallArgs &a, unsigned items,
                        hipStream_t s);
// Matrix-core pass A (N1 = N2 = 1024, int8, two pols): k_fs_a's contract;
// a.grid_a workgroups loop over the items.
int fs_a_mfma_ok(int n1, int n2, int nbps, int np);
int launch_fs_a_mfma(int nbps, int np, const FsArgs &a, hipStream_t s);
// Matrix-core pass B (N1 = 1024, two pols): a workgroup's 4 columns are one
// k2 tile of Y (tby = fs_tile_b(1024) = 4).
int fs_b_mfma_ok(int n1, int np);
int launch_fs_b_mfma(int np, int pmode, const FsArgs &a, hipStream_t s);
int launch_reduce(const ReduceArgs &a, hipStream_t s);
int launch_ics(const IcsArgs &a, hipStream_t s);
int launch_expand4(const uint8_t *stage, uint8_t *dst, unsigned Nc,
                   unsigned long long chan_bytes4, unsigned long long dst_chan_stride,
                   unsigned long long dst_block_off, hipStream_t s);

}  // namespace rsk
