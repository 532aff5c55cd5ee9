// Thermal-scattering surrogate: an MLP (ln E, T, u1, u2) -> (y1, y2) that
// replaces the POD inverse-CDF sampling of ThermalScattering::Scatter
// (ThermalScattering.cpp:159-171, :271-463).  y1 = beta kT / (E + kT) and
// y2 = (alpha - alpha_min(beta)) / (alpha_max(beta) - alpha_min(beta)).
//
// - POD transport map (training targets): the same POD tables and
//   inverse-CDF arithmetic as SampleBeta/SampleAlpha, with linear
//   interpolation between grid energies / betas instead of a random pick.
// - Training: FP32 tiled GEMMs on the vector ALUs, SiLU, MSE loss, Adam.
// - Inference: one fused persistent kernel (KMlpF).  Layer 0 (4 -> 256) and
//   layer 3 (256 -> 2) run on the vector ALUs; the two 256 x 256 layers run
//   on v_mfma_f32_32x32x16_bf16 with each wave's share of the weights held
//   in registers for the whole launch and the activations passed between
//   layers through LDS.  Each layer is computed transposed (Z^T = W H^T).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/minimc_amd.h"
#include "mmc_mlp.hpp"

using namespace mmc;

hipStream_t MmcStream(mmc_ctx* ctx);
MlpState& MmcMlp(mmc_ctx* ctx);
const DevWorld& MmcWorld(mmc_ctx* ctx);
int MmcFail(mmc_ctx* ctx, int code, const char* msg);
int MmcDevice(mmc_ctx* ctx);
int MmcLaunchCheck(mmc_ctx* ctx, const char* where);
int MmcTslCount(mmc_ctx* ctx);

// Waits for the context's stream after a batch of launches; a launch that
// failed at configuration time or a fault is an error.
static int SyncChecked(mmc_ctx* ctx, const char* where) {
  if (const int lc = MmcLaunchCheck(ctx, where)) return lc;
  if (hipStreamSynchronize(MmcStream(ctx)) != hipSuccess)
    return MmcFail(ctx, MMC_ERR_HIP, where);
  return MMC_OK;
}

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

namespace mmc {

constexpr int kHid = 256;           // hidden width of the fused kernel
constexpr int kFragElems = 8 * 16 * 64 * 8;  // bf16 per 256x256 layer (65536)

// The fused kernel's small parameters, FP32, one block staged in LDS:
//   [kW0, +1024)  layer-0 weights as float4 [h 2][kb 16][q 8]
//   [kB0, +256)   layer-0 bias [h][kb][q]
//   [kB1, +256)   layer-1 bias [h 2][t 8][r 16]
//   [kB2, +256)   layer-2 bias
//   [kW3, +512)   layer-3 weights [h 2][t 8][r 16][2]
constexpr int kW0 = 0, kB0 = 1024, kB1 = 1280, kB2 = 1536, kW3 = 1792, kSmall = 2304;

// Device view of the network for the fused kernel.
struct MlpDev {
  const bf16x8* frag1;  // layer 1 A fragments [t 8][kb 16][lane 64]
  const bf16x8* frag2;  // layer 2
  const float* small;   // kSmall floats, layout above
  float b3[2];
  float in_shift[4], in_scale[4], out_shift[2], out_scale[2];
};

// Input feature of the K-permuted B fragment (K block kb, lane half h,
// element q); equal to the accumulator feature of tile kb/2, register
// 8 (kb % 2) + q, lane half h.
__host__ __device__ __forceinline__ int FeatB(int kb, int h, int q) {
  return 32 * (kb / 2) + (2 * (kb % 2) + q / 4) * 8 + 4 * h + q % 4;
}
// Output feature of accumulator register r of tile t in lane half h.
__host__ __device__ __forceinline__ int FeatC(int t, int r, int h) {
  return 32 * t + (r / 4) * 8 + 4 * h + r % 4;
}

__device__ __forceinline__ float Silu(float z) { return z / (1.0f + __expf(-z)); }
// SiLU for the BF16 inference path: hardware exp2 and reciprocal (about
// 1 ulp FP32 each, far below the BF16 rounding that follows).
__device__ __forceinline__ float SiluFast(float z) {
  return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}

void MlpFree(MlpState& m) {
  void* ps[] = {m.params, m.wbf16, m.adam_m, m.adam_v, m.grads, m.act, m.loss,
                m.xin, m.yout, m.berr, m.fb, m.packed};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m = MlpState{};
}

// ---------------------------------------------------------------------------
// BF16 MFMA inference (KMlpF below): a block holds kMlpWaves waves, two per
// SIMD, so that one wave's matrix instructions overlap another's vector work.
// In the tracker the kernel gathers (ln E, T, u1, u2) of the thermal-
// scattering queue itself and applies the scatter.  The rare rows the
// network cannot serve wait in a fallback queue that the POD sampler empties
// once per batch of iterations, so that its latency is not paid every
// iteration.
constexpr int kMlpWaves = 8;

// Returns false when the network's outgoing energy leaves the trained box
// (E' = E + beta kT below e_min, which includes E' <= 0, or not a number):
// the caller then uses the POD sampler for that collision.  A network that
// undershoots the lower end of the beta law (E' -> 0) would otherwise send
// particles to energies where the POD sampler's rejection loop
// (ThermalScattering.cpp:302-337) runs out of resamples.
__device__ __forceinline__ bool OutputsToAlphaBeta(double y1, double y2, double E, double T,
                                                   double awr, double beta_cutoff,
                                                   double e_min, double& alpha,
                                                   double& beta) {
  const double kT = kBoltzmann * T;
  double b = y1 * (E + kT) / kT;
  if (!(E + b * kT >= e_min) || !(E + b * kT > 0)) return false;
  if (b > beta_cutoff) b = beta_cutoff;
  const double Ep = E + b * kT;
  const double sE = sqrt(E), sp = sqrt(Ep > 0 ? Ep : 0.0);
  const double akT = awr * kT;
  const double lo = (sE - sp) * (sE - sp) / akT, hi = (sE + sp) * (sE + sp) / akT;
  const double f = y2 < 0 ? 0 : (y2 > 1 ? 1 : y2);
  alpha = lo + f * (hi - lo);
  beta = b;
  return true;
}

// Tracker-mode arguments of the two kernels.
struct TrackIO {
  DevWorld W;
  RunState R;
  int iter, qin;
  uint32_t* fb;  // fallback queue, length in RunState::misc[kMiscFallback]
  double e_min, e_max, t_lo, t_hi;
};
__device__ __forceinline__ bool InBoxT(const TrackIO& io, double E, double T) {
  return E >= io.e_min && E <= io.e_max && T >= io.t_lo && T <= io.t_hi;
}

// ---------------------------------------------------------------------------
// Fused inference: both 256 x 256 layers in one persistent kernel, with the
// weights held in registers instead of LDS.  Wave w of a block owns output
// features 32w .. 32w + 31 of layers 1 and 2 and keeps its 16 A fragments of
// each (128 registers) for the whole launch.  The activations of a block's
// kFRows rows pass between layers through LDS, stored in the K-permuted
// order of the B fragments (position 16 kb + 8 h + q holds input feature
// FeatB(kb, h, q)), so that a B fragment is one 16-byte LDS read and an
// accumulator quad one 8-byte LDS write.  One iteration of a block:
//   gather (waves 0 .. kFRows/64 - 1)          -> X      | barrier
//   layer 0 on the vector ALUs, every wave      -> H0     | barrier
//   layer 1: W1 slice x H0, bias, SiLU          -> H1     | barrier
//   layer 2: W2 slice x H1, bias, SiLU, then the wave's share of
//   layer 3 (its 32 features)                   -> parts  | barrier
//   finish (waves 0 .. kFRows/64 - 1): the sum of the 8 shares, then the
//   output rows or, in the tracker, the scatter.
// The layer-1 activations never go to memory, and a block reads each weight
// once per launch.
constexpr int kFRows = 64;         // rows per block iteration
constexpr int kFCb = kFRows / 32;  // 32-row column blocks per wave
constexpr int kHStride = 264;      // bf16 per LDS activation row (256 + 8 padding)
static_assert(kMlpWaves * 32 == kHid, "one wave per 32 output features");
static_assert(kFRows % 64 == 0 && (64 * kMlpWaves) % kFRows == 0, "row tiling");
// LDS: small parameters | X [kFRows] float4 | parts [waves][kFRows] float2 |
// pidx [kFRows] | H0, H1 [kFRows][kHStride] bf16 | each wave's layer-2 weight
// fragments of k-blocks kA2Regs..15 [waves][16 - kA2Regs][64] bf16x8.  A block
// of 8 waves at 256 VGPRs fills a CU, so the LDS beyond the activations is
// free; holding most of the layer-2 weights there leaves the registers that
// the tracker's FP64 finish needs (7 in registers: no spill).
constexpr int kA2Regs = 7;  // k-blocks of the layer-2 weights held in registers
constexpr size_t kFusedLds = sizeof(float) * (kSmall + 4 * kFRows + 2 * kMlpWaves * kFRows) +
                             sizeof(uint32_t) * kFRows +
                             sizeof(__bf16) * 2 * kFRows * kHStride +
                             sizeof(bf16x8) * kMlpWaves * (16 - kA2Regs) * 64;
static_assert(kFusedLds <= 160 * 1024, "LDS of one CU");

// LDS position (B-fragment order) of the feature that accumulator register r
// of output tile t holds in lane half h, FeatC(t, r, h) = FeatB(2t + j/2, h,
// 4 (j % 2) + r % 4) with j = r / 4.
__host__ __device__ __forceinline__ int PosC(int t, int r, int h) {
  const int j = r / 4;
  return 16 * (2 * t + j / 2) + 8 * h + 4 * (j % 2) + r % 4;
}
// MfmaRows with the weights of k-blocks kA2Regs..15 read from LDS (this
// wave's slice, As[(kb - kA2Regs) * 64 + lane]) and the others from A.
__device__ __forceinline__ void MfmaRowsSplit(const __bf16* H, const bf16x8 A[kA2Regs],
                                              const bf16x8* As, int lane, int h, int j,
                                              f32x16 acc[kFCb]) {
#pragma unroll
  for (int c = 0; c < kFCb; c++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[c][r] = 0.0f;
#pragma unroll
  for (int kb = 0; kb < 16; kb++) {
    const bf16x8 a = kb < kA2Regs ? A[kb] : As[(kb - kA2Regs) * 64 + lane];
#pragma unroll
    for (int c = 0; c < kFCb; c++) {
      const bf16x8 b =
          *reinterpret_cast<const bf16x8*>(H + (32 * c + j) * kHStride + 16 * kb + 8 * h);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[c], 0, 0, 0);
    }
  }
}
// acc[c] = A (this wave's 32 output features) x H rows 32c .. 32c + 31.
__device__ __forceinline__ void MfmaRows(const __bf16* H, const bf16x8 A[16], int h, int j,
                                         f32x16 acc[kFCb]) {
#pragma unroll
  for (int c = 0; c < kFCb; c++)
#pragma unroll
    for (int r = 0; r < 16; r++) acc[c][r] = 0.0f;
#pragma unroll
  for (int kb = 0; kb < 16; kb++) {
#pragma unroll
    for (int c = 0; c < kFCb; c++) {
      const bf16x8 b =
          *reinterpret_cast<const bf16x8*>(H + (32 * c + j) * kHStride + 16 * kb + 8 * h);
      acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[kb], b, acc[c], 0, 0, 0);
    }
  }
}

// (ln E, T, u1, u2) of queued particle p, with u1 and u2 drawn from its
// stream (the gather of the event loop); advances the stored RNG state.
__device__ __forceinline__ float4 TslInputs(const RunState& R, uint32_t p) {
  const double E = R.B.E[p];
  const double T = R.B.tslT[p];
  uint32_t g = R.B.rng[p];
  const double u1 = Canonical(g), u2 = Canonical(g);
  R.B.rng[p] = g;
  return make_float4(float(log(E)), float(T), float(u1), float(u2));
}
__device__ __forceinline__ float4 Normalize(const MlpDev& m, float4 v) {
  return make_float4((v.x - m.in_shift[0]) * m.in_scale[0], (v.y - m.in_shift[1]) * m.in_scale[1],
                     (v.z - m.in_shift[2]) * m.in_scale[2], (v.w - m.in_shift[3]) * m.in_scale[3]);
}

template <bool kTrack>
__global__ void __launch_bounds__(64 * kMlpWaves)
    KMlpF(MlpDev m, const float* __restrict__ x, float* __restrict__ y, const uint32_t* n_ptr,
          uint32_t n_static, TrackIO io) {
  extern __shared__ float4 lds_f4[];
  float* sp = reinterpret_cast<float*>(lds_f4);
  float4* X = reinterpret_cast<float4*>(sp + kSmall);
  float2* parts = reinterpret_cast<float2*>(X + kFRows);
  uint32_t* pidx = reinterpret_cast<uint32_t*>(parts + kMlpWaves * kFRows);
  __bf16* H0 = reinterpret_cast<__bf16*>(pidx + kFRows);
  __bf16* H1 = H0 + kFRows * kHStride;
  bf16x8* A2s = reinterpret_cast<bf16x8*>(H1 + kFRows * kHStride);
  const uint32_t n = n_ptr ? *n_ptr : n_static;
  const uint32_t n_it = (n + kFRows - 1) / kFRows;
  if (blockIdx.x >= n_it) return;  // no rows: read nothing
  const int wave = threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  const int h = lane / 32;
  const int j = lane % 32;
  for (int i = threadIdx.x; i < kSmall / 4; i += blockDim.x)
    reinterpret_cast<float4*>(sp)[i] = reinterpret_cast<const float4*>(m.small)[i];
  // The tracker keeps the layer-2 weights of k-blocks kA2Regs..15 in LDS (a
  // wave reads back only the LDS weights it wrote: no barrier); batch
  // inference, which has no FP64 finish, keeps them all in registers.
  constexpr int kRegs2 = kTrack ? kA2Regs : 16;
  bf16x8 A1[16], A2[kRegs2];
  bf16x8* const A2w = A2s + wave * (16 - kA2Regs) * 64;
#pragma unroll
  for (int kb = 0; kb < 16; kb++) {
    A1[kb] = m.frag1[(wave * 16 + kb) * 64 + lane];
    const bf16x8 w2 = m.frag2[(wave * 16 + kb) * 64 + lane];
    if (kb < kRegs2) A2[kb] = w2;
    else A2w[(kb - kA2Regs) * 64 + lane] = w2;
  }
  const float* bias1 = sp + kB1 + h * 128 + wave * 16;
  const float* bias2 = sp + kB2 + h * 128 + wave * 16;
  const float2* w3 = reinterpret_cast<const float2*>(sp + kW3) + h * 128 + wave * 16;
  const float4* w0 = reinterpret_cast<const float4*>(sp + kW0);
  const float* b0 = sp + kB0;
  const bool io_wave = wave < kFRows / 64;
  const int lr_io = wave * 64 + lane;  // the row an io wave's lane gathers and finishes
  const RunState& R = io.R;
  // The first iteration's inputs.  Later iterations' inputs are loaded
  // while the previous iteration computes (software pipelining): the queue
  // entry before layer 0, the particle state before layer 1, and X and pidx
  // are written after the finish.
  if (io_wave) {
    const uint32_t row = blockIdx.x * kFRows + lr_io;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    uint32_t p = 0;
    if (row < n) {
      if (kTrack) {
        p = R.QT[row];
        v = TslInputs(R, p);
      } else {
        v = reinterpret_cast<const float4*>(x)[row];
      }
    }
    X[lr_io] = Normalize(m, v);
    if (kTrack) pidx[lr_io] = p;
  }
  WaveTally t_tsl, t_fall, t_net;  // tracker: the io waves' event counts
  for (uint32_t it = blockIdx.x; it < n_it; it += gridDim.x) {
    const uint32_t row = it * kFRows + lr_io;
    const bool valid = io_wave && row < n;
    const uint32_t row_nx = (it + gridDim.x) * kFRows + lr_io;
    const bool pre = io_wave && it + gridDim.x < n_it && row_nx < n;
    // next iteration: queue entry (tracker) or input row (batch)
    uint32_t p_nx = 0;
    float4 v_nx = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (pre) {
      if (kTrack) p_nx = R.QT[row_nx];
      else v_nx = reinterpret_cast<const float4*>(x)[row_nx];
    }
    // this iteration's finish: the particle's energy, temperature and table
    uint32_t p = 0;
    double fE = 0, fT = 0;
    int fI = 0;
    if (kTrack && valid) {
      p = pidx[lr_io];
      fE = R.B.E[p];
      fT = R.B.tslT[p];
      fI = R.B.tslI[p];
    }
    __syncthreads();
    // layer 0 (4 -> 256): thread = (row, 64 consecutive positions)
    {
      const int lr = threadIdx.x % kFRows;
      const int c0 = threadIdx.x / kFRows;
      const float4 xn = X[lr];
#pragma unroll
      for (int g = 0; g < (256 * kFRows) / (64 * kMlpWaves) / 8; g++) {
        const int P0 = (256 * kFRows / (64 * kMlpWaves)) * c0 + 8 * g;
        const int kb = P0 / 16, hh = (P0 % 16) / 8;
        const int i0 = (hh * 16 + kb) * 8;  // packed index of FeatB(kb, hh, 0..7)
        bf16x8 v;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const float4 w = w0[i0 + q];
          const float z = b0[i0 + q] + w.x * xn.x + w.y * xn.y + w.z * xn.z + w.w * xn.w;
          v[q] = (__bf16)SiluFast(z);
        }
        *reinterpret_cast<bf16x8*>(H0 + lr * kHStride + P0) = v;
      }
    }
    __syncthreads();
    // next iteration (tracker): the queued particle's state; this
    // iteration's finish: its table's constants
    double nE = 0, nT = 0;
    uint32_t ng = 0;
    if (kTrack && pre) {
      nE = R.B.E[p_nx];
      nT = R.B.tslT[p_nx];
      ng = R.B.rng[p_nx];
    }
    double awr = 0, b_cut = 0;
    if (kTrack && valid) {
      awr = io.W.tsl[fI].awr;
      b_cut = io.W.tsl[fI].beta_cutoff;
    }
    f32x16 acc[kFCb];
    // layer 1 -> H1
    MfmaRows(H0, A1, h, j, acc);
#pragma unroll
    for (int c = 0; c < kFCb; c++)
#pragma unroll
      for (int jq = 0; jq < 4; jq++) {
        typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
        bf16x4 v;
#pragma unroll
        for (int e = 0; e < 4; e++)
          v[e] = (__bf16)SiluFast(acc[c][4 * jq + e] + bias1[4 * jq + e]);
        *reinterpret_cast<bf16x4*>(H1 + (32 * c + j) * kHStride + PosC(wave, 4 * jq, h)) = v;
      }
    __syncthreads();
    // layer 2, then this wave's share of layer 3 (256 -> 2) in FP32
    if constexpr (kTrack)
      MfmaRowsSplit(H1, A2, A2w, lane, h, j, acc);
    else
      MfmaRows(H1, A2, h, j, acc);
#pragma unroll
    for (int c = 0; c < kFCb; c++) {
      float o0 = 0.0f, o1 = 0.0f;
#pragma unroll
      for (int r = 0; r < 16; r++) {
        const float a = SiluFast(acc[c][r] + bias2[r]);
        const float2 w = w3[r];
        o0 += w.x * a;
        o1 += w.y * a;
      }
      o0 += __shfl_xor(o0, 32);
      o1 += __shfl_xor(o1, 32);
      if (h == 0) parts[wave * kFRows + 32 * c + j] = make_float2(o0, o1);
    }
    __syncthreads();
    if (!io_wave) continue;
    // finish
    float o0 = 0.0f, o1 = 0.0f;
#pragma unroll
    for (int w = 0; w < kMlpWaves; w++) {
      const float2 s = parts[w * kFRows + lr_io];
      o0 += s.x;
      o1 += s.y;
    }
    const float y1 = (o0 + m.b3[0]) * m.out_scale[0] + m.out_shift[0];
    const float y2 = (o1 + m.b3[1]) * m.out_scale[1] + m.out_shift[1];
    if (!kTrack) {
      if (valid) {
        y[size_t(row) * 2 + 0] = y1;
        y[size_t(row) * 2 + 1] = y2;
      }
      if (pre) X[lr_io] = Normalize(m, v_nx);
      continue;
    }
    // KTslApply of the event loop: the scatter with the network's (alpha,
    // beta); a particle outside the trained box or the kinematic support is
    // parked in the fallback queue for the POD sampler (MlpFlushFallback)
    bool alive = false, fall = false;
    if (valid) {
      const double E = fE, T = fT;
      // the particle's direction and random state load while the network's
      // outputs are mapped to (alpha, beta)
      V3 dir = LoadDir(R.B, p);
      uint32_t g = R.B.rng[p];
      double a = 0, b = 0;
      if (InBoxT(io, E, T) && OutputsToAlphaBeta(y1, y2, E, T, awr, b_cut, io.e_min, a, b)) {
        const double E_p = E + b * kBoltzmann * T;
        const double mu = (E + E_p - a * awr * kBoltzmann * T) / (2 * sqrt(E * E_p));
        const double phi = 2 * kPi * Canonical(g);
        dir = Rotate(dir, mu > 1 ? 1.0 : (mu < -1 ? -1.0 : mu), phi);
        if (io.W.n_est_tsl) ScoreAlphaBeta(io.W, R, R.B.hist[p], a, b);
        R.B.u[p] = dir.x; R.B.v[p] = dir.y; R.B.w[p] = dir.z;
        R.B.E[p] = E_p;
        R.B.rng[p] = g;
        alive = true;
      } else {
        fall = true;
      }
    }
    // the next iteration's inputs, from the state loaded above
    if (pre) {
      const double u1 = Canonical(ng), u2 = Canonical(ng);
      R.B.rng[p_nx] = ng;
      X[lr_io] = Normalize(m, make_float4(float(log(nE)), float(nT), float(u1), float(u2)));
      pidx[lr_io] = p_nx;
    }
    uint32_t tot;
    const uint32_t iq = WavePush(Ring(R, io.iter, kQ), alive, tot);
    if (alive) R.Q[io.qin ^ 1][iq] = p;
    const uint32_t ifb = WavePush(R.misc + kMiscFallback, fall, tot);
    if (fall) io.fb[ifb] = p;
    t_tsl.add(valid);
    t_fall.add(fall);
    t_net.add(valid && !fall);
  }
  if (kTrack && io_wave) {
    t_tsl.flush(Ev(io.R, kEvTsl));
    t_fall.flush(Ev(io.R, kEvFallback));
    t_net.flush(Ev(io.R, kEvMlp));
  }
}

// Packs the FP32 master parameters into the fused kernel's layouts.
__global__ void KMlpPack(const float* P, int w0_off, int b0_off, int w1_off, int b1_off,
                         int w2_off, int b2_off, int w3_off, __bf16* frag1, __bf16* frag2,
                         float* small) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kFragElems) {
    // i = ((t * 16 + kb) * 64 + lane) * 8 + q
    const int q = i % 8, lane = (i / 8) % 64, kb = (i / 512) % 16, t = i / 8192;
    const int row = 32 * t + lane % 32;
    const int col = FeatB(kb, lane / 32, q);
    frag1[i] = (__bf16)P[w1_off + row * kHid + col];
    frag2[i] = (__bf16)P[w2_off + row * kHid + col];
  }
  if (i < 2 * 16 * 8) {  // i = (h * 16 + kb) * 8 + q
    const int q = i % 8, kb = (i / 8) % 16, h = i / 128;
    const int f = FeatB(kb, h, q);
    for (int k = 0; k < 4; k++) small[kW0 + i * 4 + k] = P[w0_off + f * 4 + k];
    small[kB0 + i] = P[b0_off + f];
  }
  if (i < 2 * 8 * 16) {  // i = (h * 8 + t) * 16 + r
    const int r = i % 16, t = (i / 16) % 8, h = i / 128;
    const int f = FeatC(t, r, h);
    small[kB1 + i] = P[b1_off + f];
    small[kB2 + i] = P[b2_off + f];
    small[kW3 + i * 2 + 0] = P[w3_off + 0 * kHid + f];
    small[kW3 + i * 2 + 1] = P[w3_off + 1 * kHid + f];
  }
}

// ---------------------------------------------------------------------------
// The POD transport map (training targets): the quantile functions of the
// distributions the POD sampler draws from.  SampleBeta picks the lower
// (upper) tabulated energy with probability 1 - r (r), draws beta from that
// grid's piecewise-linear inverse CDF and rejects beta < -E/kT; the beta
// law is therefore a mixture of two truncated piecewise-linear laws, and
// y1 comes from inverting its CDF by bisection.  SampleAlpha picks a beta
// grid with the reference's interpolation factor, draws F uniformly between
// the CDF values of the kinematic bounds and rescales alpha to the exact
// bounds; the fraction y2 = (alpha - alpha_min) / (alpha_max - alpha_min)
// follows the mixture of the two grids' fraction laws, inverted the same way.

// POD coefficients of one (partition, x index, T): value(F_j) =
// sum_r cdf[j][r] c[r] (linear in T between the tabulated temperatures).
struct PodRow {
  const double* cdf;
  const double* F;
  int nF, rank;
  double c[16];
};
__device__ void MakeRow(const DevWorld& W, const DTslPart& P, int x_i, double T, PodRow& row) {
  const double* Ts = W.hot + P.T_off;
  const int cand = UpperBoundArr(Ts, P.nT, T);
  const int t_hi = cand == P.nT ? cand - 1 : cand;
  const int t_lo = t_hi == 0 ? t_hi : t_hi - 1;
  const double w = t_hi == t_lo ? 0.0 : (T - Ts[t_lo]) / (Ts[t_hi] - Ts[t_lo]);
  const double* S = W.hot + P.S_off;
  const double* vh = W.pool + P.xt_off + (x_i * P.nT + t_hi) * P.rank;
  const double* vl = W.pool + P.xt_off + (x_i * P.nT + t_lo) * P.rank;
  row.rank = P.rank < 16 ? P.rank : 16;
  for (int r = 0; r < row.rank; r++) row.c[r] = S[r] * (vl[r] + w * (vh[r] - vl[r]));
  row.cdf = W.hot + P.cdf_off;
  row.F = W.hot + P.F_off;
  row.nF = P.nF;
}
__device__ __forceinline__ double RowValue(const PodRow& row, int j) {
  const double* u = row.cdf + j * row.rank;
  double v = 0;
  for (int r = 0; r < row.rank; r++) v += u[r] * row.c[r];
  return v;
}
// CDF of the piecewise-linear law with knots (F_j, value_j), end points
// (0, lo_end) and (1, hi_end), evaluated at v.
__device__ double RowCdf(const PodRow& row, double lo_end, double hi_end, double v) {
  int first = 0, len = row.nF;
  while (len > 0) {
    const int half = len >> 1;
    const int mid = first + half;
    if (v < RowValue(row, mid)) {
      len = half;
    } else {
      first = mid + 1;
      len = len - half - 1;
    }
  }
  const int hi = first;
  const double F_lo = hi != 0 ? row.F[hi - 1] : 0.0;
  const double F_hi = hi != row.nF ? row.F[hi] : 1.0;
  const double a = hi != 0 ? RowValue(row, hi - 1) : lo_end;
  const double b = hi != row.nF ? RowValue(row, hi) : hi_end;
  if (!(b > a)) return v < a ? F_lo : F_hi;
  double c = F_lo + (v - a) * (F_hi - F_lo) / (b - a);
  return c < 0 ? 0 : (c > 1 ? 1 : c);
}

// Measure of F in [Fa, Fb] with phi(F) <= v, phi the piecewise-linear map
// through (0, lo_end), (F_j, value_j), (1, hi_end).  Exact also where the
// POD reconstruction makes phi non-monotone (then the sampler's law is not
// the inverse of phi, but this measure).
__device__ double RowMeasure(const PodRow& row, double lo_end, double hi_end, double v,
                             double Fa, double Fb) {
  double f0 = 0.0, a = lo_end, m = 0.0;
  for (int j = 0; j <= row.nF; j++) {
    const double f1 = j < row.nF ? row.F[j] : 1.0;
    const double b = j < row.nF ? RowValue(row, j) : hi_end;
    const double lo = f0 > Fa ? f0 : Fa, hi = f1 < Fb ? f1 : Fb;
    if (hi > lo && f1 > f0) {
      // phi on [lo, hi] is linear between pa and pb
      const double pa = a + (b - a) * (lo - f0) / (f1 - f0);
      const double pb = a + (b - a) * (hi - f0) / (f1 - f0);
      if (pa <= v && pb <= v) {
        m += hi - lo;
      } else if (pa <= v || pb <= v) {
        const double tcross = (v - pa) / (pb - pa);
        m += (pa <= v ? tcross : 1.0 - tcross) * (hi - lo);
      }
    }
    f0 = f1;
    a = b;
  }
  return m;
}

__device__ double PodMapBeta(const DevWorld& W, const DTsl& t, double E, double T, double u) {
  const double* Es = W.hot + t.Es_off;
  int e_hi = UpperBoundArr(Es, t.nEs, E);
  if (e_hi >= t.nEs) e_hi = t.nEs - 1;
  int g[2];
  double wgt[2];
  int ng = 0;
  if (e_hi == 0) {
    g[0] = 0; wgt[0] = 1; ng = 1;
  } else {
    const double r = (E - Es[e_hi - 1]) / (Es[e_hi] - Es[e_hi - 1]);
    g[0] = e_hi - 1; wgt[0] = 1 - r;
    g[1] = e_hi; wgt[1] = r;
    ng = 2;
  }
  const double kT = kBoltzmann * T;
  const double b_min = -E / kT;
  PodRow rows[2];
  double lo_end[2], c0[2];
  for (int q = 0; q < ng; q++) {
    int local;
    const int p_i = PartitionOf(W.hot + t.E_ends_off, t.n_bp, g[q], local);
    MakeRow(W, W.parts[t.bp_begin + p_i], local, T, rows[q]);
    lo_end[q] = -Es[g[q]] / kT;
    c0[q] = RowMeasure(rows[q], lo_end[q], t.beta_cutoff, b_min, 0.0, 1.0);  // rejected
  }
  auto mix = [&](double b) {
    double m = 0;
    for (int q = 0; q < ng; q++) {
      const double c = RowMeasure(rows[q], lo_end[q], t.beta_cutoff, b, 0.0, 1.0);
      const double d = 1 - c0[q];
      m += wgt[q] * (d > 0 ? (c - c0[q]) / d : 1.0);
    }
    return m;
  };
  double lo = b_min, hi = t.beta_cutoff;
  for (int it = 0; it < 34; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mix(mid) < u) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

__device__ double PodMapAlphaFrac(const DevWorld& W, const DTsl& t, double b, double E,
                                  double T, double u) {
  const double* betas = W.hot + t.betas_off;
  const double kT = kBoltzmann * T;
  const double ab = fabs(b);
  const int sgn = (0 < b) - (b < 0);
  int b_hi = UpperBoundArr(betas, t.nbetas, ab);
  if (b_hi >= t.nbetas) b_hi = t.nbetas - 1;
  // grid choice of SampleAlpha (ThermalScattering.cpp:355-367, quirk Q2)
  const bool snap_lower = (sgn == 1 && t.beta_cutoff <= betas[b_hi]) ||
                          (sgn == -1 && -betas[b_hi] < -E / kT);
  const bool snap_min = b_hi == 0;
  double r = snap_lower ? 0
             : snap_min ? 1
                        : (ab - betas[b_hi - 1] / (betas[b_hi] - betas[b_hi - 1]));
  r = r < 0 ? 0 : (r > 1 ? 1 : r);
  int g[2];
  double wgt[2];
  int ng = 0;
  if (b_hi == 0) {
    g[0] = 0; wgt[0] = 1; ng = 1;
  } else {
    g[0] = b_hi - 1; wgt[0] = 1 - r;
    g[1] = b_hi; wgt[1] = r;
    ng = 2;
  }
  const double sqrt_E = sqrt(E);
  const double akT = t.awr * kT;
  PodRow rows[2];
  double amin[2], amax[2], Fmin[2], Fmax[2], m0[2], m1[2];
  for (int q = 0; q < ng; q++) {
    double b_s = sgn * betas[g[q]];
    if (E + b_s * kT < 0) b_s = b;
    const double root = sqrt(E + b_s * kT);
    amin[q] = (sqrt_E - root) * (sqrt_E - root) / akT;
    amax[q] = (sqrt_E + root) * (sqrt_E + root) / akT;
    int local;
    const int p_i = PartitionOf(W.hot + t.beta_ends_off, t.n_ap, g[q], local);
    MakeRow(W, W.parts[t.ap_begin + p_i], local, T, rows[q]);
    // find_cdf of the reference (monotone search) fixes the F window ...
    Fmin[q] = RowCdf(rows[q], 0.0, t.alpha_cutoff, amin[q]);
    Fmax[q] = RowCdf(rows[q], 0.0, t.alpha_cutoff, amax[q]);
    // ... and values outside (a_min, a_max) are rejected (resampled)
    m0[q] = RowMeasure(rows[q], 0.0, t.alpha_cutoff, amin[q], Fmin[q], Fmax[q]);
    m1[q] = RowMeasure(rows[q], 0.0, t.alpha_cutoff, amax[q], Fmin[q], Fmax[q]);
  }
  auto mix = [&](double f) {
    double m = 0;
    for (int q = 0; q < ng; q++) {
      const double d = m1[q] - m0[q];
      if (!(d > 0) || !(amax[q] > amin[q])) {
        m += wgt[q] * (f >= 0.5 ? 1.0 : 0.0);
        continue;
      }
      const double c = RowMeasure(rows[q], 0.0, t.alpha_cutoff,
                                  amin[q] + f * (amax[q] - amin[q]), Fmin[q], Fmax[q]);
      m += wgt[q] * (c - m0[q]) / d;
    }
    return m;
  };
  double lo = 0, hi = 1;
  for (int it = 0; it < 26; it++) {
    const double mid = 0.5 * (lo + hi);
    if (mix(mid) < u) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

__device__ void PodMap(const DevWorld& W, const DTsl& t, double E, double T, double u1,
                       double u2, double& y1, double& y2) {
  const double b = PodMapBeta(W, t, E, T, u1);
  y1 = b * (kBoltzmann * T) / (E + kBoltzmann * T);
  y2 = PodMapAlphaFrac(W, t, b, E, T, u2);
}
// (y1, y2) -> (alpha, beta) in the trained parametrisation.
__global__ void KPodMap(DevWorld W, int tsl, uint32_t n, const float* x, float* y) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DTsl t = W.tsl[tsl];
  const double E = exp(double(x[size_t(i) * 4 + 0]));
  const double T = x[size_t(i) * 4 + 1];
  double y1, y2;
  PodMap(W, t, E, T, x[size_t(i) * 4 + 2], x[size_t(i) * 4 + 3], y1, y2);
  y[size_t(i) * 2 + 0] = float(y1);
  y[size_t(i) * 2 + 1] = float(y2);
}

// ---------------------------------------------------------------------------
// Tracker step with the surrogate.
__device__ __forceinline__ bool InBox(const MlpState& m, double E, double T) {
  return E >= m.e_min && E <= m.e_max && T >= m.t_lo && T <= m.t_hi;
}
// The POD sampler for the parked fallback particles; the survivors join the
// live queue of iteration `iter` (Q[qin], length in ring slot iter - 1).
// kLds: the sampler's tables are staged in LDS first (as KTslPod<true>): the
// few parked samples each wait on a chain of bisection probes.
template <bool kLds>
__global__ void __launch_bounds__(256) KTslPodFallback(DevWorld W, RunState R, int iter,
                                                        int qin, const uint32_t* fb) {
  extern __shared__ double lds_hot[];
  const uint32_t n = R.misc[kMiscFallback];
  if (blockIdx.x * blockDim.x >= n) return;  // no work: skip the staging too
  const double* H = kLds ? StageHot(W.hot, lds_hot, W.hot_n) : W.hot;
  for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
    const uint32_t q = base + threadIdx.x;
    bool alive = false;
    uint32_t p = 0;
    if (q < n) {
      p = fb[q];
      alive = TslPodEvent(W, R, p, H);
    }
    uint32_t tot;
    const uint32_t iq = WavePush(Ring(R, iter - 1, kQ), alive, tot);
    if (alive) R.Q[qin][iq] = p;
  }
}
__global__ void KClearFallback(RunState R) { R.misc[kMiscFallback] = 0; }
// Clears one error word (a kernel, not a memset: no memset node in graphs).
__global__ void KClearWord(uint32_t* w) { *w = 0u; }

static MlpDev DevView(const MlpState& m) {
  MlpDev d;
  d.frag1 = reinterpret_cast<const bf16x8*>(m.packed);
  d.frag2 = reinterpret_cast<const bf16x8*>(m.packed) + kFragElems / 8;
  d.small = reinterpret_cast<const float*>(
      reinterpret_cast<const char*>(m.packed) + 2 * kFragElems * sizeof(__bf16));
  d.b3[0] = m.b3[0];
  d.b3[1] = m.b3[1];
  for (int k = 0; k < 4; k++) {
    d.in_shift[k] = m.in_shift[k];
    d.in_scale[k] = m.in_scale[k];
  }
  for (int k = 0; k < 2; k++) {
    d.out_shift[k] = m.out_shift[k];
    d.out_scale[k] = m.out_scale[k];
  }
  return d;
}

// Batch inference of n rows.
static void LaunchInfer(const MlpState& m, const float* x, float* y, uint32_t n, hipStream_t s) {
  const TrackIO none{};
  const uint32_t grid = uint32_t(std::min<uint64_t>((uint64_t(n) + kFRows - 1) / kFRows, 256));
  hipLaunchKernelGGL(KMlpF<false>, dim3(std::max(grid, 1u)), dim3(64 * kMlpWaves), kFusedLds,
                     s, DevView(m), x, y, nullptr, n, none);
}
static int InferChunked(mmc_ctx* ctx, MlpState& m, const float* x, float* y, uint64_t n) {
  constexpr uint64_t kChunk = 1ull << 30;  // rows per launch (32-bit row indices)
  for (uint64_t c0 = 0; c0 < n; c0 += kChunk) {
    const uint32_t c = uint32_t(std::min<uint64_t>(n - c0, kChunk));
    LaunchInfer(m, x + c0 * 4, y + c0 * 2, c, MmcStream(ctx));
  }
  return MMC_OK;
}

int MlpTslStep(MlpState& m, const DevWorld& W, const RunState& R, int iter, int qin,
               uint32_t grid, hipStream_t s, void (*mark)(void*, int), void* mark_arg) {
  if (!m.ready || !m.fb) return 1;
  TrackIO io{W, R, iter, qin, m.fb, m.e_min, m.e_max, m.t_lo, m.t_hi};
  // both kernels read the queue length themselves: one block per CU
  mark(mark_arg, MMC_K_MLP);
  hipLaunchKernelGGL(KMlpF<true>, dim3(256), dim3(64 * kMlpWaves), kFusedLds, s, DevView(m),
                     nullptr, nullptr, Ring(R, iter, kT), 0u, io);
  mark(mark_arg, -1);
  return 0;
}

int MlpFlushFallback(MlpState& m, const DevWorld& W, const RunState& R, int iter, int qin,
                     size_t hot_lds, hipStream_t s) {
  if (!m.ready || !m.fb) return 1;
  if (hot_lds)
    hipLaunchKernelGGL(KTslPodFallback<true>, dim3(64), dim3(256), hot_lds, s, W, R, iter, qin,
                       m.fb);
  else
    hipLaunchKernelGGL(KTslPodFallback<false>, dim3(64), dim3(256), 0, s, W, R, iter, qin,
                       m.fb);
  hipLaunchKernelGGL(KClearFallback, dim3(1), dim3(1), 0, s, R);
  return 0;
}

// ---------------------------------------------------------------------------
// FP32 training kernels.
// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[N]); optional SiLU copy.
// op(A)(m, k) = TA ? A[k * lda + m] : A[m * lda + k]
// op(B)(k, n) = TB ? B[n * ldb + k] : B[k * ldb + n]
template <bool TA, bool TB>
__global__ void __launch_bounds__(256)
    KGemm(int M, int N, int K, const float* __restrict__ A, int lda,
          const float* __restrict__ Bm, int ldb, float* __restrict__ C, int ldc,
          const float* __restrict__ bias, float* __restrict__ H) {
  __shared__ float As[16][64 + 4];
  __shared__ float Bs[16][64 + 4];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  // split K over gridDim.z (partial sums added atomically when > 1)
  const int kc = (K + gridDim.z - 1) / gridDim.z;
  const int k_begin = blockIdx.z * kc;
  const int k_end = min(K, k_begin + kc);
  float acc[4][4] = {};
  for (int k0 = k_begin; k0 < k_end; k0 += 16) {
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
      int mm, kk;
      if (TA) {
        mm = e % 64;
        kk = e / 64;
      } else {
        kk = e % 16;
        mm = e / 16;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      float v = 0.0f;
      if (gm < M && gk < k_end) v = TA ? A[size_t(gk) * lda + gm] : A[size_t(gm) * lda + gk];
      As[kk][mm] = v;
      int nn, kb;
      if (TB) {
        kb = e % 16;
        nn = e / 16;
      } else {
        nn = e % 64;
        kb = e / 64;
      }
      const int gn = n0 + nn, gk2 = k0 + kb;
      float w = 0.0f;
      if (gn < N && gk2 < k_end) w = TB ? Bm[size_t(gn) * ldb + gk2] : Bm[size_t(gk2) * ldb + gn];
      Bs[kb][nn] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int i = 0; i < 4; i++) b[i] = Bs[kk][tx * 4 + i];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int jj = 0; jj < 4; jj++) acc[i][jj] += a[i] * b[jj];
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int gm = m0 + ty * 4 + i;
    if (gm >= M) continue;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int gn = n0 + tx * 4 + jj;
      if (gn >= N) continue;
      float v = acc[i][jj];
      if (gridDim.z > 1) {
        atomicAdd(C + size_t(gm) * ldc + gn, v);
        continue;
      }
      if (bias) v += bias[gn];
      C[size_t(gm) * ldc + gn] = v;
      if (H) H[size_t(gm) * ldc + gn] = v / (1.0f + expf(-v));
    }
  }
}

template <bool TA, bool TB>
static void Gemm(hipStream_t s, int M, int N, int K, const float* A, int lda, const float* B,
                 int ldb, float* C, int ldc, const float* bias = nullptr, float* H = nullptr,
                 int split = 1) {
  dim3 grid((N + 63) / 64, (M + 63) / 64, split);
  hipLaunchKernelGGL((KGemm<TA, TB>), grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc,
                     bias, H);
}

// dZ = dH * silu'(Z)
__global__ void KSiluBack(const float* Z, const float* dH, float* dZ, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const float z = Z[i];
    const float sg = 1.0f / (1.0f + expf(-z));
    dZ[i] = dH[i] * sg * (1.0f + z * (1.0f - sg));
  }
}
// dZ = (Z - Y) / (rows * outs); loss partial sums
__global__ void KLossGrad(const float* Z, const float* Y, const float* w, float* dZ, size_t n,
                          float inv, float* loss) {
  float acc = 0.0f;
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const float d = Z[i] - Y[i];
    const float wd = w ? w[i] * d : d;
    dZ[i] = wd * inv;
    acc += wd * d;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if (threadIdx.x % 64 == 0) atomicAdd(loss, acc * 0.5f * inv);
}
// db[n] += sum over rows of dZ[row][n]
__global__ void KColSum(const float* dZ, int rows, int N, float* db) {
  const int n = blockIdx.x * 64 + threadIdx.x % 64;
  const int r0 = blockIdx.y * 1024 + threadIdx.x / 64;
  float acc = 0.0f;
  if (n < N)
    for (int r = r0; r < min(rows, int(blockIdx.y + 1) * 1024); r += 4) acc += dZ[size_t(r) * N + n];
  __shared__ float s[256];
  s[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64 && n < N)
    atomicAdd(db + n, s[threadIdx.x] + s[threadIdx.x + 64] + s[threadIdx.x + 128] +
                          s[threadIdx.x + 192]);
}
__global__ void KAdam(float* p, const float* g, float* m, float* v, size_t n, float lr,
                      float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}
__global__ void KZero(float* p, size_t n) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x)
    p[i] = 0.0f;
}
// x_norm = (x - shift) * scale, y_norm = (y - shift) / scale, in place
__global__ void KNormalize(float* x, float* y, size_t n, MlpDev d) {
  for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n;
       i += size_t(gridDim.x) * blockDim.x) {
    for (int k = 0; k < 4; k++) x[i * 4 + k] = (x[i * 4 + k] - d.in_shift[k]) * d.in_scale[k];
    if (y)
      for (int k = 0; k < 2; k++) y[i * 2 + k] = (y[i * 2 + k] - d.out_shift[k]) / d.out_scale[k];
  }
}

}  // namespace mmc

// ---------------------------------------------------------------------------
// Host side
static bool FusedShape(const MlpState& m) {
  return m.n_layers == 4 && m.widths[0] == 4 && m.widths[1] == kHid && m.widths[2] == kHid &&
         m.widths[3] == kHid && m.widths[4] == 2;
}

static int Repack(mmc_ctx* ctx) {
  MlpState& m = MmcMlp(ctx);
  if (!FusedShape(m)) return MMC_OK;
  const int total = kFragElems;
  char* base = reinterpret_cast<char*>(m.packed);
  __bf16* f1 = reinterpret_cast<__bf16*>(base);
  __bf16* f2 = f1 + kFragElems;
  float* small = reinterpret_cast<float*>(base + 2 * kFragElems * sizeof(__bf16));
  hipLaunchKernelGGL(KMlpPack, dim3((total + 255) / 256), dim3(256), 0, MmcStream(ctx),
                     m.params, m.w_off[0], m.b_off[0], m.w_off[1], m.b_off[1], m.w_off[2],
                     m.b_off[2], m.w_off[3], f1, f2, small);
  // layer-3 bias lives in the kernel arguments
  float b3[2];
  if (hipMemcpyAsync(b3, m.params + m.b_off[3], sizeof(b3), hipMemcpyDeviceToHost,
                     MmcStream(ctx)) != hipSuccess ||
      hipStreamSynchronize(MmcStream(ctx)) != hipSuccess)
    return MmcFail(ctx, MMC_ERR_HIP, "mlp repack failed");
  m.b3[0] = b3[0];
  m.b3[1] = b3[1];
  return MMC_OK;
}

extern "C" int mmc_mlp_set(mmc_ctx* ctx, int tsl, const mmc_mlp_desc* d, const float* params,
                           uint64_t n_params) {
  if (!ctx || !d || d->n_layers < 1 || d->n_layers > kMlpMaxLayers) return MMC_ERR_ARG;
  if (tsl < 0 || tsl >= MmcTslCount(ctx)) return MmcFail(ctx, MMC_ERR_ARG, "no such tsl table");
  if (hipSetDevice(MmcDevice(ctx)) != hipSuccess) return MMC_ERR_HIP;
  MlpState& m = MmcMlp(ctx);
  MlpFree(m);
  m.tsl = tsl;
  m.n_layers = d->n_layers;
  long off = 0;
  for (int l = 0; l <= d->n_layers; l++) {
    if (d->widths[l] < 1) return MmcFail(ctx, MMC_ERR_ARG, "bad mlp width");
    m.widths[l] = d->widths[l];
  }
  for (int l = 0; l < d->n_layers; l++) {
    m.w_off[l] = int(off);
    off += long(m.widths[l + 1]) * m.widths[l];
    m.b_off[l] = int(off);
    off += m.widths[l + 1];
  }
  m.n_params = off;
  if (params && n_params != uint64_t(off)) return MmcFail(ctx, MMC_ERR_ARG, "params size");
  std::memcpy(m.in_shift, d->in_shift, sizeof(m.in_shift));
  std::memcpy(m.in_scale, d->in_scale, sizeof(m.in_scale));
  std::memcpy(m.out_shift, d->out_shift, sizeof(m.out_shift));
  std::memcpy(m.out_scale, d->out_scale, sizeof(m.out_scale));
  m.e_min = d->e_min;
  m.e_max = d->e_max;
  m.t_lo = d->t_lo;
  m.t_hi = d->t_hi;
  const size_t pbytes = sizeof(float) * off;
  if (hipMalloc(&m.params, pbytes) != hipSuccess || hipMalloc(&m.grads, pbytes) != hipSuccess ||
      hipMalloc(&m.adam_m, pbytes) != hipSuccess || hipMalloc(&m.adam_v, pbytes) != hipSuccess ||
      hipMalloc(&m.loss, sizeof(float)) != hipSuccess ||
      hipMalloc(&m.packed, 2 * kFragElems * sizeof(__bf16) + kSmall * sizeof(float)) != hipSuccess)
    return MmcFail(ctx, MMC_ERR_HIP, "mlp allocation failed");
  hipMemsetAsync(m.adam_m, 0, pbytes, MmcStream(ctx));
  hipMemsetAsync(m.adam_v, 0, pbytes, MmcStream(ctx));
  m.adam_t = 0;
  if (params) {
    if (hipMemcpy(m.params, params, pbytes, hipMemcpyHostToDevice) != hipSuccess)
      return MmcFail(ctx, MMC_ERR_HIP, "mlp upload failed");
  } else {
    // He-style uniform initialisation from a fixed LCG
    std::vector<float> h(off);
    uint64_t st = 0x9E3779B97F4A7C15ULL;
    auto rnd = [&]() {
      st = st * 6364136223846793005ULL + 1442695040888963407ULL;
      return double(st >> 11) / double(1ULL << 53);
    };
    for (int l = 0; l < m.n_layers; l++) {
      const double lim = std::sqrt(6.0 / m.widths[l]);
      for (long i = 0; i < long(m.widths[l + 1]) * m.widths[l]; i++)
        h[m.w_off[l] + i] = float((2 * rnd() - 1) * lim);
      for (int i = 0; i < m.widths[l + 1]; i++) h[m.b_off[l] + i] = 0.0f;
    }
    if (hipMemcpy(m.params, h.data(), pbytes, hipMemcpyHostToDevice) != hipSuccess)
      return MmcFail(ctx, MMC_ERR_HIP, "mlp upload failed");
  }
  m.ready = FusedShape(m);
  const int rc = Repack(ctx);
  if (rc) return rc;
  return SyncChecked(ctx, "mmc_mlp_set");
}

extern "C" int mmc_mlp_get(mmc_ctx* ctx, float* params, uint64_t n_params) {
  if (!ctx || !params) return MMC_ERR_ARG;
  MlpState& m = MmcMlp(ctx);
  if (!m.params || n_params != uint64_t(m.n_params)) return MMC_ERR_ARG;
  if (hipStreamSynchronize(MmcStream(ctx)) != hipSuccess ||
      hipMemcpy(params, m.params, sizeof(float) * m.n_params, hipMemcpyDeviceToHost) !=
          hipSuccess)
    return MMC_ERR_HIP;
  return MMC_OK;
}

extern "C" int mmc_mlp_infer(mmc_ctx* ctx, uint64_t n, const float* x, float* y) {
  if (!ctx || !x || !y) return MMC_ERR_ARG;
  MlpState& m = MmcMlp(ctx);
  if (!m.ready) return MmcFail(ctx, MMC_ERR_UNSUPPORTED, "no 4-256-256-256-2 network set");
  if (n == 0) return MMC_OK;
  if (n > 0xffffffffull) return MmcFail(ctx, MMC_ERR_ARG, "batch too large");
  const int rc = InferChunked(ctx, m, x, y, n);
  if (rc) return rc;
  return SyncChecked(ctx, "mmc_mlp_infer");
}

extern "C" int mmc_tsl_pod_map(mmc_ctx* ctx, int tsl, uint64_t n, const float* x, float* y) {
  if (!ctx || !x || !y) return MMC_ERR_ARG;
  if (tsl < 0 || tsl >= MmcTslCount(ctx)) return MmcFail(ctx, MMC_ERR_ARG, "no such tsl table");
  if (n == 0) return MMC_OK;
  if (n > 0xffffffffull) return MMC_ERR_ARG;
  hipLaunchKernelGGL(KPodMap, dim3(uint32_t((n + 255) / 256)), dim3(256), 0, MmcStream(ctx),
                     MmcWorld(ctx), tsl, uint32_t(n), x, y);
  return SyncChecked(ctx, "mmc_tsl_pod_map");
}

// Forward + backward of one minibatch (normalised x, y) into m.grads.
static int ForwardBackward(mmc_ctx* ctx, uint64_t rows, const float* x, const float* y,
                           const float* w = nullptr) {
  MlpState& m = MmcMlp(ctx);
  hipStream_t s = MmcStream(ctx);
  const int L = m.n_layers;
  int wmax = 0;
  for (int l = 0; l <= L; l++) wmax = std::max(wmax, m.widths[l]);
  const size_t need = size_t(rows) * wmax * (3 * L + 2);
  if (m.act_cap < need) {
    if (m.act) (void)hipFree(m.act);
    m.act = nullptr;
    if (hipMalloc(&m.act, need * sizeof(float)) != hipSuccess)
      return MmcFail(ctx, MMC_ERR_HIP, "activation allocation failed");
    m.act_cap = need;
  }
  // per layer l: Z_l and H_l (rows x widths[l+1]); dZ scratch, dH scratch
  std::vector<float*> Z(L), H(L);
  float* p = m.act;
  for (int l = 0; l < L; l++) {
    Z[l] = p;
    p += size_t(rows) * wmax;
    H[l] = p;
    p += size_t(rows) * wmax;
  }
  float* dZ = p;
  p += size_t(rows) * wmax;
  float* dH = p;
  const int R = int(rows);
  for (int l = 0; l < L; l++) {
    const float* in = l == 0 ? x : H[l - 1];
    Gemm<false, true>(s, R, m.widths[l + 1], m.widths[l], in, m.widths[l], m.params + m.w_off[l],
                      m.widths[l], Z[l], m.widths[l + 1], m.params + m.b_off[l],
                      l + 1 < L ? H[l] : nullptr);
  }
  hipLaunchKernelGGL(KZero, dim3(256), dim3(256), 0, s, m.grads, size_t(m.n_params));
  hipLaunchKernelGGL(KZero, dim3(1), dim3(64), 0, s, m.loss, size_t(1));
  const int no = m.widths[L];
  hipLaunchKernelGGL(KLossGrad, dim3(1024), dim3(256), 0, s, Z[L - 1], y, w, dZ,
                     size_t(R) * no, 1.0f / float(R * no), m.loss);
  for (int l = L - 1; l >= 0; l--) {
    const int nout = m.widths[l + 1], nin = m.widths[l];
    const float* in = l == 0 ? x : H[l - 1];
    // dW_l = dZ^T in, K = rows split over blocks (grads were zeroed)
    const int tiles = ((nout + 63) / 64) * ((nin + 63) / 64);
    int split = std::max(1, std::min(64, 512 / tiles));
    split = std::min(split, std::max(1, R / 256));
    Gemm<true, false>(s, nout, nin, R, dZ, nout, in, nin, m.grads + m.w_off[l], nin, nullptr,
                      nullptr, split);
    hipLaunchKernelGGL(KColSum, dim3((nout + 63) / 64, (R + 1023) / 1024), dim3(256), 0, s, dZ,
                       R, nout, m.grads + m.b_off[l]);
    if (l > 0) {
      // dH_{l-1} = dZ W_l ; dZ_{l-1} = dH * silu'(Z_{l-1})
      Gemm<false, false>(s, R, nin, nout, dZ, nout, m.params + m.w_off[l], nin, dH, nin);
      hipLaunchKernelGGL(KSiluBack, dim3(1024), dim3(256), 0, s, Z[l - 1], dH, dZ,
                         size_t(R) * nin);
    }
  }
  return MMC_OK;
}

extern "C" int mmc_mlp_grad(mmc_ctx* ctx, uint64_t n, const float* x_host, const float* y_host,
                            float* loss, float* grads) {
  if (!ctx || !x_host || !y_host || n == 0) return MMC_ERR_ARG;
  MlpState& m = MmcMlp(ctx);
  if (!m.params) return MMC_ERR_ARG;
  const int ni = m.widths[0], no = m.widths[m.n_layers];
  float *dx = nullptr, *dy = nullptr;
  if (hipMalloc(&dx, sizeof(float) * n * ni) != hipSuccess ||
      hipMalloc(&dy, sizeof(float) * n * no) != hipSuccess)
    return MMC_ERR_HIP;
  hipMemcpy(dx, x_host, sizeof(float) * n * ni, hipMemcpyHostToDevice);
  hipMemcpy(dy, y_host, sizeof(float) * n * no, hipMemcpyHostToDevice);
  int rc = ForwardBackward(ctx, n, dx, dy);
  if (!rc) {
    hipStreamSynchronize(MmcStream(ctx));
    if (loss) hipMemcpy(loss, m.loss, sizeof(float), hipMemcpyDeviceToHost);
    if (grads) hipMemcpy(grads, m.grads, sizeof(float) * m.n_params, hipMemcpyDeviceToHost);
  }
  (void)hipFree(dx);
  (void)hipFree(dy);
  return rc;
}

extern "C" int mmc_mlp_train(mmc_ctx* ctx, uint64_t n, const float* x, const float* y,
                             const mmc_train_opts* o, float* last_loss) {
  if (!ctx || !x || !y || !o || o->batch == 0 || o->batch > n) return MMC_ERR_ARG;
  MlpState& m = MmcMlp(ctx);
  if (!m.params) return MMC_ERR_ARG;
  hipStream_t s = MmcStream(ctx);
  uint64_t off = o->offset % n;
  const int ni = m.widths[0], no = m.widths[m.n_layers];
  for (uint64_t step = 0; step < o->steps; step++) {
    if (off + o->batch > n) off = 0;
    int rc = ForwardBackward(ctx, o->batch, x + off * ni, y + off * no,
                             o->weights ? o->weights + off * no : nullptr);
    if (rc) return rc;
    m.adam_t++;
    const float c1 = 1.0f - std::pow(o->beta1, float(m.adam_t));
    const float c2 = 1.0f - std::pow(o->beta2, float(m.adam_t));
    hipLaunchKernelGGL(KAdam, dim3(256), dim3(256), 0, s, m.params, m.grads, m.adam_m, m.adam_v,
                       size_t(m.n_params), o->lr, o->beta1, o->beta2, o->eps, c1, c2);
    off += o->batch;
  }
  const int rc = Repack(ctx);
  if (rc) return rc;
  if (last_loss) hipMemcpyAsync(last_loss, m.loss, sizeof(float), hipMemcpyDeviceToHost, s);
  return SyncChecked(ctx, "mmc_mlp_train");
}

// Device batch: normalise raw inputs/targets in place (training data prep).
extern "C" int mmc_mlp_normalize(mmc_ctx* ctx, uint64_t n, float* x, float* y) {
  if (!ctx || !x) return MMC_ERR_ARG;
  MlpState& m = MmcMlp(ctx);
  if (!m.params) return MMC_ERR_ARG;
  hipLaunchKernelGGL(KNormalize, dim3(1024), dim3(256), 0, MmcStream(ctx), x, y, size_t(n),
                     DevView(m));
  return SyncChecked(ctx, "mmc_mlp_normalize");
}

__global__ void KSurrogateBatch(DevWorld W, MlpState m, int tsl, uint32_t n, const double* E,
                                const double* T, uint32_t* rng, float* xin) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t g = rng[i];
  const double u1 = Canonical(g), u2 = Canonical(g);
  rng[i] = g;
  xin[size_t(i) * 4 + 0] = float(log(E[i]));
  xin[size_t(i) * 4 + 1] = float(T[i]);
  xin[size_t(i) * 4 + 2] = float(u1);
  xin[size_t(i) * 4 + 3] = float(u2);
}
__global__ void KSurrogateFinish(DevWorld W, MlpState m, int tsl, uint32_t n, const double* E,
                                 const double* T, uint32_t* rng, const float* yout,
                                 double* alpha, double* beta, uint32_t* err) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const DTsl t = W.tsl[tsl];
  double a, b;
  if (!InBox(m, E[i], T[i]) ||
      !OutputsToAlphaBeta(yout[size_t(i) * 2], yout[size_t(i) * 2 + 1], E[i], T[i], t.awr,
                          t.beta_cutoff, m.e_min, a, b)) {
    uint32_t g = rng[i];
    uint32_t e = SampleBeta(W, W.hot, t, g, E[i], T[i], b) ? 0u : uint32_t(kErrResample);
    if (!e) e = SampleAlpha(W, W.hot, t, g, b, E[i], T[i], a);
    if (e) {
      atomicOr(err, e);
      a = b = __builtin_nan("");
    }
    rng[i] = g;
  }
  alpha[i] = a;
  beta[i] = b;
}

extern "C" int mmc_tsl_sample_surrogate(mmc_ctx* ctx, int tsl, uint64_t n, const double* E,
                                        const double* T, uint32_t* rng, double* alpha,
                                        double* beta) {
  if (!ctx || !E || !T || !rng || !alpha || !beta) return MMC_ERR_ARG;
  if (tsl < 0 || tsl >= MmcTslCount(ctx)) return MmcFail(ctx, MMC_ERR_ARG, "no such tsl table");
  MlpState& m = MmcMlp(ctx);
  if (!m.ready) return MmcFail(ctx, MMC_ERR_UNSUPPORTED, "no surrogate network set");
  if (n == 0) return MMC_OK;
  if (n > 0xffffffffull) return MMC_ERR_ARG;
  hipStream_t s = MmcStream(ctx);
  // the network's inputs and outputs and the error word, kept by the context
  // and grown to the largest batch seen
  if (m.batch_cap < n) {
    if (hipStreamSynchronize(s) != hipSuccess) return MMC_ERR_HIP;
    for (void* p : {static_cast<void*>(m.xin), static_cast<void*>(m.yout),
                    static_cast<void*>(m.berr)})
      if (p) (void)hipFree(p);
    m.xin = m.yout = nullptr;
    m.berr = nullptr;
    m.batch_cap = 0;
    if (hipMalloc(&m.xin, sizeof(float) * 4 * n) != hipSuccess ||
        hipMalloc(&m.yout, sizeof(float) * 2 * n) != hipSuccess ||
        hipMalloc(&m.berr, sizeof(uint32_t)) != hipSuccess)
      return MmcFail(ctx, MMC_ERR_HIP, "surrogate staging allocation failed");
    m.batch_cap = n;
  }
  float* xin = m.xin;
  float* yout = m.yout;
  uint32_t* err = m.berr;
  hipLaunchKernelGGL(KClearWord, dim3(1), dim3(1), 0, s, err);
  const uint32_t nb = uint32_t((n + 255) / 256);
  hipLaunchKernelGGL(KSurrogateBatch, dim3(nb), dim3(256), 0, s, MmcWorld(ctx), m, tsl,
                     uint32_t(n), E, T, rng, xin);
  if (InferChunked(ctx, m, xin, yout, n)) return MMC_ERR_HIP;
  hipLaunchKernelGGL(KSurrogateFinish, dim3(nb), dim3(256), 0, s, MmcWorld(ctx), m, tsl,
                     uint32_t(n), E, T, rng, yout, alpha, beta, err);
  uint32_t herr = 0;
  if (hipMemcpyAsync(&herr, err, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
    return MmcFail(ctx, MMC_ERR_HIP, "error word readback failed");
  if (const int rc = SyncChecked(ctx, "mmc_tsl_sample_surrogate")) return rc;
  if (herr) return MmcFail(ctx, MMC_ERR_RESAMPLE, "POD fallback failed");
  return MMC_OK;
}

// Staging buffers for the tracker's surrogate step (sized to the bank).
int MlpEnsureStaging(mmc_ctx* ctx, uint32_t cap) {
  MlpState& m = MmcMlp(ctx);
  if (m.stage_cap >= cap) return MMC_OK;
  if (m.fb) (void)hipFree(m.fb);
  m.fb = nullptr;
  if (hipMalloc(&m.fb, sizeof(uint32_t) * size_t(cap)) != hipSuccess)
    return MmcFail(ctx, MMC_ERR_HIP, "surrogate staging allocation failed");
  m.stage_cap = cap;
  return MMC_OK;
}

// ---------------------------------------------------------------------------
// Surrogate sampling without allocation (mmc_tsl_sample_surrogate_noalloc),
// for callers that cannot allocate device memory.  The staging of
// mmc_tsl_sample_surrogate (the network's inputs and outputs, the error
// word) lives in `scratch`, a file-scope device array that each GPU holds
// until the process ends: kNoAllocRows rows, 6 MiB.  Its host-side name is a
// stand-in, not a device address; MlpScratchInit looks the address up once
// per device (hipGetSymbolAddress, with that device current) and every
// launch below uses that address.  Every call on one GPU shares the array:
// calls that overlap on one GPU race, and each call clears the error word
// before its kernels run.
constexpr uint64_t kNoAllocRows = 1u << 18;
constexpr size_t kScratchX = 0;                                                // float4 rows
constexpr size_t kScratchY = kScratchX + sizeof(float) * 4 * kNoAllocRows;    // float2 rows
constexpr size_t kScratchErr = kScratchY + sizeof(float) * 2 * kNoAllocRows;  // u32
constexpr size_t kScratchBytes = kScratchErr + 256;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
static void* g_scratch[kScratchDevices] = {};
static std::mutex g_scratch_mu;

// The address of `scratch` on `device`, which must be current; 0 on success.
int MlpScratchInit(int device) {
  if (device < 0 || device >= kScratchDevices) return MMC_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return MMC_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MMC_ERR_HIP;
  g_scratch[device] = p;
  return MMC_OK;
}

extern "C" int mmc_tsl_sample_surrogate_noalloc(mmc_ctx* ctx, int tsl, uint64_t n,
                                                const double* E, const double* T,
                                                uint32_t* rng, double* alpha, double* beta) {
  if (!ctx || !E || !T || !rng || !alpha || !beta) return MMC_ERR_ARG;
  if (tsl < 0 || tsl >= MmcTslCount(ctx)) return MmcFail(ctx, MMC_ERR_ARG, "no such tsl table");
  MlpState& m = MmcMlp(ctx);
  if (!m.ready) return MmcFail(ctx, MMC_ERR_UNSUPPORTED, "no surrogate network set");
  if (n > kNoAllocRows)
    return MmcFail(ctx, MMC_ERR_ARG,
                   "batch larger than the static scratch of the no-allocation sampler "
                   "(262144 rows)");
  if (n == 0) return MMC_OK;
  const int dev = MmcDevice(ctx);
  if (hipSetDevice(dev) != hipSuccess) return MmcFail(ctx, MMC_ERR_HIP, "hipSetDevice failed");
  if (dev >= kScratchDevices || (!g_scratch[dev] && MlpScratchInit(dev)))
    return MmcFail(ctx, MMC_ERR_HIP, "no address for the static scratch");
  unsigned char* base = static_cast<unsigned char*>(g_scratch[dev]);
  float* xin = reinterpret_cast<float*>(base + kScratchX);
  float* yout = reinterpret_cast<float*>(base + kScratchY);
  uint32_t* err = reinterpret_cast<uint32_t*>(base + kScratchErr);
  hipStream_t s = MmcStream(ctx);
  hipLaunchKernelGGL(KClearWord, dim3(1), dim3(1), 0, s, err);
  const uint32_t nb = uint32_t((n + 255) / 256);
  hipLaunchKernelGGL(KSurrogateBatch, dim3(nb), dim3(256), 0, s, MmcWorld(ctx), m, tsl,
                     uint32_t(n), E, T, rng, xin);
  LaunchInfer(m, xin, yout, uint32_t(n), s);
  hipLaunchKernelGGL(KSurrogateFinish, dim3(nb), dim3(256), 0, s, MmcWorld(ctx), m, tsl,
                     uint32_t(n), E, T, rng, yout, alpha, beta, err);
  uint32_t herr = 0;
  if (hipMemcpyAsync(&herr, err, sizeof(uint32_t), hipMemcpyDeviceToHost, s) != hipSuccess)
    return MmcFail(ctx, MMC_ERR_HIP, "error word readback failed");
  if (const int rc = SyncChecked(ctx, "mmc_tsl_sample_surrogate_noalloc")) return rc;
  if (herr) return MmcFail(ctx, MMC_ERR_RESAMPLE, "POD fallback failed");
  return MMC_OK;
}

