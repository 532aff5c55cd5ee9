// Thermal-scattering surrogate: state shared between the tracker and the
// MLP kernels (mmc_mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "mmc_kernels.hpp"

namespace mmc {

constexpr int kMlpMaxLayers = 8;

struct MlpState {
  bool ready = false;
  int tsl = 0;
  int n_layers = 0;
  int widths[kMlpMaxLayers + 1] = {};
  int w_off[kMlpMaxLayers] = {}, b_off[kMlpMaxLayers] = {};
  long n_params = 0;
  float in_shift[4] = {}, in_scale[4] = {}, out_shift[2] = {}, out_scale[2] = {};
  double e_min = 0, e_max = 0, t_lo = 0, t_hi = 0;
  float* params = nullptr;        // FP32 master parameters
  void* wbf16 = nullptr;          // unused
  void* packed = nullptr;         // inference layouts (BF16 fragments + FP32)
  float b3[2] = {};
  float *adam_m = nullptr, *adam_v = nullptr, *grads = nullptr;
  float* act = nullptr;           // training activations
  size_t act_cap = 0;
  float* loss = nullptr;
  long adam_t = 0;
  float* xin = nullptr;           // batch sampler: network inputs [batch_cap x 4]
  float* yout = nullptr;          // and outputs [batch_cap x 2]
  uint32_t* berr = nullptr;       // and its error word
  uint64_t batch_cap = 0;
  uint32_t* fb = nullptr;         // parked fallback particles [cap]
  uint32_t stage_cap = 0;         // rows of the tracker buffers
};

void MlpFree(MlpState& m);
}  // namespace mmc
struct mmc_ctx;
// Staging buffers of the surrogate step for `cap` particle slots.
int MlpEnsureStaging(mmc_ctx* ctx, uint32_t cap);
// The device address of the no-allocation sampler's static scratch on
// `device` (current), looked up once per device; 0 on success.
int MlpScratchInit(int device);
namespace mmc {
// One thermal-scattering step of the tracker with the surrogate: one network
// kernel that gathers (log E, T, u1, u2) of the TSL queue and applies the
// scatter (the POD sampler serves the rows the network cannot).
// `mark(k)` is called before each launch with the profiling kind of the
// next kernel (MMC_K_TSL or MMC_K_MLP) and with -1 after the last one.
int MlpTslStep(MlpState& m, const DevWorld& W, const RunState& R, int iter, int qin,
               uint32_t grid, hipStream_t stream, void (*mark)(void*, int), void* mark_arg);
// The POD sampler for the particles the surrogate parked; they rejoin the
// live queue of iteration `iter` (Q[qin]).  Called before every read of the
// live count, so that a parked particle is never missed.  hot_lds: bytes of
// the sampler's tables to stage in LDS (0: read them from HBM).
int MlpFlushFallback(MlpState& m, const DevWorld& W, const RunState& R, int iter, int qin,
                     size_t hot_lds, hipStream_t stream);

}  // namespace mmc
