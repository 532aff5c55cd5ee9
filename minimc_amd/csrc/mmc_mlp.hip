// Thermal-scattering surrogate MLP (placeholder until the kernels land).
#include <hip/hip_runtime.h>

#include "../../include/minimc_amd.h"
#include "mmc_mlp.hpp"

namespace mmc {
void MlpFree(MlpState& m) { m.ready = false; }
int MlpTslStep(MlpState&, const DevWorld&, const RunState&, int, int, uint32_t, hipStream_t) {
  return 1;
}
}  // namespace mmc

extern "C" {
int mmc_mlp_set(mmc_ctx*, int, const mmc_mlp_desc*, const float*, uint64_t) {
  return MMC_ERR_UNSUPPORTED;
}
int mmc_mlp_get(mmc_ctx*, float*, uint64_t) { return MMC_ERR_UNSUPPORTED; }
int mmc_mlp_infer(mmc_ctx*, uint64_t, const float*, float*) { return MMC_ERR_UNSUPPORTED; }
int mmc_tsl_sample_surrogate(mmc_ctx*, int, uint64_t, const double*, const double*,
                             uint32_t*, double*, double*) {
  return MMC_ERR_UNSUPPORTED;
}
int mmc_mlp_train(mmc_ctx*, uint64_t, const float*, const float*, const mmc_train_opts*,
                  float*) {
  return MMC_ERR_UNSUPPORTED;
}
int mmc_mlp_grad(mmc_ctx*, uint64_t, const float*, const float*, float*, float*) {
  return MMC_ERR_UNSUPPORTED;
}
}
