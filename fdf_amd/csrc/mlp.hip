// mlp.hip -- placeholder, replaced by the MFMA implementation.
#include "picgpu_internal.hpp"

namespace picgpu {
int mlp_guess_device(picgpu_ctx *c, int64_t, const double *, const int32_t *, int32_t *) {
  (void)c;
  set_error("MLP not available");
  return PICGPU_E_UNSUPPORTED;
}
}  // namespace picgpu

extern "C" {
int64_t picgpu_mlp_param_count(void) { return 64 * 3 + 64 + 64 * 64 + 64 + 3 * 64 + 3; }
int picgpu_mlp_set_params(picgpu_ctx *, const float *) { return PICGPU_E_UNSUPPORTED; }
int picgpu_mlp_get_params(picgpu_ctx *, float *) { return PICGPU_E_UNSUPPORTED; }
int picgpu_mlp_train(picgpu_ctx *, int64_t, int, int, float, uint64_t, float *) { return PICGPU_E_UNSUPPORTED; }
int picgpu_mlp_forward(picgpu_ctx *, int64_t, const float *, float *) { return PICGPU_E_UNSUPPORTED; }
int picgpu_mlp_loss_grad(picgpu_ctx *, int64_t, const float *, const float *, float *, float *) {
  return PICGPU_E_UNSUPPORTED;
}
int picgpu_mlp_guess(picgpu_ctx *, int64_t, const double *, int64_t *) { return PICGPU_E_UNSUPPORTED; }
}
