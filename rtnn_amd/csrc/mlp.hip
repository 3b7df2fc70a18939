// Learned radius proposal (placeholder until the MLP kernels land).
#include "../../include/rtnn_hip.h"
#include "mlp.h"

namespace rtnn {
void mlp_release(MlpState& m) {
  if (m.dev) (void)hipFree(m.dev);
  m.dev = nullptr;
  m.bytes = 0;
  m.ready = false;
}
}  // namespace rtnn

extern "C" {
int rtnn_mlp_train(rtnn_ctx, float, uint32_t, uint32_t, uint32_t, uint32_t, float, float,
                   uint64_t, float*) { return RTNN_ERR_UNSUPPORTED; }
int rtnn_mlp_get_weights(rtnn_ctx, float*) { return RTNN_ERR_UNSUPPORTED; }
int rtnn_mlp_set_weights(rtnn_ctx, const float*) { return RTNN_ERR_UNSUPPORTED; }
int rtnn_search_learned(rtnn_ctx, const float*, uint64_t, int, float, uint32_t, int32_t*,
                        float*, int, rtnn_stats*) { return RTNN_ERR_UNSUPPORTED; }
int rtnn_mlp_forward(rtnn_ctx, const float*, uint64_t, float*) { return RTNN_ERR_UNSUPPORTED; }
int rtnn_mlp_backward(rtnn_ctx, const float*, const float*, uint64_t, float, float*, float*) {
  return RTNN_ERR_UNSUPPORTED;
}
int rtnn_mlp_features(rtnn_ctx, const float*, uint64_t, float, float*) {
  return RTNN_ERR_UNSUPPORTED;
}
}
