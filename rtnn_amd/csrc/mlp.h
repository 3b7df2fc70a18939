// Radius-proposal MLP state (30 -> 64 -> 64 -> 1, GELU, BF16 weights).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtnn {

struct MlpState {
  void* dev = nullptr;  // device parameter block (see mlp.hip)
  size_t bytes = 0;
  bool ready = false;   // weights present (trained or loaded)
};

void mlp_release(MlpState& m);

}  // namespace rtnn
