// Radius-proposal MLP: 30 -> 64 -> 64 -> 1, GELU (tanh form), BF16 weights
// and activations into FP32 accumulation.  Inputs per query: log1p point
// counts of the 27 cells (edge r) around the query's cell, scaled, and the
// query's offset inside its cell.  Output o ~ (k-th neighbour distance) / r;
// the proposed radius is r * clamp(o * margin, kMinFrac, 1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"

namespace rtnn {

constexpr int kMlpH = 64;     // hidden width
constexpr int kMlpF = 30;     // features
constexpr int kMlpFP = 32;    // features padded to the MFMA K step
constexpr int kOW1 = 0;       // [64][30]
constexpr int kOB1 = 1920;
constexpr int kOW2 = 1984;    // [64][64]
constexpr int kOB2 = 6080;
constexpr int kOW3 = 6144;
constexpr int kOB3 = 6208;
constexpr int kMlpNP = 6209;  // parameters
constexpr int kMlpNBlob = 6211;
constexpr float kMinFrac = 0.02f;
constexpr int kGradBlock = 64;  // samples per gradient block

struct MlpState {
  float* params = nullptr;   // FP32 master weights (kMlpNP)
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  float* partial = nullptr;  // per-block gradient partials, (kMlpNP + 1) each
  uint32_t partial_blocks = 0;
  float count_scale = 0.25f;
  float margin = 1.0f;
  bool ready = false;
};

hipError_t mlp_alloc(MlpState& m, uint32_t max_grad_blocks);
void mlp_release(MlpState& m);
// Seeded Glorot-uniform weights, zero biases, b3 = 0.3.
void mlp_init_host(float* params, uint64_t seed);

// feats[i] (kMlpFP floats) of query i (xyz AoS) on the coarse count grid.
hipError_t mlp_features(const float* qxyz, uint64_t n, const GridDesc* coarse,
                        const uint32_t* counts, float count_scale, float* feats,
                        hipStream_t st);
// MFMA forward of feature rows -> o (raw network output).
hipError_t mlp_forward(const MlpState& m, const float* feats, uint64_t n, float* out,
                       hipStream_t st, int num_cus);
// Fused features + MFMA forward for queries in processing order; writes the
// proposed radius per processing position.  morton: *coarse is a level of a
// hierarchical Morton grid (counts = its table: end - start), else a raster
// count grid.  cap: each proposal is at most the radius qradius holds on entry.
hipError_t mlp_propose(const MlpState& m, const float4* qproc, uint32_t nq,
                       const GridDesc* coarse, bool morton, const uint32_t* counts, float r,
                       bool cap, float* qradius, hipStream_t st, int num_cus);
// Pinball-loss gradient partials over `count` samples (rows sel[i] of
// feats/targets, or rows 0..count-1 if sel == nullptr, or seeded random rows
// of [0, n) if rand_seed != 0).  norm = 1 / batch size.
hipError_t mlp_grad(const MlpState& m, const float* feats, const float* targets, uint64_t n,
                    uint32_t count, uint64_t rand_seed, float tau, float norm, float* partial,
                    hipStream_t st);
hipError_t mlp_reduce(const float* partial, uint32_t nblocks, float* grads, float* loss,
                      hipStream_t st);
hipError_t mlp_adam(MlpState& m, uint32_t nblocks, float lr, int t, hipStream_t st);
// t_i = sqrt(d2[i*k + k-1]) / r, or 1 when the k-th slot is empty.
hipError_t mlp_targets(const float* d2, uint32_t k, uint64_t n, float r, float* t,
                       hipStream_t st);
// Training queries: seeded uniform draws (with replacement) of the points.
hipError_t mlp_sample_points(const float* pts, uint64_t npts, uint64_t n, uint64_t seed,
                             float* out, hipStream_t st);

}  // namespace rtnn
