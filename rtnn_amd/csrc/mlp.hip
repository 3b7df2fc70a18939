// Learned radius proposal for KNN on gfx950.
//
// Inference (k_mlp_propose, k_mlp_forward): MFMA v_mfma_f32_16x16x16_bf16.
// The layers are computed transposed, H^T = W * X^T: the weights are the A
// operand (staged once per block in LDS, in fragment order) and the activations
// of 16 queries are the B operand.  For this instruction lane l holds
//   A[l % 16][4 * (l / 16) + j],  B[4 * (l / 16) + j][l % 16],
//   D[4 * (l / 16) + j][l % 16]      (j = 0..3),
// so a layer's output tile is already laid out as the next layer's B operand:
// GELU and the BF16 rounding happen in registers, no LDS transpose.  The last
// layer (64 -> 1) is a VALU dot product reduced over the four lane groups.
//
// Training (k_mlp_grad, k_mlp_adam): FP32 VALU, one sample per lane, the
// same BF16 rounding of weights and activations in the forward pass as the
// MFMA path (straight-through gradients to the FP32 master weights).
// Per-block gradient partials are summed in a fixed order, so training is
// deterministic.
//
// GELU is Hendrycks & Gimpel's sigmoid form (2016, "Gaussian Error Linear
// Units", section 2), x sigma(1.702 x), with the logistic replaced by the
// hard sigmoid of Howard et al.'s hard-swish (2019, "Searching for
// MobileNetV3", section 5.2: x ReLU6(x + 3) / 6): gelu(x) = x clamp(1.702 x
// / 6 + 1/2, 0, 1), within 0.069 of the exact x Phi(x) (0.021 for the
// sigmoid form).  It keeps gelu(x) - gelu(-x) = x and is exactly x for
// x >= 3 / 1.702 and exactly 0 below -3 / 1.702.  Two full-rate VALU
// operations (a fused multiply-add with the clamp, a multiply) where the
// sigmoid took an exp2 and a reciprocal, both issued at a quarter of the
// VALU rate: the activations had been the largest share of the kernel's
// issue time.  Measured at 10M queries against a smoothstep in place of
// the ramp (0.031 from x Phi(x), six operations): 0.45 vs 0.52 ms per
// proposal pass, and the network trained with it leaves fewer rows short
// (0.72M vs 0.79M).  Training differentiates the same function.  The last
// layer (64 -> 1, VALU) takes the FP32 activations.
#include <math.h>

#include "../../include/rtnn_hip.h"
#include "common.h"
#include "internal.h"
#include "mlp.h"

namespace rtnn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16 with round-to-nearest-even (finite inputs): on the device the
// gfx950 conversion instruction (v_cvt_pk_bf16_f32, round to nearest even).
__host__ __device__ inline uint16_t f2bf(float f) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_bit_cast(uint16_t, (__bf16)f);
#else
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
#endif
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ inline float round_bf(float f) { return bf2f(f2bf(f)); }

constexpr float kGeluA = 1.702f / 6.0f;  // slope of the ramp

__device__ inline float gelu_ramp(float x) { return fmaf(kGeluA, x, 0.5f); }
__device__ inline float gelu(float x) { return x * fminf(fmaxf(gelu_ramp(x), 0.0f), 1.0f); }
// 0 below the ramp, 1 above it, 2 a x + 1/2 on it
__device__ inline float gelu_grad(float x) {
  const float h = gelu_ramp(x);
  return h <= 0.0f ? 0.0f : h >= 1.0f ? 1.0f : fmaf(2.0f * kGeluA, x, 0.5f);
}

// ------------------------------------------------------------ features --

constexpr float kLn2 = 0.6931471805599453f;

// log1p(count) * count_scale as log2(1 + count) * (count_scale ln 2): one
// hardware log2 (v_log_f32) and one multiply per count feature.
__device__ inline float count_feature(uint32_t c, float scale_ln2) {
  return __builtin_amdgcn_logf(1.0f + (float)c) * scale_ln2;
}


// Feature f of a query: f < 27 -> log1p(count of coarse cell c + (dx,dy,dz)),
// (dx,dy,dz) = (f%3-1, f/3%3-1, f/9-1), times count_scale; 27..29 -> offset of
// the query inside its coarse cell; 30, 31 -> 0.
struct QueryCell {
  int cx, cy, cz;
  float ox, oy, oz;
};

// Coarse cell (clamped to one cell outside the grid for far or NaN queries)
// and the offset of the query inside it, clamped to [0, 1].
__device__ inline void axis_cell(float p, float mn, float inv, int dim, int& c, float& off) {
  float t = (p - mn) * inv;
  float f = floorf(t);
  f = fminf(fmaxf(f, -2.0f), (float)dim + 1.0f);
  c = (int)f;
  float o = t - f;
  off = (o >= 0.0f && o <= 1.0f) ? o : (o > 1.0f ? 1.0f : 0.0f);
}

__device__ inline QueryCell query_cell(const GridDesc& g, float x, float y, float z) {
  QueryCell qc;
  axis_cell(x, g.mn[0], g.inv, g.dim[0], qc.cx, qc.ox);
  axis_cell(y, g.mn[1], g.inv, g.dim[1], qc.cy, qc.oy);
  axis_cell(z, g.mn[2], g.inv, g.dim[2], qc.cz, qc.oz);
  return qc;
}

__device__ inline float feature(const GridDesc& g, const uint32_t* __restrict__ counts,
                                const QueryCell& qc, int f, float count_scale) {
  if (f < 27) {
    int x = qc.cx + f % 3 - 1, y = qc.cy + (f / 3) % 3 - 1, z = qc.cz + f / 9 - 1;
    if (x < 0 || y < 0 || z < 0 || x >= g.dim[0] || y >= g.dim[1] || z >= g.dim[2])
      return 0.0f;
    uint32_t c;
    if (g.morton) {  // cell table of a Morton level: count = end - start
      const uint32_t id = g.cell_base + morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
      c = counts[id + 1] - counts[id];
    } else {
      c = counts[((uint32_t)z * (uint32_t)g.dim[1] + (uint32_t)y) * (uint32_t)g.dim[0] +
                 (uint32_t)x];
    }
    return count_feature(c, count_scale * kLn2);
  }
  if (f == 27) return qc.ox;
  if (f == 28) return qc.oy;
  if (f == 29) return qc.oz;
  return 0.0f;
}

// The 8 features 16ks + 4*grp + j (ks < 2, j < 4) of one query, as the two
// BF16 B-operand fragments of layer 1, in two steps so a caller can overlap
// the count loads with other work: feature_loads forms every count address
// and issues all loads (branch-free; an out of range cell loads entry 0 and
// is masked), feature_frag turns the counts into the fragments.  Same values
// as feature().
struct FeatureLoads {
  uint32_t c0[8], c1[8];
  uint32_t valid;  // bit t: feature t is a count inside the grid
};

// What the 8 features of a lane are, fixed by its lane group: bit t of cnt
// says feature t is a cell count (f < 27), off[t] is that cell's raster
// offset from the query's cell, and features 27, 28, 29 (the query's offset
// in its cell; all in the second fragment) are w{x,y,z}[t - 4] = 1.
struct FeatureLane {
  int off[8];
  uint32_t cnt;
  float wx[4], wy[4], wz[4];
};

__device__ inline void feature_lane(const GridDesc& g, int grp, FeatureLane& fn) {
  fn.cnt = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int f = 16 * (t / 4) + 4 * grp + (t % 4);
    const int dx = f % 3 - 1, dy = (f / 3) % 3 - 1, dz = f / 9 - 1;
    fn.off[t] = f < 27 ? (dz * g.dim[1] + dy) * g.dim[0] + dx : 0;
    fn.cnt |= f < 27 ? (1u << t) : 0u;
    if (t >= 4) {
      fn.wx[t - 4] = f == 27 ? 1.0f : 0.0f;
      fn.wy[t - 4] = f == 28 ? 1.0f : 0.0f;
      fn.wz[t - 4] = f == 29 ? 1.0f : 0.0f;
    }
  }
}

template <bool MORTON>
__device__ inline void feature_loads(const GridDesc& g, const uint32_t* __restrict__ counts,
                                     const QueryCell& qc, int grp, const FeatureLane& fn,
                                     FeatureLoads& fl) {
  // Cell index of each of the lane's 8 features (0 and not valid for the
  // offsets and padding), then all loads at once: one assignment per field,
  // so the struct stays in registers (an early return from the fast path
  // left part of it in scratch memory).
  uint32_t id[8];
  uint32_t valid = 0;
  bool fast = false;
  if constexpr (!MORTON) {
    // When every query of the wave has its 27 cells inside the grid (the
    // common case) a feature's cell is the query's cell plus a fixed offset.
    const bool inner = qc.cx >= 1 && qc.cy >= 1 && qc.cz >= 1 && qc.cx + 1 < g.dim[0] &&
                       qc.cy + 1 < g.dim[1] && qc.cz + 1 < g.dim[2];
    fast = __ballot(!inner) == 0;
  }
  if (fast) {
    const int lin = (qc.cz * g.dim[1] + qc.cy) * g.dim[0] + qc.cx;
#pragma unroll
    for (int t = 0; t < 8; ++t) id[t] = (fn.cnt >> t) & 1u ? (uint32_t)(lin + fn.off[t]) : 0u;
    valid = fn.cnt;
  } else {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
      const int f = 16 * (t / 4) + 4 * grp + (t % 4);
      const int cx = qc.cx + f % 3 - 1, cy = qc.cy + (f / 3) % 3 - 1, cz = qc.cz + f / 9 - 1;
      const bool v = f < 27 && cx >= 0 && cy >= 0 && cz >= 0 && cx < g.dim[0] &&
                     cy < g.dim[1] && cz < g.dim[2];
      uint32_t cell;
      if (MORTON)
        cell = g.cell_base + morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
      else
        cell = ((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) * (uint32_t)g.dim[0] +
               (uint32_t)cx;
      id[t] = v ? cell : 0u;
      valid |= v ? (1u << t) : 0u;
    }
  }
  fl.valid = valid;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    fl.c0[t] = counts[id[t]];
    fl.c1[t] = MORTON ? counts[id[t] + 1] : 0u;
  }
}

template <bool MORTON>
__device__ inline void feature_frag(const QueryCell& qc, float scale_ln2, const FeatureLane& fn,
                                    const FeatureLoads& fl, bf16x4 (&x)[2]) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const uint32_t c = MORTON ? fl.c1[t] - fl.c0[t] : fl.c0[t];
    float v = (fl.valid >> t) & 1u ? count_feature(c, scale_ln2) : 0.0f;
    if (t >= 4) v = fmaf(fn.wx[t - 4], qc.ox, fmaf(fn.wy[t - 4], qc.oy, fmaf(fn.wz[t - 4], qc.oz, v)));
    x[t / 4][t % 4] = (short)f2bf(v);
  }
}

__global__ __launch_bounds__(256) void k_mlp_features(const float* __restrict__ q, uint64_t n,
                                                      const GridDesc* __restrict__ coarse,
                                                      const uint32_t* __restrict__ counts,
                                                      float count_scale,
                                                      float* __restrict__ feats) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const GridDesc g = *coarse;
  QueryCell qc = query_cell(g, q[3 * i], q[3 * i + 1], q[3 * i + 2]);
  for (int f = 0; f < kMlpFP; ++f) feats[i * kMlpFP + f] = feature(g, counts, qc, f, count_scale);
}

// ------------------------------------------------------- MFMA inference --

// Weights of the two MFMA layers in LDS, one copy per block, in A-fragment
// order: lane l reads its bf16x4 of tile (s, ks) at [s][ks][l], 8 consecutive
// bytes per lane (conflict-free ds_read_b64).  Biases and the last layer by
// output index (each read is a 4-address broadcast within the wave).  Keeping
// them out of VGPRs leaves room for more waves to hide the feature gathers.
struct MfmaLds {
  bf16x4 a1[4][2][64];
  bf16x4 a2[4][4][64];
  float b1[kMlpH], b2[kMlpH], w3[kMlpH];  // w3 BF16-rounded
  float b3;
};

// Whole block; ends with a barrier.
__device__ inline void stage_weights(const float* __restrict__ P, MfmaLds& w) {
  for (int e = threadIdx.x; e < 4 * 2 * 64; e += blockDim.x) {
    const int s = e / 128, ks = (e / 64) % 2, l = e % 64, row = l & 15, grp = l >> 4;
    bf16x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = 16 * ks + 4 * grp + j;
      v[j] = (short)f2bf(col < kMlpF ? P[kOW1 + (16 * s + row) * kMlpF + col] : 0.0f);
    }
    w.a1[s][ks][l] = v;
  }
  for (int e = threadIdx.x; e < 4 * 4 * 64; e += blockDim.x) {
    const int s = e / 256, ks = (e / 64) % 4, l = e % 64, row = l & 15, grp = l >> 4;
    bf16x4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      v[j] = (short)f2bf(P[kOW2 + (16 * s + row) * kMlpH + 16 * ks + 4 * grp + j]);
    w.a2[s][ks][l] = v;
  }
  for (int e = threadIdx.x; e < kMlpH; e += blockDim.x) {
    w.b1[e] = P[kOB1 + e];
    w.b2[e] = P[kOB2 + e];
    w.w3[e] = round_bf(P[kOW3 + e]);
  }
  if (threadIdx.x == 0) w.b3 = P[kOB3];
  __syncthreads();
}

// Forward of one 16-query tile given the B operands of layer 1 (x[ks] holds
// features 16ks + 4*(l/16) + j of query l%16).  Returns o for query l%16 in
// every lane of the column.
__device__ inline float mlp_tile(const MfmaLds& w, const bf16x4 (&x)[2]) {
  const int l = threadIdx.x & 63, grp = l >> 4;
  // Compiler-only barrier: keeps the weight reads inside the tile loop (hoisted
  // out of it they would occupy ~100 VGPRs and halve the waves per SIMD).
  __asm__ __volatile__("" ::: "memory");
  // The biases start the accumulators: D = W X + b in one MFMA chain.
  f32x4 acc[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[s][j] = w.b1[16 * s + 4 * grp + j];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
      acc[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.a1[s][ks][l], x[ks], acc[s], 0, 0, 0);
  }
  bf16x4 h1[4];
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j) h1[s][j] = (short)f2bf(gelu(acc[s][j]));
#pragma unroll
  for (int s = 0; s < 4; ++s) {
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[s][j] = w.b2[16 * s + 4 * grp + j];
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
      acc[s] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w.a2[s][ks][l], h1[ks], acc[s], 0, 0, 0);
  }
  float part = 0.0f;
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      part = fmaf(w.w3[16 * s + 4 * grp + j], gelu(acc[s][j]), part);
  part += __shfl_xor(part, 16, 64);
  part += __shfl_xor(part, 32, 64);
  return part + w.b3;
}

// Persistent: each wave handles tiles of 16 rows, grid-strided.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ P,
                                                     const float* __restrict__ feats, uint64_t n,
                                                     float* __restrict__ out) {
  __shared__ MfmaLds w;
  stage_weights(P, w);
  const int l = threadIdx.x & 63, row = l & 15, grp = l >> 4;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / 64;
  const uint64_t ntiles = (n + 15) / 16;
  for (uint64_t t = wave; t < ntiles; t += nwaves) {
    const uint64_t qi = t * 16 + row;
    bf16x4 x[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        x[ks][j] = (short)(qi < n ? f2bf(feats[qi * kMlpFP + 16 * ks + 4 * grp + j]) : 0);
    float o = mlp_tile(w, x);
    if (grp == 0 && qi < n) out[qi] = o;
  }
}

template <bool MORTON>
__global__ __launch_bounds__(256) void k_mlp_propose(const float* __restrict__ P,
                                                     const float4* __restrict__ qproc,
                                                     uint32_t nq,
                                                     const GridDesc* __restrict__ coarse,
                                                     const uint32_t* __restrict__ counts,
                                                     float count_scale, float margin, float r,
                                                     int cap, float* __restrict__ qradius) {
  __shared__ MfmaLds w;
  stage_weights(P, w);
  const GridDesc g = *coarse;
  const int l = threadIdx.x & 63, row = l & 15, grp = l >> 4;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const uint32_t nwaves = gridDim.x * blockDim.x / 64;
  const uint32_t ntiles = (nq + 15) / 16;
  // Software pipeline over this wave's tiles: while tile t runs its MFMA
  // layers, the count loads of tile t + 1 and the query of tile t + 2 are in
  // flight (two tiles of count loads ahead measured no faster at 10M
  // queries, 0.623 vs 0.622 ms).  A position past nq reads query 0 (its
  // result is not written).
  auto load_q = [&](uint32_t t) -> float4 {
    const uint64_t qi = (uint64_t)t * 16 + row;
    return qi < nq ? qproc[qi] : qproc[0];
  };
  uint32_t t = wave;
  if (t >= ntiles) return;
  FeatureLane fn;
  feature_lane(g, grp, fn);
  const float scale_ln2 = count_scale * kLn2;
  QueryCell qa = query_cell(g, 0.f, 0.f, 0.f);
  {
    const float4 q = load_q(t);
    qa = query_cell(g, q.x, q.y, q.z);
  }
  FeatureLoads la;
  feature_loads<MORTON>(g, counts, qa, grp, fn, la);
  float4 qb = load_q(t + nwaves);
  for (; t < ntiles; t += nwaves) {
    const QueryCell qn = query_cell(g, qb.x, qb.y, qb.z);
    FeatureLoads ln;
    feature_loads<MORTON>(g, counts, qn, grp, fn, ln);
    const float4 qc = load_q(t + 2 * nwaves);
    const uint32_t qi = t * 16 + row;
    const float bound = cap && grp == 0 && qi < nq ? qradius[qi] : r;
    bf16x4 x[2];
    feature_frag<MORTON>(qa, scale_ln2, fn, la, x);
    const float o = mlp_tile(w, x);
    if (grp == 0 && qi < nq) {
      const float frac = fminf(fmaxf(o * margin, kMinFrac), 1.0f);
      qradius[qi] = fminf(r * frac, bound);
    }
    qa = qn;
    la = ln;
    qb = qc;
  }
}

// ------------------------------------------------------------ training --

__device__ inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

constexpr int kLd = 65;  // padded row stride (conflict-free column access)

// One sample per lane (kGradBlock lanes).  partial[block] receives this
// block's gradient sums (kMlpNP floats) and loss sum.
__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(const float* __restrict__ P,
                                                         const float* __restrict__ feats,
                                                         const float* __restrict__ targets,
                                                         uint64_t n, uint32_t count,
                                                         uint64_t seed, float tau, float norm,
                                                         float* __restrict__ partial) {
  __shared__ float sW1[kMlpH][kMlpF + 1];
  __shared__ float sW2[kMlpH][kLd];
  __shared__ float sB1[kMlpH], sB2[kMlpH], sW3[kMlpH];
  __shared__ float sX[kGradBlock][kMlpF + 1];
  __shared__ float sZ1[kGradBlock][kLd];
  __shared__ float sH1[kGradBlock][kLd];  // bf16(gelu(z1)), the input of layer 2
  __shared__ float sZ2[kGradBlock][kLd];
  __shared__ float sD2[kGradBlock][kLd];
  __shared__ float sD1[kGradBlock][kLd];
  __shared__ float sG[kGradBlock];
  __shared__ float sLoss[kGradBlock];
  const int t = threadIdx.x;

  for (int i = t; i < kMlpH * kMlpF; i += kGradBlock) sW1[i / kMlpF][i % kMlpF] = round_bf(P[kOW1 + i]);
  for (int i = t; i < kMlpH * kMlpH; i += kGradBlock) sW2[i / kMlpH][i % kMlpH] = round_bf(P[kOW2 + i]);
  sB1[t] = P[kOB1 + t];
  sB2[t] = P[kOB2 + t];
  sW3[t] = round_bf(P[kOW3 + t]);
  const float b3 = P[kOB3];
  __syncthreads();

  const uint32_t si = blockIdx.x * kGradBlock + t;
  const bool valid = si < count;
  uint64_t row = si;
  if (seed) row = mix64(seed ^ ((uint64_t)si * 0x100000001b3ull)) % n;
  float x[kMlpF];
#pragma unroll
  for (int f = 0; f < kMlpF; ++f) {
    x[f] = valid ? round_bf(feats[row * kMlpFP + f]) : 0.0f;
    sX[t][f] = x[f];
  }
  // forward
  for (int o = 0; o < kMlpH; ++o) {
    float z = 0.0f;
#pragma unroll
    for (int f = 0; f < kMlpF; ++f) z += sW1[o][f] * x[f];
    z += sB1[o];
    sZ1[t][o] = z;
    sH1[t][o] = round_bf(gelu(z));
  }
  for (int o = 0; o < kMlpH; ++o) {
    float z = 0.0f;
    for (int i = 0; i < kMlpH; ++i) z += sW2[o][i] * sH1[t][i];
    sZ2[t][o] = z + sB2[o];
  }
  float out = 0.0f;
  for (int o = 0; o < kMlpH; ++o) out += sW3[o] * gelu(sZ2[t][o]);
  out += b3;
  const float tgt = valid ? targets[row] : 0.0f;
  const float diff = tgt - out;
  const float loss = valid ? (diff >= 0.0f ? tau * diff : (tau - 1.0f) * diff) : 0.0f;
  const float g = valid ? norm * (out > tgt ? (1.0f - tau) : (out < tgt ? -tau : 0.0f)) : 0.0f;
  sG[t] = g;
  sLoss[t] = loss;
  // backward through layer 3 and 2
  for (int o = 0; o < kMlpH; ++o) sD2[t][o] = g * sW3[o] * gelu_grad(sZ2[t][o]);
  for (int i = 0; i < kMlpH; ++i) {
    float d = 0.0f;
    for (int o = 0; o < kMlpH; ++o) d += sW2[o][i] * sD2[t][o];
    sD1[t][i] = d * gelu_grad(sZ1[t][i]);
  }
  __syncthreads();

  // weight gradients: lane t owns output unit t of every layer
  float* dst = partial + (uint64_t)blockIdx.x * (kMlpNP + 1);
  {
    float gb2 = 0.0f, gw3 = 0.0f, gb1 = 0.0f;
    for (int s = 0; s < kGradBlock; ++s) {
      gb2 += sD2[s][t];
      gb1 += sD1[s][t];
      gw3 += sG[s] * gelu(sZ2[s][t]);
    }
    dst[kOB2 + t] = gb2;
    dst[kOB1 + t] = gb1;
    dst[kOW3 + t] = gw3;
    for (int i = 0; i < kMlpH; ++i) {
      float acc = 0.0f;
      for (int s = 0; s < kGradBlock; ++s) acc += sD2[s][t] * sH1[s][i];
      dst[kOW2 + t * kMlpH + i] = acc;
    }
    for (int f = 0; f < kMlpF; ++f) {
      float acc = 0.0f;
      for (int s = 0; s < kGradBlock; ++s) acc += sD1[s][t] * sX[s][f];
      dst[kOW1 + t * kMlpF + f] = acc;
    }
  }
  if (t == 0) {
    float gb3 = 0.0f, ls = 0.0f;
    for (int s = 0; s < kGradBlock; ++s) {
      gb3 += sG[s];
      ls += sLoss[s];
    }
    dst[kOB3] = gb3;
    dst[kMlpNP] = ls * norm;
  }
}

// grads[p] = sum over blocks of partial[b][p] (fixed order); loss likewise.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial,
                                                    uint32_t nblocks, float* __restrict__ grads,
                                                    float* __restrict__ loss) {
  uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p > (uint32_t)kMlpNP) return;
  float s = 0.0f;
  for (uint32_t b = 0; b < nblocks; ++b) s += partial[(uint64_t)b * (kMlpNP + 1) + p];
  if (p < (uint32_t)kMlpNP) grads[p] = s;
  else if (loss) *loss = s;
}

__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ P, float* __restrict__ m,
                                                  float* __restrict__ v,
                                                  const float* __restrict__ partial,
                                                  uint32_t nblocks, float lr, float bc1,
                                                  float bc2) {
  uint32_t p = blockIdx.x * 256 + threadIdx.x;
  if (p >= (uint32_t)kMlpNP) return;
  float g = 0.0f;
  for (uint32_t b = 0; b < nblocks; ++b) g += partial[(uint64_t)b * (kMlpNP + 1) + p];
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float mm = b1 * m[p] + (1.0f - b1) * g;
  float vv = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  P[p] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

__global__ __launch_bounds__(256) void k_mlp_targets(const float* __restrict__ d2, uint32_t k,
                                                     uint64_t n, float r,
                                                     float* __restrict__ t) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float d = d2[i * k + (k - 1)];
  t[i] = (d < __builtin_huge_valf()) ? sqrtf(d) / r : 1.0f;
}

__global__ __launch_bounds__(256) void k_sample_points(const float* __restrict__ pts,
                                                       uint64_t npts, uint64_t n, uint64_t seed,
                                                       float* __restrict__ out) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint64_t j = mix64(seed * 0x9e3779b97f4a7c15ull + i) % npts;
  out[3 * i] = pts[3 * j];
  out[3 * i + 1] = pts[3 * j + 1];
  out[3 * i + 2] = pts[3 * j + 2];
}

// ------------------------------------------------------------- host --

hipError_t mlp_alloc(MlpState& m, uint32_t max_grad_blocks) {
  hipError_t e;
  if (!m.params) {
    if ((e = dev_alloc((void**)&m.params, 4 * kMlpNP)) != hipSuccess) return e;
    if ((e = dev_alloc((void**)&m.adam_m, 4 * kMlpNP)) != hipSuccess) return e;
    if ((e = dev_alloc((void**)&m.adam_v, 4 * kMlpNP)) != hipSuccess) return e;
  }
  if (m.partial_blocks < max_grad_blocks) {
    if (m.partial) {
      (void)hipDeviceSynchronize();
      dev_free(m.partial, 4ull * (kMlpNP + 1) * m.partial_blocks);
      m.partial = nullptr;
    }
    if ((e = dev_alloc((void**)&m.partial, 4ull * (kMlpNP + 1) * max_grad_blocks)) !=
        hipSuccess) {
      m.partial_blocks = 0;
      return e;
    }
    m.partial_blocks = max_grad_blocks;
  }
  return hipSuccess;
}

void mlp_release(MlpState& m) {
  float* ptrs[] = {m.params, m.adam_m, m.adam_v};
  for (float* p : ptrs)
    if (p) dev_free(p, 4 * kMlpNP);
  if (m.partial) dev_free(m.partial, 4ull * (kMlpNP + 1) * m.partial_blocks);
  m = MlpState{};
}

void mlp_init_host(float* P, uint64_t seed) {
  uint64_t s = seed;
  auto u01 = [&s]() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
  };
  auto fill = [&](int off, int fan_out, int fan_in) {
    double lim = std::sqrt(6.0 / (fan_in + fan_out));
    for (int i = 0; i < fan_out * fan_in; ++i) P[off + i] = (float)((2 * u01() - 1) * lim);
  };
  for (int i = 0; i < kMlpNP; ++i) P[i] = 0.0f;
  fill(kOW1, kMlpH, kMlpF);
  fill(kOW2, kMlpH, kMlpH);
  fill(kOW3, 1, kMlpH);
  P[kOB3] = 0.3f;
}

static uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + 255) / 256); }

hipError_t mlp_features(const float* qxyz, uint64_t n, const GridDesc* coarse,
                        const uint32_t* counts, float count_scale, float* feats,
                        hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_features, dim3(blocks_for(n)), dim3(256), 0, st, qxyz, n, coarse,
                     counts, count_scale, feats);
  return hipGetLastError();
}

static dim3 persistent_grid(uint64_t tiles16, int num_cus) {
  uint64_t want = (tiles16 + 3) / 4;  // 4 waves per block
  // two blocks per resident slot: one resident round (4 blocks per CU)
  // measured 5 % slower at 10M queries
  uint64_t cap = (uint64_t)(num_cus > 0 ? num_cus : 256) * 8;
  return dim3((uint32_t)(want < cap ? (want ? want : 1) : cap));
}

hipError_t mlp_forward(const MlpState& m, const float* feats, uint64_t n, float* out,
                       hipStream_t st, int num_cus) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, persistent_grid((n + 15) / 16, num_cus), dim3(256), 0, st,
                     m.params, feats, n, out);
  return hipGetLastError();
}

hipError_t mlp_propose(const MlpState& m, const float4* qproc, uint32_t nq,
                       const GridDesc* coarse, bool morton, const uint32_t* counts, float r,
                       bool cap, float* qradius, hipStream_t st, int num_cus) {
  if (nq == 0) return hipSuccess;
  const dim3 grid = persistent_grid((nq + 15) / 16, num_cus);
  if (morton)
    hipLaunchKernelGGL(k_mlp_propose<true>, grid, dim3(256), 0, st, m.params, qproc, nq, coarse,
                       counts, m.count_scale, m.margin, r, cap ? 1 : 0, qradius);
  else
    hipLaunchKernelGGL(k_mlp_propose<false>, grid, dim3(256), 0, st, m.params, qproc, nq, coarse,
                       counts, m.count_scale, m.margin, r, cap ? 1 : 0, qradius);
  return hipGetLastError();
}

hipError_t mlp_grad(const MlpState& m, const float* feats, const float* targets, uint64_t n,
                    uint32_t count, uint64_t rand_seed, float tau, float norm, float* partial,
                    hipStream_t st) {
  uint32_t nb = (count + kGradBlock - 1) / kGradBlock;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(kGradBlock), 0, st, m.params, feats, targets, n,
                     count, rand_seed, tau, norm, partial);
  return hipGetLastError();
}

hipError_t mlp_reduce(const float* partial, uint32_t nblocks, float* grads, float* loss,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks_for(kMlpNP + 1)), dim3(256), 0, st, partial,
                     nblocks, grads, loss);
  return hipGetLastError();
}

hipError_t mlp_adam(MlpState& m, uint32_t nblocks, float lr, int t, hipStream_t st) {
  float bc1 = 1.0f - std::pow(0.9f, (float)t);
  float bc2 = 1.0f - std::pow(0.999f, (float)t);
  hipLaunchKernelGGL(k_mlp_adam, dim3(blocks_for(kMlpNP)), dim3(256), 0, st, m.params, m.adam_m,
                     m.adam_v, m.partial, nblocks, lr, bc1, bc2);
  return hipGetLastError();
}

hipError_t mlp_targets(const float* d2, uint32_t k, uint64_t n, float r, float* t,
                       hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_targets, dim3(blocks_for(n)), dim3(256), 0, st, d2, k, n, r, t);
  return hipGetLastError();
}

hipError_t mlp_sample_points(const float* pts, uint64_t npts, uint64_t n, uint64_t seed,
                             float* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_sample_points, dim3(blocks_for(n)), dim3(256), 0, st, pts, npts, n, seed,
                     out);
  return hipGetLastError();
}

}  // namespace rtnn
