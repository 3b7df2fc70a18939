// Density-based launch radius per query: the reference's query partitioning
// (src/optixNSearch/grid.cu:33-44,84-211 calcSearchSize; sort.cpp:346-391
// genBatches; util.cpp:408-426) as one kernel per query instead of one thread
// per occupied cell plus a histogram and per-batch BVHs.
//
// For the query's cell c on a count grid of edge `cell`, grow the cube of
// (2i+1)^3 cells shell by shell until it holds >= k+1 points (the count may
// include the query itself) or its width (2i+2)*cell exceeds 2r/sqrt(3) (the
// cube inscribed in the radius-r sphere).  Every point of that cube lies within
// sqrt(3)*(i+1)*cell of the query, so the k nearest others do too: the launch
// radius is min(r, sqrt(3)*(i+1)*cell) (the reference's exact mode,
// radiusFromMegacell with approxMode 0).  Queries that still see fewer than k
// hits are re-run at r by the caller, so the result stays exact.
#include "common.h"
#include "internal.h"

namespace rtnn {

__device__ inline uint32_t cell_count(const GridDesc& g, const uint32_t* __restrict__ counts,
                                      int x, int y, int z) {
  if (x < 0 || y < 0 || z < 0 || x >= g.dim[0] || y >= g.dim[1] || z >= g.dim[2]) return 0;
  return counts[((uint32_t)z * (uint32_t)g.dim[1] + (uint32_t)y) * (uint32_t)g.dim[0] +
                (uint32_t)x];
}

// Point count of the cells at Chebyshev distance exactly i from (cx, cy, cz).
__device__ inline uint64_t shell_count(const GridDesc& g, const uint32_t* __restrict__ counts,
                                       int cx, int cy, int cz, int i) {
  if (i == 0) return cell_count(g, counts, cx, cy, cz);
  uint64_t s = 0;
  for (int dz = -i; dz <= i; ++dz) {
    const int z = cz + dz;
    if (z < 0 || z >= g.dim[2]) continue;
    for (int dy = -i; dy <= i; ++dy) {
      const int y = cy + dy;
      if (y < 0 || y >= g.dim[1]) continue;
      const bool face = (dz == -i || dz == i || dy == -i || dy == i);
      if (face) {
        for (int dx = -i; dx <= i; ++dx) s += cell_count(g, counts, cx + dx, y, z);
      } else {
        s += cell_count(g, counts, cx - i, y, z) + cell_count(g, counts, cx + i, y, z);
      }
    }
  }
  return s;
}

__global__ __launch_bounds__(256) void k_partition_radius(const float4* __restrict__ qproc,
                                                          uint32_t nq,
                                                          const GridDesc* __restrict__ grid,
                                                          const uint32_t* __restrict__ counts,
                                                          uint32_t k, float r,
                                                          float* __restrict__ qradius,
                                                          uint32_t* __restrict__ iter_hist,
                                                          int hist_bins) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  const GridDesc g = *grid;
  const float4 q = qproc[i];
  int c[3];
  const float p[3] = {q.x, q.y, q.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    float f = floorf((p[a] - g.mn[a]) * g.inv);
    f = fminf(fmaxf(f, -1.0f), (float)g.dim[a]);  // outside / NaN: border-adjacent
    c[a] = (int)f;
  }
  const float max_width = 2.0f * r / 1.7320508f;
  uint64_t count = 0;
  int it = 0;
  float R = r;
  for (;;) {
    count += shell_count(g, counts, c[0], c[1], c[2], it);
    const float width = (float)(2 * it + 2) * g.cell;
    if (width > max_width) break;  // megacell too wide: search at r
    if (count >= (uint64_t)k + 1) {
      R = fminf(r, 1.7320508f * (float)(it + 1) * g.cell * (1.0f + 1e-6f));
      break;
    }
    ++it;
  }
  qradius[i] = R;
  if (iter_hist) atomicAdd(&iter_hist[min(it, hist_bins - 1)], 1u);
}

// The same megacell rule on a hierarchical Morton grid (levels l, cell
// c_l = 2^l c_0, point counts from the level tables): at the finest level
// where the query's own cell holds >= k+1 points the radius is sqrt(3) c_l
// (every point of a cell containing the query is within its diagonal); where
// the 3x3x3 block around it does, 2 sqrt(3) c_l.  Queries outside the grid box
// and queries no level satisfies get r.
__global__ __launch_bounds__(256) void k_partition_hier(const float4* __restrict__ qproc,
                                                        uint32_t nq,
                                                        const GridDesc* __restrict__ levels,
                                                        int nlevels, int level_min,
                                                        const uint32_t* __restrict__ table,
                                                        uint32_t k, float r,
                                                        float* __restrict__ qradius) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  const GridDesc g0 = levels[0];
  const float4 q = qproc[i];
  const float p[3] = {q.x, q.y, q.z};
  int c[3];
  bool inside = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float f = floorf((p[a] - g0.mn[a]) * g0.inv);
    inside = inside && f >= 0.0f && f < (float)g0.dim[a];
    c[a] = inside ? (int)f : 0;
  }
  float R = r;
  if (inside) {
    const float kSqrt3 = 1.7320508f * (1.0f + 1e-6f);
    for (int l = level_min; l < nlevels; ++l) {
      const float cell = ldexpf(g0.cell, l);
      if (kSqrt3 * cell >= r) break;
      const GridDesc& gl = levels[l];
      const int x = c[0] >> l, y = c[1] >> l, z = c[2] >> l;
      const uint32_t id = gl.cell_base + morton3((uint32_t)x, (uint32_t)y, (uint32_t)z);
      if (table[id + 1] - table[id] >= k + 1) {
        R = kSqrt3 * cell;
        break;
      }
      if (2.0f * kSqrt3 * cell >= r) continue;  // the block test could not beat r
      uint64_t cnt = 0;
      for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dx = -1; dx <= 1; ++dx) {
            const int xx = x + dx, yy = y + dy, zz = z + dz;
            if (xx < 0 || yy < 0 || zz < 0 || xx >= gl.dim[0] || yy >= gl.dim[1] ||
                zz >= gl.dim[2])
              continue;
            const uint32_t j = gl.cell_base + morton3((uint32_t)xx, (uint32_t)yy, (uint32_t)zz);
            cnt += table[j + 1] - table[j];
          }
      if (cnt >= (uint64_t)k + 1) {
        R = 2.0f * kSqrt3 * cell;
        break;
      }
    }
  }
  qradius[i] = fminf(R, r);
}

// Density buckets of the proposal methods: a query whose proposed radius is
// at least snap_r belongs to the sparse bucket and is searched at r at once,
// so its row is final after the first launch (no search at r' that is bound
// to come up short, and no fallback stages); the others keep r'.  The
// queries stay in one launch in their Morton order: laying the launch out
// bucket by bucket (the reference's batches, sort.cpp:346-391) cost the
// Levy-flight KNN 2.3x in L2 locality.
__global__ __launch_bounds__(256) void k_bucket_snap(float* __restrict__ qradius, uint32_t nq,
                                                     float snap_r, float r) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= nq) return;
  const float R = qradius[i];
  if (R >= snap_r && R != r) qradius[i] = r;
}

hipError_t bucket_snap(float* qradius, uint32_t nq, float snap_r, float r, hipStream_t st) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bucket_snap, dim3((nq + 255) / 256), dim3(256), 0, st, qradius, nq,
                     snap_r, r);
  return hipGetLastError();
}

hipError_t partition_hier(const float4* qproc, uint32_t nq, const GridDesc* levels, int nlevels,
                          int level_min, const uint32_t* table, uint32_t k, float r,
                          float* qradius, hipStream_t st) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_partition_hier, dim3((nq + 255) / 256), dim3(256), 0, st, qproc, nq,
                     levels, nlevels, level_min, table, k, r, qradius);
  return hipGetLastError();
}

hipError_t partition_radius(const float4* qproc, uint32_t nq, const GridDesc* grid,
                            const uint32_t* counts, uint32_t k, float r, float* qradius,
                            uint32_t* iter_hist, int hist_bins, hipStream_t st) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_partition_radius, dim3((nq + 255) / 256), dim3(256), 0, st, qproc, nq,
                     grid, counts, k, r, qradius, iter_hist, hist_bins);
  return hipGetLastError();
}

}  // namespace rtnn
