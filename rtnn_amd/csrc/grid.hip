// Uniform-grid cell list for gfx950.
//
// Replaces the reference's grid build (src/optixNSearch/sort.cpp:34-136,
// 500-627; grid.cu:13-31,213-329): integer bbox via atomics, meta-grid Morton
// cell ids, atomicAdd local indices (non-deterministic within a cell),
// Thrust scan and sort.  Here:
//   bbox        float min/max, block reduction (no atomics on 6 addresses);
//   cell_keys   raster cell id per point (x fastest), or the level-0 Morton
//               code of a hierarchical grid (morton_keys);
//   radix sort  stable sort of (cell, index) -> points ordered by (cell, index);
//   cell_bounds cell_start[c] = first sorted position with key >= c: one
//               binary search per 4096-cell tile, then each tile's cells
//               searched inside its own key range staged in LDS, so the work
//               per thread is bounded however many cells are empty; a Morton
//               level l uses key >> 3l and the same sorted keys;
//   gather      sorted points as float4 (x, y, z, bits(original index)) so one
//               16-byte load yields a candidate and its output index.
// Query order: Morton code of the query's cell (helper_mortonCode.h:19-34
// restated), radix-sorted, so a wave's queries walk neighbouring cells.
#include <float.h>

#include "common.h"
#include "internal.h"

namespace rtnn {

constexpr int kBboxThreads = 256;

__device__ inline void wave_minmax(float& mn, float& mx) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, o, kWave));
    mx = fmaxf(mx, __shfl_xor(mx, o, kWave));
  }
}

__device__ inline void block_bbox(float v[6], float* sh /*6 * waves*/) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  constexpr int nw = kBboxThreads / kWave;
#pragma unroll
  for (int a = 0; a < 3; ++a) wave_minmax(v[a], v[a + 3]);
  if (lane == 0)
#pragma unroll
    for (int a = 0; a < 6; ++a) sh[a * nw + w] = v[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 3; ++a) {
      float mn = sh[a * nw], mx = sh[(a + 3) * nw];
      for (int i = 1; i < nw; ++i) {
        mn = fminf(mn, sh[a * nw + i]);
        mx = fmaxf(mx, sh[(a + 3) * nw + i]);
      }
      v[a] = mn;
      v[a + 3] = mx;
    }
  }
}

__device__ inline void accum_point(const float* xyz, uint64_t i, float v[6], uint32_t& bad) {
  float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
  if (!(isfinite(x) && isfinite(y) && isfinite(z))) {
    ++bad;
    return;
  }
  v[0] = fminf(v[0], x); v[1] = fminf(v[1], y); v[2] = fminf(v[2], z);
  v[3] = fmaxf(v[3], x); v[4] = fmaxf(v[4], y); v[5] = fmaxf(v[5], z);
}

__global__ __launch_bounds__(kBboxThreads) void k_bbox(const float* __restrict__ xyz,
                                                       uint64_t n, float* __restrict__ part,
                                                       uint32_t* __restrict__ nonfinite) {
  __shared__ float sh[6 * kBboxThreads / kWave];
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  uint32_t bad = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBboxThreads + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * kBboxThreads)
    accum_point(xyz, i, v, bad);
  if (bad) atomicAdd(nonfinite, bad);
  block_bbox(v, sh);
  if (threadIdx.x == 0)
    for (int a = 0; a < 6; ++a) part[6 * blockIdx.x + a] = v[a];
}

__global__ __launch_bounds__(kBboxThreads) void k_bbox_final(const float* __restrict__ part,
                                                             int np, float* __restrict__ out) {
  __shared__ float sh[6 * kBboxThreads / kWave];
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = threadIdx.x; i < np; i += kBboxThreads)
    for (int a = 0; a < 3; ++a) {
      v[a] = fminf(v[a], part[6 * i + a]);
      v[a + 3] = fmaxf(v[a + 3], part[6 * i + a + 3]);
    }
  block_bbox(v, sh);
  if (threadIdx.x == 0)
    for (int a = 0; a < 6; ++a) out[a] = v[a];
}

hipError_t bbox(const float* xyz, uint64_t n, float* part, float* out, uint32_t* nonfinite,
                hipStream_t st) {
  uint64_t want = (n + kBboxThreads * 8 - 1) / (kBboxThreads * 8);
  int nb = (int)(want < (uint64_t)kBboxBlocks ? (want ? want : 1) : kBboxBlocks);
  hipLaunchKernelGGL(k_bbox, dim3(nb), dim3(kBboxThreads), 0, st, xyz, n, part, nonfinite);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBboxThreads), 0, st, part, nb, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(kBboxThreads) void k_bbox_seg(const float* __restrict__ xyz,
                                                           const uint64_t* __restrict__ off,
                                                           float* __restrict__ out,
                                                           uint32_t* __restrict__ nonfinite) {
  __shared__ float sh[6 * kBboxThreads / kWave];
  const uint32_t c = blockIdx.x;
  float v[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
  uint32_t bad = 0;
  for (uint64_t i = off[c] + threadIdx.x; i < off[c + 1]; i += kBboxThreads)
    accum_point(xyz, i, v, bad);
  if (bad) atomicAdd(nonfinite, bad);
  block_bbox(v, sh);
  if (threadIdx.x == 0)
    for (int a = 0; a < 6; ++a) out[6 * c + a] = v[a];
}

hipError_t bbox_segmented(const float* xyz, const uint64_t* offsets, uint32_t nclouds,
                          float* out, uint32_t* nonfinite, hipStream_t st) {
  if (nclouds == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bbox_seg, dim3(nclouds), dim3(kBboxThreads), 0, st, xyz, offsets, out,
                     nonfinite);
  return hipGetLastError();
}

// ---- device-side planning of the batched grids (no host round trip, so the
// whole batched pipeline can be captured in a graph).

__device__ inline uint64_t cells_sat(int d0, int d1, int d2) {
  const double c = (double)d0 * (double)d1 * (double)d2;
  return c > 1e15 ? (uint64_t)1e15 : (uint64_t)d0 * (uint64_t)d1 * (uint64_t)d2;
}

// Dims of one cloud's grid; the host planner's arithmetic (-ffp-contract=off).
__device__ inline void plan_dims(const float* bb, float inv, bool empty, int d[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float e = (bb[a + 3] - bb[a]) * inv;
    const float f = floorf(e) + 1.0f;
    d[a] = empty ? 1 : (int)fminf(fmaxf(f, 1.0f), 2.0e9f);  // NaN, -inf -> 1
  }
}

constexpr int kPlanThreads = 1024;

__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t* sh) {
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  uint64_t t = 0;
  for (int i = 0; i < kPlanThreads / kWave; ++i) t += sh[i];
  return t;
}

// One block.  Grows the edge by 1.25x until the total cell count fits cap
// (the host planner's loop), then writes every cloud's GridDesc with its
// cell_base.  If no edge fits within 200 steps, every cloud gets a single
// cell (inv = 0): slow but exact, and status[1] is set.
__global__ __launch_bounds__(kPlanThreads) void k_plan_batched(
    const float* __restrict__ bboxes, const uint64_t* __restrict__ off, uint32_t nclouds,
    float cell, const float* __restrict__ cells, uint64_t cap, GridDesc* __restrict__ grids,
    uint32_t* __restrict__ status) {
  __shared__ uint64_t sh[kPlanThreads / kWave];
  __shared__ uint64_t s_scan[kPlanThreads];
  // Cloud i's edge: cells[i] (size classes) or cell, times the common growth.
  float gf = 1.0f;
  bool ok = false;
  for (int tries = 0; tries <= 200 && !ok; ++tries) {
    uint64_t tot = 0;
    for (uint32_t i = threadIdx.x; i < nclouds; i += kPlanThreads) {
      int d[3];
      plan_dims(bboxes + 6 * i, 1.0f / ((cells ? cells[i] : cell) * gf), off[i + 1] == off[i], d);
      tot = min(tot + cells_sat(d[0], d[1], d[2]), (uint64_t)1e15);  // no wrap
    }
    tot = block_sum_u64(tot, sh);
    ok = tot <= cap && tot < 0xfffffff0ull;
    if (!ok) gf *= 1.25f;
  }
  if (!ok && threadIdx.x == 0) status[1] = 1u;
  uint64_t carry = 0;
  for (uint32_t i0 = 0; i0 < nclouds; i0 += kPlanThreads) {
    const uint32_t i = i0 + threadIdx.x;
    int d[3] = {1, 1, 1};
    bool empty = true;
    float c = FLT_MAX, inv = 0.0f;
    if (i < nclouds) {
      empty = off[i + 1] == off[i];
      if (ok) {
        c = (cells ? cells[i] : cell) * gf;
        inv = 1.0f / c;
        plan_dims(bboxes + 6 * i, inv, empty, d);
      }
    }
    const uint64_t mine = i < nclouds ? cells_sat(d[0], d[1], d[2]) : 0;
    // inclusive scan of this chunk (Hillis-Steele in LDS)
    __syncthreads();
    s_scan[threadIdx.x] = mine;
    __syncthreads();
    for (int o = 1; o < kPlanThreads; o <<= 1) {
      const uint64_t t = threadIdx.x >= (uint32_t)o ? s_scan[threadIdx.x - o] : 0;
      __syncthreads();
      s_scan[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < nclouds) {
      GridDesc g;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        g.mn[a] = empty ? 0.0f : bboxes[6 * i + a];
        g.dim[a] = d[a];
      }
      g.inv = inv;
      g.cell = c;
      g.cell_base = (uint32_t)(carry + s_scan[threadIdx.x] - mine);
      g.point_base = (uint32_t)off[i];
      g.npoints = (uint32_t)(off[i + 1] - off[i]);
      g.morton = 0;
      grids[i] = g;
    }
    carry += s_scan[kPlanThreads - 1];
  }
  if (threadIdx.x == 0) status[2] = (uint32_t)carry;
}

hipError_t plan_batched(const float* bboxes, const uint64_t* offsets, uint32_t nclouds,
                        float cell, const float* cells, uint64_t cap, GridDesc* grids,
                        uint32_t* status, hipStream_t st) {
  if (nclouds == 0) return hipSuccess;
  hipLaunchKernelGGL(k_plan_batched, dim3(1), dim3(kPlanThreads), 0, st, bboxes, offsets,
                     nclouds, cell, cells, cap, grids, status);
  return hipGetLastError();
}

__device__ inline int cell_of(float x, float mn, float inv, int dim) {
  float c = floorf((x - mn) * inv);
  c = fminf(fmaxf(c, 0.0f), (float)(dim - 1));
  return (int)c;
}

__device__ inline uint32_t find_cloud(const uint64_t* off, uint32_t nclouds, uint64_t i) {
  // last c with off[c] <= i
  uint32_t lo = 0, hi = nclouds - 1;
  while (lo < hi) {
    uint32_t m = (lo + hi + 1) >> 1;
    if (off[m] <= i) lo = m; else hi = m - 1;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_cell_keys(const float* __restrict__ xyz, uint64_t n,
                                                   const GridDesc* __restrict__ grids,
                                                   const uint64_t* __restrict__ off,
                                                   uint32_t nclouds,
                                                   uint32_t* __restrict__ keys,
                                                   uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ counts) {
  // The block's first cloud by one binary search; a thread steps forward from
  // it (clouds are contiguous, so a block of 256 points spans few of them).
  __shared__ uint32_t s_c0;
  const uint64_t i0 = (uint64_t)blockIdx.x * 256;
  const uint64_t i = i0 + threadIdx.x;
  if (threadIdx.x == 0) s_c0 = nclouds > 1 ? find_cloud(off, nclouds, i0) : 0u;
  __syncthreads();
  if (i >= n) return;
  uint32_t c = s_c0;
  while (c + 1 < nclouds && off[c + 1] <= i) ++c;
  const GridDesc g = grids[c];
  int cx = cell_of(xyz[3 * i], g.mn[0], g.inv, g.dim[0]);
  int cy = cell_of(xyz[3 * i + 1], g.mn[1], g.inv, g.dim[1]);
  int cz = cell_of(xyz[3 * i + 2], g.mn[2], g.inv, g.dim[2]);
  uint32_t key = g.cell_base + ((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) *
                                   (uint32_t)g.dim[0] + (uint32_t)cx;
  keys[i] = key;
  vals[i] = (uint32_t)i;
  if (counts) atomicAdd(&counts[key], 1u);
}

hipError_t cell_keys(const float* xyz, uint64_t n, const GridDesc* grids,
                     const uint64_t* offsets, uint32_t nclouds, uint32_t* keys,
                     uint32_t* vals, uint32_t* counts, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_cell_keys, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, xyz, n,
                     grids, offsets, nclouds, keys, vals, counts);
  return hipGetLastError();
}

// One cloud: the grid descriptor comes by value (no host-to-device copy before
// the first kernel) and thread 0 stores it to gstore for the kernels after it.
__global__ __launch_bounds__(256) void k_cell_keys1(const float* __restrict__ xyz, uint64_t n,
                                                    const GridDesc g, GridDesc* __restrict__ gstore,
                                                    uint32_t* __restrict__ keys,
                                                    uint32_t* __restrict__ vals) {
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i == 0) *gstore = g;
  if (i >= n) return;
  const int cx = cell_of(xyz[3 * i], g.mn[0], g.inv, g.dim[0]);
  const int cy = cell_of(xyz[3 * i + 1], g.mn[1], g.inv, g.dim[1]);
  const int cz = cell_of(xyz[3 * i + 2], g.mn[2], g.inv, g.dim[2]);
  keys[i] = g.cell_base + ((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) * (uint32_t)g.dim[0] +
            (uint32_t)cx;
  vals[i] = (uint32_t)i;
}

hipError_t cell_keys_single(const float* xyz, uint64_t n, const GridDesc& g, GridDesc* gstore,
                            uint32_t* keys, uint32_t* vals, hipStream_t st) {
  const uint32_t blocks = (uint32_t)std::max<uint64_t>(1, (n + 255) / 256);
  hipLaunchKernelGGL(k_cell_keys1, dim3(blocks), dim3(256), 0, st, xyz, n, g, gstore, keys, vals);
  return hipGetLastError();
}

// cell_start[c] = first sorted position whose key (>> shift) is >= c, for c
// in [0, ncells] (entry ncells holds n), so a cell's points are
// [cell_start[c], cell_start[c+1]) and no counts, atomics or scan are needed.
// The cells are cut into tiles: k_tile_bounds finds each tile's first sorted
// position by binary search (one thread per tile), then one block per tile
// writes its cells, each a lower_bound over the tile's keys staged in LDS.
// Every block writes exactly one tile, however the points are spread (empty
// space costs the same as occupied space).  Tiles hold 4096 cells, fewer for
// small tables (down to 256) so that at least ~1024 blocks share the work:
// a 125k-cell table in 31 blocks of 4096 left most CUs idle.
constexpr uint32_t kCellTileMax = 4096;
constexpr uint32_t kCellTileMin = 256;
constexpr uint32_t kCellTileKeys = 4096;  // keys staged in LDS per tile

uint32_t cell_tile(uint64_t ncells1) {
  uint32_t t = kCellTileMax;
  while (t > kCellTileMin && ncells1 / t < 1024) t >>= 1;
  return t;
}

__device__ inline uint64_t lower_bound_shifted(const uint32_t* __restrict__ keys, uint64_t lo,
                                               uint64_t hi, uint64_t c, int shift) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if ((uint64_t)(keys[mid] >> shift) < c) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_tile_bounds(const uint32_t* __restrict__ keys, uint64_t n,
                                                     uint64_t ncells1, uint32_t ntiles,
                                                     uint32_t tile, int shift,
                                                     uint64_t* __restrict__ tb) {
  const uint32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t > ntiles) return;
  const uint64_t c = min((uint64_t)t * tile, ncells1);
  tb[t] = lower_bound_shifted(keys, 0, n, c, shift);
}

__global__ __launch_bounds__(256) void k_cell_fill(const uint32_t* __restrict__ keys,
                                                   uint64_t ncells1, uint32_t tile, int shift,
                                                   const uint64_t* __restrict__ tb,
                                                   uint32_t* __restrict__ cs) {
  __shared__ uint32_t s_keys[kCellTileKeys];
  const uint32_t t = blockIdx.x;
  const uint64_t c0 = (uint64_t)t * tile;
  const uint64_t j0 = tb[t], j1 = tb[t + 1];
  const uint64_t m = j1 - j0;
  const bool staged = m <= kCellTileKeys;
  if (staged)
    for (uint32_t i = threadIdx.x; i < m; i += 256) s_keys[i] = keys[j0 + i] >> shift;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < tile; i += 256) {
    const uint64_t c = c0 + i;
    if (c >= ncells1) break;
    uint64_t pos;
    if (staged) {
      uint32_t lo = 0, hi = (uint32_t)m;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)s_keys[mid] < c) lo = mid + 1; else hi = mid;
      }
      pos = j0 + lo;
    } else {
      pos = lower_bound_shifted(keys, j0, j1, c, shift);
    }
    cs[c] = (uint32_t)pos;
  }
}

uint64_t cell_bounds_tmp_words(uint64_t ncells) {
  const uint32_t tile = cell_tile(ncells + 1);
  return 2 * ((ncells + 1 + tile - 1) / tile + 1) + 2;
}

hipError_t cell_bounds(const uint32_t* sorted_keys, uint64_t n, uint32_t ncells,
                       uint32_t* cell_start, uint32_t* tmp, hipStream_t st, int shift) {
  const uint64_t ncells1 = (uint64_t)ncells + 1;
  const uint32_t tile = cell_tile(ncells1);
  const uint32_t ntiles = (uint32_t)((ncells1 + tile - 1) / tile);
  uint64_t* tb = reinterpret_cast<uint64_t*>(tmp);
  hipLaunchKernelGGL(k_tile_bounds, dim3((ntiles + 1 + 255) / 256), dim3(256), 0, st,
                     sorted_keys, n, ncells1, ntiles, tile, shift, tb);
  hipLaunchKernelGGL(k_cell_fill, dim3(ntiles), dim3(256), 0, st, sorted_keys, ncells1, tile,
                     shift, tb, cell_start);
  return hipGetLastError();
}

// Level-0 Morton key of each point's cell on grid g (cubic, 2^bits cells per
// axis); the sort by this key makes every coarser level's cells contiguous.
__global__ __launch_bounds__(256) void k_morton_keys(const float* __restrict__ xyz, uint64_t n,
                                                     const GridDesc* __restrict__ grid,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const GridDesc g = *grid;
  uint32_t cx = (uint32_t)cell_of(xyz[3 * i], g.mn[0], g.inv, g.dim[0]);
  uint32_t cy = (uint32_t)cell_of(xyz[3 * i + 1], g.mn[1], g.inv, g.dim[1]);
  uint32_t cz = (uint32_t)cell_of(xyz[3 * i + 2], g.mn[2], g.inv, g.dim[2]);
  keys[i] = morton3(cx, cy, cz);
  vals[i] = (uint32_t)i;
}

hipError_t morton_keys(const float* xyz, uint64_t n, const GridDesc* g0, uint32_t* keys,
                       uint32_t* vals, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton_keys, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, xyz,
                     n, g0, keys, vals);
  return hipGetLastError();
}

// counts[cell of spts[j]] over the sorted points; consecutive points share a
// cell, so each wave adds one count per run of equal cells (one atomic per
// run instead of per point).
__global__ __launch_bounds__(256) void k_count_sorted(const float4* __restrict__ spts,
                                                      uint64_t n,
                                                      const GridDesc* __restrict__ grid,
                                                      uint32_t* __restrict__ counts) {
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  const GridDesc g = *grid;
  uint32_t key = 0xffffffffu;
  if (j < n) {
    const float4 p = spts[j];
    const int cx = cell_of(p.x, g.mn[0], g.inv, g.dim[0]);
    const int cy = cell_of(p.y, g.mn[1], g.inv, g.dim[1]);
    const int cz = cell_of(p.z, g.mn[2], g.inv, g.dim[2]);
    key = ((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) * (uint32_t)g.dim[0] + (uint32_t)cx;
  }
  const uint32_t prev = __shfl_up(key, 1, kWave);
  const bool head = j < n && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(head);
  if (head) {
    // run ends at the next head (or at the end of the wave / of the data)
    const uint64_t after = lane == kWave - 1 ? 0ull : (heads >> (lane + 1)) << (lane + 1);
    const int next = after ? __ffsll((long long)after) - 1 : kWave;
    const uint64_t j0 = j - (uint64_t)lane;
    const uint64_t valid = min((uint64_t)kWave, n - j0);  // lanes holding points
    const uint32_t len = (uint32_t)(min((uint64_t)next, valid) - (uint64_t)lane);
    atomicAdd(&counts[key], len);
  }
}

hipError_t count_sorted(const float4* spts, uint64_t n, const GridDesc* g, uint32_t* counts,
                        hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_sorted, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, spts,
                     n, g, counts);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_count_cells(const float* __restrict__ xyz, uint64_t n,
                                                     const GridDesc* __restrict__ grid,
                                                     uint32_t* __restrict__ counts) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const GridDesc g = *grid;
  int cx = cell_of(xyz[3 * i], g.mn[0], g.inv, g.dim[0]);
  int cy = cell_of(xyz[3 * i + 1], g.mn[1], g.inv, g.dim[1]);
  int cz = cell_of(xyz[3 * i + 2], g.mn[2], g.inv, g.dim[2]);
  atomicAdd(&counts[((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) * (uint32_t)g.dim[0] +
                    (uint32_t)cx],
            1u);
}

hipError_t count_cells(const float* xyz, uint64_t n, const GridDesc* g, uint32_t* counts,
                       hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_count_cells, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, xyz,
                     n, g, counts);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ xyz,
                                                const uint32_t* __restrict__ vals, uint64_t n,
                                                float4* __restrict__ spts) {
  uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  uint32_t i = vals[j];
  spts[j] = make_float4(xyz[3ull * i], xyz[3ull * i + 1], xyz[3ull * i + 2],
                        __uint_as_float(i));
}

// kWave / kChunkPts chunks of kChunkPts sorted points per wave: each chunk's
// bounding box (min at box[2c], max at box[2c+1]) by a reduction over its
// lanes (xor shuffles below kChunkPts stay inside the chunk).
__global__ __launch_bounds__(256) void k_chunk_boxes(const float4* __restrict__ spts, uint64_t n,
                                                     float4* __restrict__ box) {
  constexpr uint32_t kPerWave = kWave / kChunkPts;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint64_t c = ((uint64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave) * kPerWave +
                     lane / kChunkPts;
  const uint64_t nch = (n + kChunkPts - 1) / kChunkPts;
  const bool valid = c < nch;
  const uint64_t j = c * kChunkPts + lane % kChunkPts;
  const float4 p = spts[!valid ? 0 : (j < n ? j : c * kChunkPts)];
  float mn[3] = {p.x, p.y, p.z}, mx[3] = {p.x, p.y, p.z};
#pragma unroll
  for (int o = kChunkPts / 2; o > 0; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, kWave));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, kWave));
    }
  }
  if (valid && lane % kChunkPts == 0) {
    box[2 * c] = make_float4(mn[0], mn[1], mn[2], 0.0f);
    box[2 * c + 1] = make_float4(mx[0], mx[1], mx[2], 0.0f);
  }
}
static_assert(kWave % kChunkPts == 0, "whole chunks per wave");

hipError_t chunk_boxes(const float4* spts, uint64_t n, float4* box, hipStream_t st) {
  const uint64_t nch = (n + kChunkPts - 1) / kChunkPts;
  if (nch == 0) return hipSuccess;
  const uint64_t waves = (nch + kWave / kChunkPts - 1) / (kWave / kChunkPts);
  hipLaunchKernelGGL(k_chunk_boxes, dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, st, spts,
                     n, box);
  return hipGetLastError();
}

hipError_t gather_points(const float* xyz, const uint32_t* vals, uint64_t n, float4* spts,
                         hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, xyz, vals,
                     n, spts);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_query_morton(const float* __restrict__ q, uint64_t nq,
                                                      const GridDesc* __restrict__ grid,
                                                      int shift, uint32_t* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= nq) return;
  const GridDesc g = *grid;
  if (shift < 0) {
    // Grid planned on the device: 8 Morton bits per axis of the widest
    // axis's cell coordinate (the host's rule for a known grid).
    const uint32_t md = (uint32_t)max(max(g.dim[0], g.dim[1]), g.dim[2]) - 1u;
    const int bits = md ? 32 - __clz((int)md) : 0;
    shift = max(0, bits - 8);
  }
  // Queries outside the box (or NaN) clamp to the border cells.
  uint32_t cx = (uint32_t)cell_of(q[3 * i], g.mn[0], g.inv, g.dim[0]) >> shift;
  uint32_t cy = (uint32_t)cell_of(q[3 * i + 1], g.mn[1], g.inv, g.dim[1]) >> shift;
  uint32_t cz = (uint32_t)cell_of(q[3 * i + 2], g.mn[2], g.inv, g.dim[2]) >> shift;
  keys[i] = morton3(cx, cy, cz);
  vals[i] = (uint32_t)i;
}

hipError_t query_morton_keys(const float* q, uint64_t nq, const GridDesc* g, int shift,
                             uint32_t* keys, uint32_t* vals, hipStream_t st) {
  if (nq == 0) return hipSuccess;
  hipLaunchKernelGGL(k_query_morton, dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, st, q,
                     nq, g, shift, keys, vals);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_morton(const uint32_t* __restrict__ xyz, uint64_t n,
                                                uint32_t* __restrict__ codes) {
  uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  codes[i] = morton3(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
}

hipError_t morton3_codes(const uint32_t* xyz, uint64_t n, uint32_t* codes, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_morton, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, xyz, n,
                     codes);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    p[i] = v;
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    p[i] = (uint32_t)i;
}

static dim3 fill_grid(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return dim3((uint32_t)(b < 65536 ? (b ? b : 1) : 65536));
}

hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, fill_grid(n), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t iota_u32(uint32_t* p, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, fill_grid(n), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_scatter_rows(const float4* __restrict__ qproc,
                                                      const float* __restrict__ val, uint64_t n,
                                                      float r, float* __restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    out[__float_as_uint(qproc[i].w)] = fminf(val[i], r);
}

hipError_t scatter_rows(const float4* qproc, const float* val, uint64_t n, float r, float* out,
                        hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_scatter_rows, fill_grid(n), dim3(256), 0, st, qproc, val, n, r, out);
  return hipGetLastError();
}

}  // namespace rtnn
