// Cell-walk search kernels for gfx950.
//
// They replace the reference's OptiX programs: __raygen__radius/__raygen__knn
// (src/optixNSearch/camera.cu:39-145) and the intersection programs
// (geometry.cu:38-191).  The per-candidate test is the reference's: d2 =
// dot(q - p, q - p) in FP32 against R*R, strict; range search keeps the query
// itself (geometry.cu:65), KNN drops d2 == 0 (geometry.cu:187).  Where the
// reference keeps the first `limit` hits in traversal order (range,
// geometry.cu:72-83) or an unsorted replace-max array (KNN,
// geometry.cu:108-155), these kernels return the k smallest by (d2, index),
// sorted, so the result is deterministic.
//
// A query visits the cells overlapping [q - R, q + R] (padded, see
// cell_range).  Cells are raster ordered with x fastest, so each (y, z) row of
// the block is one contiguous run of sorted points: a 3x3x3 walk is 9 runs.
//
// k_topk_thread : one query per lane.  A sorted array of k 64-bit keys
//                 (d2 bits << 32 | index) lives in VGPRs; insertion is an
//                 unrolled compare/select pass, so nothing spills to scratch.
// k_topk_group  : G lanes per query (G = 2, 4, 8, 16, 64).  The lanes stride over the
//                 query's runs flattened into one index space, test 16-byte
//                 candidates with coalesced loads, and append hits to an LDS
//                 buffer with one wave64 ballot + prefix popcount.  When the
//                 buffer nears capacity the k smallest are kept (rank select)
//                 and later hits must beat the k-th.  The final rank sort
//                 writes the row in (d2, index) order.
#include <float.h>
#include <limits.h>

#include "common.h"
#include "internal.h"

namespace rtnn {

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t lanemask_lt64() {
  const int lane = threadIdx.x & (kWave - 1);
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

// d2 < R2 as d2 <= r2_limit(R2): the float just below R2 (-1 for R2 <= 0,
// FLT_MAX for +inf).  After a select the limit becomes the k-th key's d2:
// ties with it are appended too and resolved by the (d2, index) rank.
__device__ inline float r2_limit(float R2) {
  return R2 > 0.0f ? __uint_as_float(__float_as_uint(R2) - 1u) : -1.0f;
}

// Resolves tile t into its query range and grid.
__device__ inline void tile_of(const SearchArgs& a, uint32_t nq, uint32_t t, uint32_t& q0,
                               uint32_t& q1, uint32_t& gi) {
  if (a.tiles) {
    Tile tl = a.tiles[t];
    q0 = tl.q0; q1 = tl.q1; gi = tl.grid;
  } else {
    q0 = t * a.tile_q;
    q1 = min(q0 + a.tile_q, nq);
    gi = 0;
  }
}

// Query count of this launch (device-resident for the fallback pass).
__device__ inline uint32_t launch_nq(const SearchArgs& a) {
  return a.nq_dev ? *a.nq_dev : a.nq;
}

// Persistent launches walk tiles in XCD-contiguous chunks; plain launches map
// each block to one tile through xcd_swizzle.
struct TileIter {
  uint32_t t, end, step;
  __device__ TileIter(const SearchArgs& a, uint32_t nq) {
    uint32_t ntiles = a.nq_dev ? (nq + a.tile_q - 1) / a.tile_q : a.ntiles;
    if (a.persistent) {
      constexpr uint32_t kXcd = 8;
      uint32_t x = blockIdx.x % kXcd, j = blockIdx.x / kXcd;
      uint32_t per = (gridDim.x + kXcd - 1 - x) / kXcd;  // blocks on this XCD
      uint32_t q = ntiles / kXcd, rem = ntiles % kXcd;
      uint32_t start = x * q + (x < rem ? x : rem);
      end = start + q + (x < rem ? 1u : 0u);
      t = start + j;
      step = per;
    } else {
      t = xcd_swizzle(blockIdx.x, gridDim.x);
      end = t + 1;
      step = 1;
    }
  }
};

// Appends the processing positions of short rows (fewer than k hits) with
// one atomic per wave; a wave's entries stay together and in lane order, so
// the re-run over the list keeps the Morton locality of the queries.
__device__ inline void append_short(const SearchArgs& a, bool is_short, uint32_t qs,
                                    uint32_t hits, float R) {
  const uint64_t m = __ballot(is_short);
  if (m == 0) return;
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if ((int)(threadIdx.x & (kWave - 1)) == leader)
    base = atomicAdd(a.short_count, (uint32_t)__popcll(m));
  base = __shfl(base, leader, kWave);
  if (is_short) {
    const uint32_t pos = base + (uint32_t)__popcll(m & lanemask_lt64());
    a.short_rows[pos] = qs;
    if (a.short_hits) a.short_hits[pos] = hits;
    if (a.radius_out) a.radius_out[qs] = R;
  }
}

// Search radius of launch position qi (query qs).
__device__ inline float query_radius(const SearchArgs& a, uint32_t qi, uint32_t qs) {
  if (!a.qradius) return a.r;
  const float sc = a.qsel_hits ? hit_scale(a.qsel_hits[qi], a.k) : a.qradius_scale;
  return fminf(a.r, a.qradius[qs] * sc);
}

// Candidates loaded per lane before the first is tested (thread kernel;
// group kernel: its G lanes share a query, so each lane takes more), and
// runs (rows or cells) whose bounds each group lane loads per batch.
constexpr int kUnroll = 4;
constexpr int kGroupUnroll = 8;
static_assert(64 * kGroupUnroll * 16 <= (int)kSptsPadBytes * 2, "padding of the sorted points");
constexpr int kRunsPerLane = 1;

// --------------------------------------------------------------------------
// thread per query

template <int KC>
__device__ inline void topk_insert(uint64_t (&list)[KC], uint64_t key) {
#pragma unroll
  for (int i = KC - 1; i > 0; --i) {
    const uint64_t prev = list[i - 1];
    const uint64_t cur = list[i];
    list[i] = prev > key ? prev : (cur > key ? key : cur);
  }
  list[0] = list[0] > key ? key : list[0];
}

// Waves per SIMD asked of the compiler (as for the octree walks: a latency-
// bound walk, residency before registers).  Spheres KNN, K = 16 (10M x 10M,
// r = 10): 10.07 ms at 117 VGPRs (four waves), 8.60 ms at 90 (five, no
// spills); Gaussian-mixture KNN, K = 8: 10.01 / 10.00 ms.
template <int KC>
constexpr int thread_waves() { return KC <= 8 ? 8 : KC <= 16 ? 5 : KC <= 32 ? 3 : 1; }

template <int KC, bool KNN, bool HIER>
__global__ __launch_bounds__(kTileQ) __attribute__((amdgpu_waves_per_eu(thread_waves<KC>())))
void k_topk_thread(const SearchArgs a) {
  const uint32_t nq = launch_nq(a);
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const uint32_t qi = q0 + threadIdx.x;
    const bool active = qi < q1;
    const uint32_t qs = active ? (a.qsel ? a.qsel[qi] : qi) : 0u;
    const GridDesc g = a.grids[gi];
    const float4 qv = active ? a.queries[qs] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t row = __float_as_uint(qv.w);
    const float R = active ? query_radius(a, qi, qs) : a.r;
    const float R2 = R * R;
    int lo0, hi0, lo1, hi1, lo2, hi2;
    cell_range(g, qv.x, R, 0, lo0, hi0);
    cell_range(g, qv.y, R, 1, lo1, hi1);
    cell_range(g, qv.z, R, 2, lo2, hi2);
    uint32_t lvl_base = g.cell_base;
    if (HIER) {
      // Finest built level whose cell is >= R * alpha; the level-0 range
      // shifted right by the level is that level's range (floor division is
      // monotone and keeps empty ranges empty).
      int lvl = a.level_min;
      float c = ldexpf(g.cell, lvl);
      const float want = R * a.level_alpha;
      while (lvl < a.nlevels - 1 && c < want) {
        c *= 2.0f;
        ++lvl;
      }
      lo0 >>= lvl; hi0 >>= lvl; lo1 >>= lvl; hi1 >>= lvl; lo2 >>= lvl; hi2 >>= lvl;
      lvl_base = a.grids[lvl].cell_base;
    }
    const int nx = hi0 - lo0 + 1, ny = hi1 - lo1 + 1, nz = hi2 - lo2 + 1;
    // raster: one run per (y, z) row; hierarchical: one run per Morton cell
    const int nrows = (active && nx > 0 && ny > 0 && nz > 0) ? (HIER ? nx : 1) * ny * nz : 0;

    uint64_t list[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) list[i] = ~0ull;

    // One contiguous run of sorted points per raster row (or Morton cell).
    auto row_bounds = [&](int rr, uint32_t& s, uint32_t& e) {
      if (HIER) {
        const int cx = lo0 + rr % nx, t = rr / nx;
        const int cy = lo1 + t % ny, cz = lo2 + t / ny;
        const uint32_t id = lvl_base + morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
        s = a.cell_start[id];
        e = a.cell_start[id + 1];
      } else {
        const int cy = lo1 + rr % ny, cz = lo2 + rr / ny;
        const uint32_t rowc = lvl_base + ((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) *
                                             (uint32_t)g.dim[0];
        s = a.cell_start[rowc + lo0];
        e = a.cell_start[rowc + hi0 + 1];
      }
    };
    // The next run's bounds are loaded while the current run is scanned.
    uint32_t s_next = 0, e_next = 0;
    if (nrows > 0) row_bounds(0, s_next, e_next);
    for (int rr = 0; rr < nrows; ++rr) {
      const uint32_t s = s_next, e = e_next;
      if (rr + 1 < nrows) row_bounds(rr + 1, s_next, e_next);
      for (uint32_t j0 = s; j0 < e; j0 += kUnroll) {
        // kUnroll loads in flight before the first test: unconditional loads
        // (a slot past the run reloads its last point and is not tested).
        float4 pv[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) pv[u] = a.spts[min(j0 + (uint32_t)u, e - 1)];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          if (j0 + u >= e) break;
          const float4 p = pv[u];
          const float d2 = d2_of(qv.x, qv.y, qv.z, p.x, p.y, p.z);
          const bool ok = KNN ? (d2 > 0.0f && d2 < R2) : (d2 < R2);
          if (ok) {
            const uint64_t key = pack_key(d2, __float_as_uint(p.w));
            if (key < list[KC - 1]) topk_insert<KC>(list, key);
          }
        }
      }
    }
    if (!active) continue;

    const uint32_t k = a.k;
    const uint32_t base = a.local_index ? g.point_base : 0u;
    int32_t* oi = a.out_idx + (uint64_t)row * k;
    float* od = a.out_d2 ? a.out_d2 + (uint64_t)row * k : nullptr;
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      if ((uint32_t)i < k) {
        const bool have = list[i] != ~0ull;
        oi[i] = have ? (int32_t)(key_idx(list[i]) - base) : -1;
        if (od) od[i] = have ? key_d2(list[i]) : __builtin_huge_valf();
      }
    }
    if (a.short_rows) {
      uint32_t hits = 0;
#pragma unroll
      for (int i = 0; i < KC; ++i) hits += ((uint32_t)i < k && list[i] != ~0ull) ? 1u : 0u;
      append_short(a, hits < k && R < a.r, qs, hits, R);  // a row searched at r is final
    }
  }
}

// --------------------------------------------------------------------------
// KNN on the hierarchical Morton grid: octree walk, one query per lane
//
// A level-l cell m has children 8m .. 8m+7 at level l-1 (the Morton key
// nests), and every cell's points are one contiguous run of the sorted
// points, so the levels form an implicit octree.  From each top-level cell
// within [q - R, q + R] the walk goes depth first, the child holding the
// query first (child j visits octant j ^ the query's octant), without a
// stack: the position among siblings is recovered from the key.  A cell
// whose padded box lies farther than the current limit is skipped; a cell of
// the finest built level, or holding at most a.tree_leaf points, is scanned,
// chunk by chunk: the bounding box of every kChunkPts sorted points skips the
// chunks beyond the limit.  The leaf size is an option: on the full-size
// Levy-flight KNN (learned radius) a step took 6.9 s with leaves of up to 256
// points, 4.1 s with 4096, 1.3 s with 65536; the Gaussian-mixture KNN (1M
// points) is fastest at 256 (2.9 ms, 3.5 ms with 65536).
// Top-level cells are taken in rings of growing distance around the query's
// own, so the limit shrinks early.
// The limit starts at the float below R^2 and, once k keys are held, drops
// to the k-th key's d2, so a dense cluster beyond the k nearest is cut off
// at its coarsest cell.  The result is the same as the cube walk's: only
// cells that cannot hold a hit (d2 <= limit) are skipped.

__device__ inline uint32_t compact3(uint32_t x) {
  x &= 0x09249249u;
  x = (x | (x >> 2)) & 0x030c30c3u;
  x = (x | (x >> 4)) & 0x0300f00fu;
  x = (x | (x >> 8)) & 0xff0000ffu;
  x = (x | (x >> 16)) & 0x000003ffu;
  return x;
}

// Squared distance from q to cell (ix, iy, iz) of edge c, the box grown by
// pad on every side (a lower bound for any point the FP32 cell mapping puts
// in the cell).
__device__ inline float box_d2(const float4& q, const GridDesc& g0, float c, float pad,
                               uint32_t ix, uint32_t iy, uint32_t iz) {
  const float lx = g0.mn[0] + (float)ix * c - pad, hx = lx + c + 2.0f * pad;
  const float ly = g0.mn[1] + (float)iy * c - pad, hy = ly + c + 2.0f * pad;
  const float lz = g0.mn[2] + (float)iz * c - pad, hz = lz + c + 2.0f * pad;
  const float dx = fmaxf(fmaxf(lx - q.x, q.x - hx), 0.0f);
  const float dy = fmaxf(fmaxf(ly - q.y, q.y - hy), 0.0f);
  const float dz = fmaxf(fmaxf(lz - q.z, q.z - hz), 0.0f);
  return dx * dx + dy * dy + dz * dz;
}

// Waves per SIMD asked of the compiler for the octree walks: the walk is
// bound by the latency of its dependent loads, so residency pays more than
// registers, up to a few spilled ones.  Levy-flight KNN (30M points, 30M
// uniform queries, r = 100), K = 32, two lanes: 908 ms at 137 VGPRs (three
// waves per SIMD), 756 ms at 128 (four; 48 bytes spilled), 726 ms at 96
// (five; 176 bytes); one lane: 1094 ms at 193 VGPRs (two waves), 802 ms at
// 168 (three), 993 ms at 128 (four; 264 bytes).  K = 16, two lanes: 473 ms
// at 105 VGPRs (four waves), 411 at 96 (five), 379 at 80 (six); one lane:
// 544 ms at 154 (three), 381 at 95 (five), 351 at 80 (six).  K = 8, two
// lanes: 328 ms at 88 (five), 302 at 80 (six), 283 at 64 (eight); one lane:
// 363 ms at 106 (four), 307 at 80 (six), 295 at 64 (eight).  Six waves at
// K = 32 spill 248 bytes and take 1402 ms.  Rows are unchanged throughout.
// Beyond K = 32 no target is set (1: the compiler's own choice).
template <int KC>
constexpr int treeg_waves() { return KC <= 8 ? 8 : KC <= 16 ? 6 : KC <= 32 ? 5 : 1; }
template <int KC>
constexpr int tree_waves() { return KC <= 8 ? 8 : KC <= 16 ? 6 : KC <= 32 ? 3 : 1; }

template <int KC>
__global__ __launch_bounds__(kTileQ) __attribute__((amdgpu_waves_per_eu(tree_waves<KC>())))
void k_topk_tree(const SearchArgs a) {
  __shared__ uint32_t s_base[8];
  if (threadIdx.x < 8) s_base[threadIdx.x] = (int)threadIdx.x < a.nlevels ?
                                               a.grids[threadIdx.x].cell_base : 0u;
  __syncthreads();
  const uint32_t nq = launch_nq(a);
  const int T = a.nlevels - 1, L0 = a.level_min;
  const GridDesc g0 = a.grids[0];
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const uint32_t qi = q0 + threadIdx.x;
    if (qi >= q1) continue;
    const uint32_t qs = a.qsel ? a.qsel[qi] : qi;
    const float4 qv = a.queries[qs];
    const uint32_t row = __float_as_uint(qv.w);
    const float R = query_radius(a, qi, qs);
    float lim = r2_limit(R * R);
    uint64_t list[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) list[i] = ~0ull;
    // top-level cells of [q - R, q + R] (the level-0 range shifted by T)
    int lo0, hi0, lo1, hi1, lo2, hi2;
    cell_range(g0, qv.x, R, 0, lo0, hi0);
    cell_range(g0, qv.y, R, 1, lo1, hi1);
    cell_range(g0, qv.z, R, 2, lo2, hi2);
    // the query's level-0 cell (clamped): its octant at every level
    const int qc0 = min(max((int)floorf((qv.x - g0.mn[0]) * g0.inv), 0), g0.dim[0] - 1);
    const int qc1 = min(max((int)floorf((qv.y - g0.mn[1]) * g0.inv), 0), g0.dim[1] - 1);
    const int qc2 = min(max((int)floorf((qv.z - g0.mn[2]) * g0.inv), 0), g0.dim[2] - 1);
    const uint32_t qm = morton3((uint32_t)qc0, (uint32_t)qc1, (uint32_t)qc2);
    const float mag = fabsf(g0.mn[0]) + fabsf(g0.mn[1]) + fabsf(g0.mn[2]) + fabsf(qv.x) +
                      fabsf(qv.y) + fabsf(qv.z) + (float)g0.dim[0] * g0.cell;
    if (hi0 >= lo0 && hi1 >= lo1 && hi2 >= lo2) {
      const int tl0 = lo0 >> T, th0 = hi0 >> T, tl1 = lo1 >> T, th1 = hi1 >> T;
      const int tl2 = lo2 >> T, th2 = hi2 >> T;
      // Top-level cells in rings of growing Chebyshev distance around the
      // query's own top cell, so the nearest points come first and the
      // limit shrinks before the far cells are tested (each cell of the
      // range is on exactly one ring; the result does not depend on order).
      const int qt0 = qc0 >> T, qt1 = qc1 >> T, qt2 = qc2 >> T;
      const int nring = max(max(max(qt0 - tl0, th0 - qt0), max(qt1 - tl1, th1 - qt1)),
                            max(qt2 - tl2, th2 - qt2));
      for (int ring = 0; ring <= nring; ++ring)
      for (int tz = max(tl2, qt2 - ring); tz <= min(th2, qt2 + ring); ++tz)
        for (int ty = max(tl1, qt1 - ring); ty <= min(th1, qt1 + ring); ++ty)
          for (int tx = max(tl0, qt0 - ring); tx <= min(th0, qt0 + ring); ++tx) {
            if (max(max(abs(tx - qt0), abs(ty - qt1)), abs(tz - qt2)) != ring) continue;
            int l = T;
            uint32_t m = morton3((uint32_t)tx, (uint32_t)ty, (uint32_t)tz);
            for (;;) {
              const float c = ldexpf(g0.cell, l);
              const float pad = c * 0x1.0p-12f + mag * 0x1.0p-19f;
              const float bd = box_d2(qv, g0, c, pad, compact3(m), compact3(m >> 1),
                                      compact3(m >> 2));
              bool down = false;
              if (bd <= lim * 1.001f) {
                const uint32_t cb = s_base[l] + m;
                const uint32_t s0 = a.cell_start[cb], e = a.cell_start[cb + 1];
                if (e > s0) {
                  if (l > L0 && e - s0 > a.tree_leaf) {
                    down = true;
                  } else {
                    // A leaf of more than one chunk goes chunk by chunk: a
                    // chunk whose box lies beyond the limit (with the cells'
                    // FP32 margin) holds no hit.
                    const bool chunked = a.chunk_box && e - s0 > kChunkPts;
                    for (uint32_t c0 = s0; c0 < e;) {
                      uint32_t c1 = e;
                      if (chunked) {
                        const uint32_t ch = c0 / kChunkPts;
                        c1 = min(e, (ch + 1) * kChunkPts);
                        const float4 blo = a.chunk_box[2 * ch], bhi = a.chunk_box[2 * ch + 1];
                        const float bx = fmaxf(fmaxf(blo.x - qv.x, qv.x - bhi.x), 0.0f);
                        const float by = fmaxf(fmaxf(blo.y - qv.y, qv.y - bhi.y), 0.0f);
                        const float bz = fmaxf(fmaxf(blo.z - qv.z, qv.z - bhi.z), 0.0f);
                        if ((bx * bx + by * by) + bz * bz > lim * 1.001f) {
                          c0 = c1;
                          continue;
                        }
                      }
                      for (uint32_t j = c0; j < c1; ++j) {
                        const float4 p = a.spts[j];
                        const float d2 = d2_of(qv.x, qv.y, qv.z, p.x, p.y, p.z);
                        if (d2 > 0.0f && d2 <= lim) {
                          const uint64_t key = pack_key(d2, __float_as_uint(p.w));
                          if (key < list[KC - 1]) {
                            topk_insert<KC>(list, key);
                            if (list[KC - 1] != ~0ull) lim = fminf(lim, key_d2(list[KC - 1]));
                          }
                        }
                      }
                      c0 = c1;
                    }
                  }
                }
              }
              if (down) {
                --l;
                m = (m << 3) | ((qm >> (3 * l)) & 7u);
                continue;
              }
              // next sibling in visiting order, climbing past last children
              while (l < T && (((m & 7u) ^ ((qm >> (3 * l)) & 7u)) == 7u)) {
                m >>= 3;
                ++l;
              }
              if (l == T) break;
              const uint32_t o = (qm >> (3 * l)) & 7u;
              m = (m & ~7u) | ((((m & 7u) ^ o) + 1u) ^ o);
            }
          }
    }
    const uint32_t k = a.k;
    const uint32_t base = a.local_index ? g0.point_base : 0u;
    int32_t* oi = a.out_idx + (uint64_t)row * k;
    float* od = a.out_d2 ? a.out_d2 + (uint64_t)row * k : nullptr;
    uint32_t hits = 0;
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      if ((uint32_t)i < k) {
        const bool have = list[i] != ~0ull;
        hits += have ? 1u : 0u;
        oi[i] = have ? (int32_t)(key_idx(list[i]) - base) : -1;
        if (od) od[i] = have ? key_d2(list[i]) : __builtin_huge_valf();
      }
    }
    if (a.short_rows) append_short(a, hits < k && R < a.r, qs, hits, R);
  }
}

// --------------------------------------------------------------------------
// The octree walk with G lanes per query
//
// k_topk_tree walks one query per lane; a walk that crosses a dense cluster
// scans its leaves chunk after chunk on that one lane while the L2 latency of
// every load is exposed (the Levy-flight KNN: 0.30 of the VALU slots, 97 %
// L2 hits).  Here the G lanes of a group walk the same octree path -- every
// decision reads a limit the group shares, so their control flow is the
// same -- and deal each leaf's chunks to them round-robin (lane l tests and
// scans chunks l, l + G, ...; a leaf of one chunk has its points split), each
// into a top-k list of its own.  After each round the group takes the
// smallest of its lanes' limits: once a lane holds
// k keys with d2 <= t, no point beyond t is among the group's k nearest, so
// every lane may prune at the smallest lane limit.  At the end the G sorted
// lists are merged by repeatedly taking the group's smallest head (the
// points of one list are distinct from every other's, so each head is
// unique); the row is the first k taken.  Same result as k_topk_tree.
// On the Levy-flight KNN (30M x 30M, leaves of 65536 points; builds without
// occupancy targets: two waves per SIMD for one lane, three for several) one
// lane took 1080 ms, two lanes 900 ms, four 1062 ms, eight 1245 ms: beyond
// two, the walk every lane repeats (cell bounds, boxes) outweighs the split
// chunks.

template <int G>
__device__ inline float group_min_f(float v) {
#pragma unroll
  for (int o = 1; o < G; o <<= 1) v = fminf(v, __shfl_xor(v, o, G));
  return v;
}

template <int G>
__device__ inline uint64_t group_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    const uint64_t w = (uint64_t)__shfl_xor((unsigned long long)v, o, G);
    v = w < v ? w : v;
  }
  return v;
}

template <int KC, int G>
__global__ __launch_bounds__(kTileQ) __attribute__((amdgpu_waves_per_eu(treeg_waves<KC>())))
void k_topk_treeg(const SearchArgs a) {
  __shared__ uint32_t s_base[8];
  if (threadIdx.x < 8) s_base[threadIdx.x] = (int)threadIdx.x < a.nlevels ?
                                               a.grids[threadIdx.x].cell_base : 0u;
  __syncthreads();
  const uint32_t nq = launch_nq(a);
  const int T = a.nlevels - 1, L0 = a.level_min;
  const GridDesc g0 = a.grids[0];
  const int gq = threadIdx.x / G, l = threadIdx.x % G;
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const uint32_t qi = q0 + gq;
    if (qi >= q1) continue;  // the whole group
    const uint32_t qs = a.qsel ? a.qsel[qi] : qi;
    const float4 qv = a.queries[qs];
    const uint32_t row = __float_as_uint(qv.w);
    const float R = query_radius(a, qi, qs);
    float lim = r2_limit(R * R);
    uint64_t list[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) list[i] = ~0ull;
    // points j0, j0 + step, ... < j1 into the lane's list
    auto scan = [&](uint32_t j0, uint32_t j1, uint32_t step) {
      for (uint32_t j = j0; j < j1; j += step) {
        const float4 p = a.spts[j];
        const float d2 = d2_of(qv.x, qv.y, qv.z, p.x, p.y, p.z);
        if (d2 > 0.0f && d2 <= lim) {
          const uint64_t key = pack_key(d2, __float_as_uint(p.w));
          if (key < list[KC - 1]) {
            topk_insert<KC>(list, key);
            if (list[KC - 1] != ~0ull) lim = fminf(lim, key_d2(list[KC - 1]));
          }
        }
      }
    };
    int lo0, hi0, lo1, hi1, lo2, hi2;
    cell_range(g0, qv.x, R, 0, lo0, hi0);
    cell_range(g0, qv.y, R, 1, lo1, hi1);
    cell_range(g0, qv.z, R, 2, lo2, hi2);
    const int qc0 = min(max((int)floorf((qv.x - g0.mn[0]) * g0.inv), 0), g0.dim[0] - 1);
    const int qc1 = min(max((int)floorf((qv.y - g0.mn[1]) * g0.inv), 0), g0.dim[1] - 1);
    const int qc2 = min(max((int)floorf((qv.z - g0.mn[2]) * g0.inv), 0), g0.dim[2] - 1);
    const uint32_t qm = morton3((uint32_t)qc0, (uint32_t)qc1, (uint32_t)qc2);
    const float mag = fabsf(g0.mn[0]) + fabsf(g0.mn[1]) + fabsf(g0.mn[2]) + fabsf(qv.x) +
                      fabsf(qv.y) + fabsf(qv.z) + (float)g0.dim[0] * g0.cell;
    if (hi0 >= lo0 && hi1 >= lo1 && hi2 >= lo2) {
      const int tl0 = lo0 >> T, th0 = hi0 >> T, tl1 = lo1 >> T, th1 = hi1 >> T;
      const int tl2 = lo2 >> T, th2 = hi2 >> T;
      const int qt0 = qc0 >> T, qt1 = qc1 >> T, qt2 = qc2 >> T;
      const int nring = max(max(max(qt0 - tl0, th0 - qt0), max(qt1 - tl1, th1 - qt1)),
                            max(qt2 - tl2, th2 - qt2));
      for (int ring = 0; ring <= nring; ++ring)
      for (int tz = max(tl2, qt2 - ring); tz <= min(th2, qt2 + ring); ++tz)
        for (int ty = max(tl1, qt1 - ring); ty <= min(th1, qt1 + ring); ++ty)
          for (int tx = max(tl0, qt0 - ring); tx <= min(th0, qt0 + ring); ++tx) {
            if (max(max(abs(tx - qt0), abs(ty - qt1)), abs(tz - qt2)) != ring) continue;
            int lv = T;
            uint32_t m = morton3((uint32_t)tx, (uint32_t)ty, (uint32_t)tz);
            for (;;) {
              const float c = ldexpf(g0.cell, lv);
              const float pad = c * 0x1.0p-12f + mag * 0x1.0p-19f;
              const float bd = box_d2(qv, g0, c, pad, compact3(m), compact3(m >> 1),
                                      compact3(m >> 2));
              bool down = false;
              if (bd <= lim * 1.001f) {
                const uint32_t cb = s_base[lv] + m;
                const uint32_t s0 = a.cell_start[cb], e = a.cell_start[cb + 1];
                if (e > s0) {
                  if (lv > L0 && e - s0 > a.tree_leaf) {
                    down = true;
                  } else if (a.chunk_box && e - s0 > kChunkPts) {
                    // the leaf's chunks dealt round-robin to the lanes, each
                    // tested against its box and scanned by its lane; the
                    // group's limit after every round
                    const uint32_t first = s0 / kChunkPts, last = (e - 1) / kChunkPts;
                    for (uint32_t cr = first; cr <= last; cr += G) {
                      const uint32_t ch = cr + (uint32_t)l;
                      if (ch <= last) {
                        const float4 blo = a.chunk_box[2 * ch], bhi = a.chunk_box[2 * ch + 1];
                        const float bx = fmaxf(fmaxf(blo.x - qv.x, qv.x - bhi.x), 0.0f);
                        const float by = fmaxf(fmaxf(blo.y - qv.y, qv.y - bhi.y), 0.0f);
                        const float bz = fmaxf(fmaxf(blo.z - qv.z, qv.z - bhi.z), 0.0f);
                        if ((bx * bx + by * by) + bz * bz <= lim * 1.001f)
                          scan(max(s0, ch * kChunkPts), min(e, (ch + 1) * kChunkPts), 1u);
                      }
                      lim = group_min_f<G>(lim);
                    }
                  } else {
                    scan(s0 + (uint32_t)l, e, G);  // a small leaf: its points split
                    lim = group_min_f<G>(lim);
                  }
                }
              }
              if (down) {
                --lv;
                m = (m << 3) | ((qm >> (3 * lv)) & 7u);
                continue;
              }
              while (lv < T && (((m & 7u) ^ ((qm >> (3 * lv)) & 7u)) == 7u)) {
                m >>= 3;
                ++lv;
              }
              if (lv == T) break;
              const uint32_t o = (qm >> (3 * lv)) & 7u;
              m = (m & ~7u) | ((((m & 7u) ^ o) + 1u) ^ o);
            }
          }
    }
    // merge: the group's smallest head, k times
    const uint32_t k = a.k;
    const uint32_t base = a.local_index ? g0.point_base : 0u;
    int32_t* oi = a.out_idx + (uint64_t)row * k;
    float* od = a.out_d2 ? a.out_d2 + (uint64_t)row * k : nullptr;
    uint32_t hits = 0;
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      if ((uint32_t)i < k) {
        const uint64_t m = group_min_u64<G>(list[0]);
        const bool have = m != ~0ull;
        hits += have ? 1u : 0u;
        if (l == i % G) {
          oi[i] = have ? (int32_t)(key_idx(m) - base) : -1;
          if (od) od[i] = have ? key_d2(m) : __builtin_huge_valf();
        }
        const bool win = have && list[0] == m;
#pragma unroll
        for (int j = 0; j < KC - 1; ++j) list[j] = win ? list[j + 1] : list[j];
        list[KC - 1] = win ? ~0ull : list[KC - 1];
      }
    }
    if (a.short_rows) append_short(a, l == 0 && hits < k && R < a.r, qs, hits, R);
  }
}

// --------------------------------------------------------------------------
// G lanes per query

template <int G>
__device__ inline uint32_t group_incl_scan(uint32_t v, int l) {
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    uint32_t t = __shfl_up(v, o, G);
    if (l >= o) v += t;
  }
  return v;
}

// DPP lane moves (VALU only, no LDS round trip).  CTRL: 0x00-0xff quad_perm,
// 0x110 + n row_shr:n, 0x141 row_half_mirror.
template <int CTRL>
__device__ inline uint32_t dpp_mov(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xf, 0xf, true);
}

// Inclusive scan over the G lanes of a group.  G <= 16 divides the 16-lane
// DPP row, so row_shr never reads across a group boundary for lanes l >= n.
template <int G>
__device__ inline uint32_t group_scan_fast(uint32_t v, int l) {
  if constexpr (G > 16) {
    return group_incl_scan<G>(v, l);
  } else {
    // Every lane executes each move (a lane with EXEC off would read as 0);
    // only the add is predicated.
    uint32_t t = dpp_mov<0x111>(v);
    v += l >= 1 ? t : 0u;
    if constexpr (G > 2) {
      t = dpp_mov<0x112>(v);
      v += l >= 2 ? t : 0u;
    }
    if constexpr (G > 4) {
      t = dpp_mov<0x114>(v);
      v += l >= 4 ? t : 0u;
    }
    if constexpr (G > 8) {
      t = dpp_mov<0x118>(v);
      v += l >= 8 ? t : 0u;
    }
    return v;
  }
}

// Value of the group's last lane, in every lane of the group.
template <int G>
__device__ inline uint32_t group_last(uint32_t v, int l) {
  if constexpr (G == 2) {
    return dpp_mov<0xF5>(v);  // quad_perm [1, 1, 3, 3]
  } else if constexpr (G == 4) {
    return dpp_mov<0xFF>(v);  // quad_perm [3, 3, 3, 3]
  } else if constexpr (G == 8) {
    const uint32_t y = dpp_mov<0xFF>(v);  // lanes 4-7 hold lane 7
    const uint32_t z = dpp_mov<0x141>(y);  // lanes 0-3 read lanes 7-4
    return l >= 4 ? y : z;
  } else {
    return __shfl(v, G - 1, G);
  }
}


template <int G, int KB, bool WIDE>
struct GroupCfg {
  static constexpr int NG = kTileQ / G;  // groups per block
  // LDS keys per group: k kept + room for later hits before the next select
  // (wide: k more; compact: 2G more, a smaller LDS footprint for more waves
  // per CU when rows rarely overflow).  Odd, so the groups of a wave start on
  // different LDS banks.
  static constexpr int CAP = (WIDE ? 2 * KB + G : KB + 2 * G) | 1;
};

// Ranks among buf[0 .. cnt) of this lane's entries e = base + l + G*i,
// i < CH, in one pass over the buffer.  Keys are distinct (the index is part
// of the key).  A slot past cnt gets key 0 (no key is below it, so its rank
// is never incremented) and rank ~0, which matches no rank < k.
template <int G, int CH>
__device__ inline void rank_chunk(const uint64_t* buf, uint32_t cnt, uint32_t base, int l,
                                  uint64_t (&kv)[CH], uint32_t (&rk)[CH]) {
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    const uint32_t e = base + (uint32_t)(l + G * i);
    kv[i] = e < cnt ? buf[e] : 0ull;
    rk[i] = e < cnt ? 0u : ~0u;
  }
  // Four buffer keys per LDS round trip.
  uint32_t j = 0;
  for (; j + 4 <= cnt; j += 4) {
    const uint64_t b0 = buf[j], b1 = buf[j + 1], b2 = buf[j + 2], b3 = buf[j + 3];
#pragma unroll
    for (int i = 0; i < CH; ++i)
      rk[i] += (b0 < kv[i] ? 1u : 0u) + (b1 < kv[i] ? 1u : 0u) + (b2 < kv[i] ? 1u : 0u) +
               (b3 < kv[i] ? 1u : 0u);
  }
  for (; j < cnt; ++j) {
    const uint64_t b = buf[j];
#pragma unroll
    for (int i = 0; i < CH; ++i) rk[i] += b < kv[i] ? 1u : 0u;
  }
}
constexpr int kRankChunk = 8;

// Number of buffer entries with d2 <= t, in every lane of the group.  Called
// with the group's lanes all active (group_scan_fast's moves stay inside it).
template <int G>
__device__ inline uint32_t group_count_le(const uint64_t* buf, uint32_t cnt, float t, int l) {
  uint32_t c = 0;
  for (uint32_t e = (uint32_t)l; e < cnt; e += G) c += key_d2(buf[e]) <= t ? 1u : 0u;
  return group_last<G>(group_scan_fast<G>(c, l), l);
}

// Cheap partial select: lowers the limit to a d2 threshold t with at least k
// entries at d2 <= t and drops the entries above it (in place, order kept).
// t comes from up to four interpolation steps between the counts at 0 and at
// the limit (aiming at about k + k/4 entries); no (d2, index) ranking.  Every
// entry that can be among the k nearest stays.  Returns with cnt >= k.
template <int G>
__device__ inline void group_shrink(uint64_t* buf, uint32_t& cnt, float& lim, uint32_t k,
                                    int l, uint64_t gmask) {
  const uint32_t target = k + (k >> 2) + 1;
  float lo = 0.0f, hi = lim;
  uint32_t clo = 0, chi = cnt;
  for (int it = 0; it < 4 && chi > target; ++it) {
    const float f = fminf(fmaxf((float)(target - clo) / (float)(chi - clo), 0.05f), 0.95f);
    const float t = lo + (hi - lo) * f;
    const uint32_t c = group_count_le<G>(buf, cnt, t, l);
    if (c >= k) {
      hi = t;
      chi = c;
    } else {
      lo = t;
      clo = c;
    }
  }
  if (chi < cnt) {
    uint32_t kept = 0;
    for (uint32_t e0 = 0; e0 < cnt; e0 += G) {
      const uint32_t e = e0 + (uint32_t)l;
      const uint64_t key = e < cnt ? buf[e] : ~0ull;
      const bool keep = e < cnt && key_d2(key) <= hi;
      const uint64_t km = __ballot(keep) & gmask;
      wave_sync();
      if (keep) buf[kept + (uint32_t)__popcll(km & lanemask_lt64())] = key;
      kept += (uint32_t)__popcll(km);
      wave_sync();
    }
    cnt = kept;
    lim = hi;
  }
}

// One step of a group of G lanes: lane l tests candidate p (if valid) and
// the group appends its hits to the LDS buffer (ballot + prefix popcount).
// Entry: cnt <= CAP - G.  When the buffer nears CAP it is cut back: first by
// a d2 threshold (group_shrink) where the buffer has room above k for it,
// else, or when ties leave it too full, to the k smallest: the k-th smallest
// key T is found by ranking, the entries <= T are compacted to the front in
// place (a round writes below the entries it has read), and later hits must
// have d2 <= d2(T).
template <int G, int KB, int CAP, bool KNN>
__device__ inline void group_take(const float4& p, bool valid, const float4& qv,
                                  uint64_t gmask, uint64_t* buf, uint32_t& cnt, float& lim,
                                  uint32_t k, int l, uint64_t* thr_slot) {
  bool hit = false;
  uint64_t key = 0;
  if (valid) {
    const float d2 = d2_of(qv.x, qv.y, qv.z, p.x, p.y, p.z);
    hit = KNN ? (d2 > 0.0f && d2 <= lim) : (d2 <= lim);
    key = pack_key(d2, __float_as_uint(p.w));
  }
  const uint64_t m = __ballot(hit) & gmask;
  if (hit) buf[cnt + (uint32_t)__popcll(m & lanemask_lt64())] = key;
  cnt += (uint32_t)__popcll(m);
  if (cnt <= (uint32_t)(CAP - G)) return;
  // Where the buffer has room above k for it (decided at compile time, so
  // the other configurations keep their code as it was), a threshold step
  // first; the exact select below runs when that leaves the buffer more than
  // half way from k to full (ties).
  if constexpr (CAP - G >= KB + KB / 2) {
    if ((uint32_t)(CAP - G) >= k + (k >> 1)) {
      wave_sync();
      group_shrink<G>(buf, cnt, lim, k, l, gmask);
      if (cnt <= (k + (uint32_t)(CAP - G)) / 2) return;
    }
  }
  {
    wave_sync();
    for (uint32_t b0 = 0; b0 < cnt; b0 += G * kRankChunk) {
      uint64_t kv[kRankChunk];
      uint32_t rk[kRankChunk];
      rank_chunk<G, kRankChunk>(buf, cnt, b0, l, kv, rk);
#pragma unroll
      for (int i = 0; i < kRankChunk; ++i)
        if (rk[i] == k - 1) *thr_slot = kv[i];
    }
    wave_sync();
    const uint64_t T = *thr_slot;
    uint32_t kept = 0;
    for (uint32_t e0 = 0; e0 < cnt; e0 += G) {
      const uint32_t e = e0 + (uint32_t)l;
      const uint64_t key2 = e < cnt ? buf[e] : ~0ull;
      const bool keep = e < cnt && key2 <= T;
      const uint64_t km = __ballot(keep) & gmask;
      wave_sync();
      if (keep) buf[kept + (uint32_t)__popcll(km & lanemask_lt64())] = key2;
      kept += (uint32_t)__popcll(km);
      wave_sync();
    }
    cnt = k;
    lim = key_d2(T);
  }
}

// Entries each lane ranks per round in group_finish: enough that one round
// (G * CH entries) covers the KB + KB/4 + 1 a threshold step leaves, up to 12,
// else kRankChunk.
template <int G, int KB>
constexpr int finish_chunk() {
  constexpr int t = KB + KB / 4 + 1;
  constexpr int ch = (t + G - 1) / G;
  return ch <= 12 ? ch : kRankChunk;
}

// One ranking round of group_finish: entries b0 + l + G*i (i < C) of the
// buffer ranked and those of rank < k written to the row.
template <int G, int C>
__device__ inline void finish_round(const uint64_t* buf, uint32_t cnt, uint32_t b0, int l,
                                    uint32_t k, uint32_t base, int32_t* oi, float* od) {
  uint64_t kv[C];
  uint32_t rk[C];
  rank_chunk<G, C>(buf, cnt, b0, l, kv, rk);
#pragma unroll
  for (int i = 0; i < C; ++i)
    if (rk[i] < k) {
      oi[rk[i]] = (int32_t)(key_idx(kv[i]) - base);
      if (od) od[rk[i]] = key_d2(kv[i]);
    }
}

// Writes a group's row: the buffer's entries by rank (the k smallest), then
// empty slots; appends the query to the short list if it has fewer than k
// and was searched below r (a row searched at r is final: it holds every
// point within r).  A buffer beyond one ranking round is first cut by a d2
// threshold (all entries are within lim).
template <int G, int KB>
__device__ inline void group_finish(const SearchArgs& a, uint64_t* buf, uint32_t cnt,
                                    float lim, uint32_t k, int l, uint32_t row, uint32_t base,
                                    uint32_t qs, float R) {
  constexpr int CH = finish_chunk<G, KB>();
  wave_sync();
  if constexpr (G * CH >= KB + KB / 4 + 1) {
    if (cnt > (uint32_t)(G * CH) && k + (k >> 2) + 1 <= (uint32_t)(G * CH)) {
      const int lane = threadIdx.x & (kWave - 1);
      const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1ull) << (lane - l));
      group_shrink<G>(buf, cnt, lim, k, l, gmask);
      wave_sync();
    }
  }
  int32_t* oi = a.out_idx + (uint64_t)row * k;
  float* od = a.out_d2 ? a.out_d2 + (uint64_t)row * k : nullptr;
  // Ranking rounds: each ranks G * C entries (C per lane) against the whole
  // buffer, so a round wider than the entries left ranks empty slots.  The
  // round width is chosen per wave (one ballot each, so every group of the
  // wave runs the same code): the narrowest of CH, CH/2, CH/4 and 1 entries
  // per lane that covers the largest remainder in the wave.  (Range rows of
  // 33 hits in 4-lane groups took a second round of 32 slots for one entry.)
  constexpr int C2 = (CH + 1) / 2, C3 = (CH + 3) / 4;
  for (uint32_t b0 = 0;;) {
    const uint32_t rem = cnt > b0 ? cnt - b0 : 0u;
    if (__ballot(rem > (uint32_t)(G * C2))) {
      finish_round<G, CH>(buf, cnt, b0, l, k, base, oi, od);
      b0 += G * CH;
    } else if (__ballot(rem > (uint32_t)(G * C3))) {
      finish_round<G, C2>(buf, cnt, b0, l, k, base, oi, od);
      b0 += G * C2;
    } else if (__ballot(rem > (uint32_t)G)) {
      finish_round<G, C3>(buf, cnt, b0, l, k, base, oi, od);
      b0 += G * C3;
    } else if (__ballot(rem > 0u)) {
      finish_round<G, 1>(buf, cnt, b0, l, k, base, oi, od);
      b0 += G;
    } else {
      break;
    }
  }
  for (uint32_t s = cnt + (uint32_t)l; s < k; s += G) {
    oi[s] = -1;
    if (od) od[s] = __builtin_huge_valf();
  }
  if (a.short_rows) append_short(a, l == 0 && cnt < k && R < a.r, qs, cnt, R);
}

template <int G, int KB, bool KNN, bool HIER, bool WIDE>
__global__ __launch_bounds__(kTileQ) void k_topk_group(const SearchArgs a) {
  using C = GroupCfg<G, KB, WIDE>;
  __shared__ uint64_t s_buf[C::NG][C::CAP];
  // Each lane sets up kRunsPerLane runs per batch (their bounds loaded
  // together: one memory round trip for up to G * kRunsPerLane runs).
  constexpr int RB = G * kRunsPerLane;
  constexpr bool kRegRuns = RB <= 8;
  __shared__ uint32_t s_rs[kRegRuns ? 1 : C::NG][RB];
  __shared__ uint32_t s_ro[kRegRuns ? 1 : C::NG][RB + 1];
  __shared__ __align__(16) uint2 s_rd[kRegRuns ? C::NG : 1][RB];
  __shared__ uint64_t s_thr[C::NG];

  const int gq = threadIdx.x / G;
  const int l = threadIdx.x % G;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1ull) << (lane - l));
  uint64_t* buf = s_buf[gq];
  uint32_t* rs = s_rs[kRegRuns ? 0 : gq];
  uint32_t* ro = s_ro[kRegRuns ? 0 : gq];
  const uint32_t k = a.k;

  const uint32_t nq = launch_nq(a);
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const GridDesc g = a.grids[gi];
    for (uint32_t qi = q0 + gq; qi < q1; qi += C::NG) {
      const uint32_t qs = a.qsel ? a.qsel[qi] : qi;
      const float4 qv = a.queries[qs];
      const uint32_t row = __float_as_uint(qv.w);
      const float R = query_radius(a, qi, qs);
      const float R2 = R * R;
      int lo0, hi0, lo1, hi1, lo2, hi2;
      cell_range(g, qv.x, R, 0, lo0, hi0);
      cell_range(g, qv.y, R, 1, lo1, hi1);
      cell_range(g, qv.z, R, 2, lo2, hi2);
      uint32_t lvl_base = g.cell_base;
      if (HIER) {  // same level rule as k_topk_thread (group-uniform)
        int lvl = a.level_min;
        float c = ldexpf(g.cell, lvl);
        const float want = R * a.level_alpha;
        while (lvl < a.nlevels - 1 && c < want) {
          c *= 2.0f;
          ++lvl;
        }
        lo0 >>= lvl; hi0 >>= lvl; lo1 >>= lvl; hi1 >>= lvl; lo2 >>= lvl; hi2 >>= lvl;
        lvl_base = a.grids[lvl].cell_base;
      }
      const int nx = hi0 - lo0 + 1, ny = hi1 - lo1 + 1, nz = hi2 - lo2 + 1;
      // raster: one run per (y, z) row; hierarchical: one run per Morton cell
      const int nrows = (nx > 0 && ny > 0 && nz > 0) ? (HIER ? nx : 1) * ny * nz : 0;

      uint32_t cnt = 0;
      float lim = r2_limit(R2);
      for (int rb = 0; rb < nrows; rb += RB) {
        // runs rb + l * kRunsPerLane + u of this batch
        uint32_t s0[kRunsPerLane], len[kRunsPerLane];
#pragma unroll
        for (int u = 0; u < kRunsPerLane; ++u) {
          const int j = rb + l * kRunsPerLane + u;
          s0[u] = 0;
          len[u] = 0;
          if (j < nrows) {
            uint32_t b0, b1;
            if (HIER) {
              const int cx = lo0 + j % nx, t = j / nx;
              const int cy = lo1 + t % ny, cz = lo2 + t / ny;
              b0 = lvl_base + morton3((uint32_t)cx, (uint32_t)cy, (uint32_t)cz);
              b1 = b0 + 1;
            } else {
              const int cy = lo1 + j % ny, cz = lo2 + j / ny;
              const uint32_t rowc = lvl_base + ((uint32_t)cz * (uint32_t)g.dim[1] +
                                                (uint32_t)cy) * (uint32_t)g.dim[0];
              b0 = rowc + lo0;
              b1 = rowc + hi0 + 1;
            }
            s0[u] = a.cell_start[b0];
            len[u] = a.cell_start[b1];
          }
        }
        uint32_t mine = 0;
#pragma unroll
        for (int u = 0; u < kRunsPerLane; ++u) {
          len[u] -= s0[u];
          mine += len[u];
        }
        const uint32_t incl = group_scan_fast<G>(mine, l);
        const uint32_t total = group_last<G>(incl, l);
        // Run table: candidate v of the batch lies in run j = the last with
        // off_j <= v, at point index v + delta_j (delta_j = start_j - off_j,
        // mod 2^32).  Up to 8 runs it is held in VGPRs and searched with
        // selects; beyond, walked in LDS.
        uint2 rt[kRegRuns ? RB : 1];
        {
          uint32_t run = incl - mine;
#pragma unroll
          for (int u = 0; u < kRunsPerLane; ++u) {
            if constexpr (kRegRuns) {
              s_rd[gq][l * kRunsPerLane + u] = make_uint2(run, s0[u] - run);
            } else {
              rs[l * kRunsPerLane + u] = s0[u];
              ro[l * kRunsPerLane + u] = run;
            }
            run += len[u];
          }
          if constexpr (!kRegRuns) {
            if (l == G - 1) ro[RB] = total;
          }
        }
        wave_sync();
        if constexpr (kRegRuns) {
          const uint4* t4 = reinterpret_cast<const uint4*>(s_rd[gq]);
#pragma unroll
          for (int j = 0; j < RB / 2; ++j) {
            const uint4 w = t4[j];
            rt[2 * j] = make_uint2(w.x, w.y);
            rt[2 * j + 1] = make_uint2(w.z, w.w);
          }
        }
        int rp = 0;
        for (uint32_t v0 = 0; v0 < total; v0 += G * kGroupUnroll) {
          // Positions of kGroupUnroll candidates per lane first, then their
          // loads as one straight-line batch with no branch in between, so
          // all of them are in flight before the first test.
          float4 pv[kGroupUnroll];
          if constexpr (kRegRuns) {
            // Candidate u of lane l is v = vl + G*u.  Run j (offset off_j,
            // first point s_j) would hold v = vl at pb[j] = spts + s_j +
            // (vl - off_j), 64-bit (vl - off_j may be negative); candidate u
            // adds G*u points, a constant that goes into the load's offset
            // field.  Its run is the last j with G*u >= off_j - vl, so the
            // address read is s_j + (v - off_j) >= s_j.  A slot past the end
            // reads at most G*kGroupUnroll points beyond some run's end (the
            // buffer's padding covers it) and is not a hit.
            const uint32_t vl = v0 + (uint32_t)l;
            const float4* pb[RB];
            int tj[RB];
#pragma unroll
            for (int j = 0; j < RB; ++j) {
              tj[j] = (int)(rt[j].x - vl);
              pb[j] = a.spts + (uint64_t)(rt[j].y + rt[j].x) - (int64_t)tj[j];
            }
#pragma unroll
            for (int u = 0; u < kGroupUnroll; ++u) {
              const float4* pp = pb[0];
#pragma unroll
              for (int j = 1; j < RB; ++j) pp = G * u >= tj[j] ? pb[j] : pp;
              pv[u] = pp[G * u];
            }
          } else {
            uint32_t pos[kGroupUnroll];
#pragma unroll
            for (int u = 0; u < kGroupUnroll; ++u) {
              // A slot past the end loads position 0 and is not a hit.
              const uint32_t v = v0 + (uint32_t)(u * G + l);
              pos[u] = 0;
              if (v < total) {
                while (ro[rp + 1] <= v) ++rp;
                pos[u] = rs[rp] + (v - ro[rp]);
              }
            }
#pragma unroll
            for (int u = 0; u < kGroupUnroll; ++u) pv[u] = a.spts[pos[u]];
          }
          // Hits of this lane as a bit mask, then one group scan of the
          // counts places them; the buffer keeps CAP - G free of this batch
          // (the entry condition of group_take).  A batch that does not fit
          // goes through group_take, one candidate at a time.
          uint64_t key[kGroupUnroll];
          uint32_t bits = 0;
          const int tt = (int)(total - v0 - (uint32_t)l);  // valid while G*u < tt
#pragma unroll
          for (int u = 0; u < kGroupUnroll; ++u) {
            const float d2 = d2_of(qv.x, qv.y, qv.z, pv[u].x, pv[u].y, pv[u].z);
            const bool ok = (KNN ? (d2 > 0.0f && d2 <= lim) : (d2 <= lim)) && G * u < tt;
            key[u] = pack_key(d2, __float_as_uint(pv[u].w));
            bits |= ok ? 1u << u : 0u;
          }
          const uint32_t nh = (uint32_t)__popc(bits);
          const uint32_t hincl = group_scan_fast<G>(nh, l);
          const uint32_t htot = group_last<G>(hincl, l);
          if (cnt + htot <= (uint32_t)(C::CAP - G)) {
            uint32_t o = cnt + hincl - nh;
#pragma unroll
            for (int u = 0; u < kGroupUnroll; ++u)
              if (bits & (1u << u)) buf[o++] = key[u];
            cnt += htot;
          } else {
#pragma unroll
            for (int u = 0; u < kGroupUnroll; ++u) {
              const uint32_t v = v0 + (uint32_t)(u * G + l);
              group_take<G, KB, C::CAP, KNN>(pv[u], v < total, qv, gmask, buf, cnt, lim, k, l,
                                         &s_thr[gq]);
            }
          }
        }
        wave_sync();
      }

      group_finish<G, KB>(a, buf, cnt, lim, k, l, row, a.local_index ? g.point_base : 0u, qs, R);
      wave_sync();
    }
  }
}

// --------------------------------------------------------------------------
// The candidates of a wave's queries as one pool
//
// k_topk_group gives each query G lanes and walks its runs batch by batch, so
// a wave takes as long as its slowest query and most lanes test nothing at a
// time: queries differ in row count and run lengths.  Here a wave lists the
// runs of its QW queries in LDS first -- one row of cells per run, each row
// clipped to the cells its query's sphere reaches -- and numbers their points
// as one pool.  Lane l tests pool slots l, l + 64, ...: consecutive lanes
// read consecutive points of a run (coalesced loads), each slot against the
// query its run belongs to, and a hit goes to that query's LDS buffer
// (CAPQ keys, filled by an LDS atomic).  The G = 64 / QW lanes of a query
// then rank and write its row as k_topk_group does.  A query with more hits
// than its buffer holds appends its position to ovf_rows; the group kernel
// searches those rows again.  Raster grid, one grid per launch.

constexpr int kPoolRows = 256;    // runs listed per pass (an index fits a byte)
constexpr int kPoolSlots = 1024;  // pool slots mapped to runs per window
constexpr int kPoolUnroll = 4;    // slots per lane per step, loads in flight

template <int QW, int CAPQ>
struct PoolLds {
  float4 q[QW];                    // query xyz and d2 limit
  uint32_t cnt[QW];                // hits found (above CAPQ: overflow)
  uint32_t rcell[kPoolRows];       // row: first cell; then run: first point - pool offset
  uint32_t rlen[kPoolRows];        // row: cells
  uint8_t rq[kPoolRows];           // query of each row
  __align__(16) uint8_t map[kPoolSlots];  // run of each slot of the current window
  // CAPQ | 1 keys per query: an odd stride puts the queries' buffers on
  // different LDS banks (the finish's rank loop reads entry j of every
  // query's buffer at once)
  uint64_t buf[QW][CAPQ | 1];
};

// Cells [clo, chi] of row (cy, cz) that can hold a point within R of q: the
// row's y and z slabs, padded for the FP32 cell mapping (a point's cell index
// can round across a boundary), bound dy and dz from below, so a hit has
// |dx| <= sqrt(R'^2 - dy^2 - dz^2) with R' = R (1 + 2^-16) covering the FP32
// d2 and these bounds' rounding; cell_range then maps that half-width.  The
// border rows reach to infinity (cell indices are clamped into the grid).
__device__ inline void clip_row(const GridDesc& g, const float4& q, float R, int cy, int cz,
                                int lo0, int hi0, int& clo, int& chi) {
  const float c = g.cell;
  const float mag = fabsf(g.mn[1]) + fabsf(g.mn[2]) + fabsf(q.y) + fabsf(q.z) +
                    (float)(max(cy, cz) + 1) * c;
  const float pad = c * 0x1.0p-12f + mag * 0x1.0p-19f;
  const float inf = __builtin_huge_valf();
  const float ylo = cy == 0 ? -inf : g.mn[1] + (float)cy * c - pad;
  const float yhi = cy == g.dim[1] - 1 ? inf : g.mn[1] + (float)(cy + 1) * c + pad;
  const float zlo = cz == 0 ? -inf : g.mn[2] + (float)cz * c - pad;
  const float zhi = cz == g.dim[2] - 1 ? inf : g.mn[2] + (float)(cz + 1) * c + pad;
  const float dy = fmaxf(fmaxf(ylo - q.y, q.y - yhi), 0.0f);
  const float dz = fmaxf(fmaxf(zlo - q.z, q.z - zhi), 0.0f);
  const float Rp = R * (1.0f + 0x1.0p-16f);
  const float h2 = Rp * Rp - (dy * dy + dz * dz);
  if (!(h2 >= 0.0f)) {
    clo = 1;
    chi = 0;
    return;
  }
  int a0, b0;
  cell_range(g, q.x, sqrtf(h2), 0, a0, b0);
  clo = max(a0, lo0);
  chi = min(b0, hi0);
}

template <int QW, int KB, int CAPQ, bool KNN>
__global__ __launch_bounds__(kTileQ) void k_topk_pool(const SearchArgs a) {
  constexpr int G = kWave / QW;
  constexpr int kWaves = kTileQ / kWave;
  static_assert(4 * kWave == kPoolRows, "four rows per lane");
  using L = PoolLds<QW, CAPQ>;
  __shared__ L s_pool[kWaves];
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int gq = lane / G, l = lane % G;
  L& W = s_pool[w];
  const uint32_t k = a.k;
  const uint32_t nq = launch_nq(a);
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const GridDesc g = a.grids[gi];
    for (uint32_t c0 = q0 + (uint32_t)(w * QW); c0 < q1; c0 += (uint32_t)(kWaves * QW)) {
      // The chunk's queries, G lanes each.
      const uint32_t qi = c0 + (uint32_t)gq;
      const bool active = qi < q1;
      uint32_t qs = 0, row = 0;
      float4 qv = make_float4(0.f, 0.f, 0.f, 0.f);
      float R = a.r, lim = -1.0f;
      int lo0 = 0, hi0 = -1, lo1 = 0, hi1 = -1, lo2 = 0, hi2 = -1;
      if (active) {
        qs = a.qsel ? a.qsel[qi] : qi;
        qv = a.queries[qs];
        row = __float_as_uint(qv.w);
        R = query_radius(a, qi, qs);
        lim = r2_limit(R * R);
        cell_range(g, qv.x, R, 0, lo0, hi0);
        cell_range(g, qv.y, R, 1, lo1, hi1);
        cell_range(g, qv.z, R, 2, lo2, hi2);
      }
      const int ny = hi1 - lo1 + 1, nz = hi2 - lo2 + 1;
      const uint32_t nrows =
          (active && hi0 >= lo0 && ny > 0 && nz > 0) ? (uint32_t)(ny * nz) : 0u;
      wave_sync();  // the previous chunk's finish has read the buffers
      if (l == 0) {
        W.q[gq] = make_float4(qv.x, qv.y, qv.z, lim);
        W.cnt[gq] = 0;
      }
      // rows of the chunk numbered query by query
      const uint32_t rv = l == 0 ? nrows : 0u;
      const uint32_t rincl = group_incl_scan<kWave>(rv, lane);
      const uint32_t total_rows = __shfl(rincl, kWave - 1, kWave);
      const uint32_t rbase = __shfl(rincl - rv, lane - l, kWave);
      for (uint32_t rw = 0; rw < total_rows; rw += kPoolRows) {
        const uint32_t nr = min((uint32_t)kPoolRows, total_rows - rw);
        // each query lists its rows of this pass (row t: cy = lo1 + t % ny,
        // cz = lo2 + t / ny), clipped to its sphere
        if (nrows) {
          int cy = lo1 + l, cz = lo2;
          while (cy > hi1) {
            cy -= ny;
            ++cz;
          }
          for (uint32_t t = (uint32_t)l; t < nrows; t += G) {
            const uint32_t j = rbase + t - rw;
            if (rbase + t >= rw && j < nr) {
              int clo, chi;
              clip_row(g, qv, R, cy, cz, lo0, hi0, clo, chi);
              const uint32_t rowc = g.cell_base + ((uint32_t)cz * (uint32_t)g.dim[1] +
                                                   (uint32_t)cy) * (uint32_t)g.dim[0];
              W.rcell[j] = rowc + (uint32_t)max(clo, 0);
              W.rlen[j] = chi >= clo ? (uint32_t)(chi - clo + 1) : 0u;
              W.rq[j] = (uint8_t)gq;
            }
            cy += G;
            while (cy > hi1) {
              cy -= ny;
              ++cz;
            }
          }
        }
        wave_sync();
        // runs: lane takes rows 4 lane .. 4 lane + 3; their points are
        // numbered in row order (pool offsets by a wave scan)
        uint32_t st[4], ln[4], own = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const uint32_t j = (uint32_t)(4 * lane + u);
          uint32_t s0 = 0, e0 = 0;
          if (j < nr) {
            const uint32_t c = W.rcell[j], n = W.rlen[j];
            if (n) {
              s0 = a.cell_start[c];
              e0 = a.cell_start[c + n];
            }
          }
          st[u] = s0;
          ln[u] = e0 - s0;
          own += e0 - s0;
        }
        const uint32_t pincl = group_incl_scan<kWave>(own, lane);
        const uint32_t P = __shfl(pincl, kWave - 1, kWave);
        uint32_t of[4];
        {
          uint32_t o = pincl - own;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            of[u] = o;
            const uint32_t j = (uint32_t)(4 * lane + u);
            if (j < nr) W.rcell[j] = st[u] - o;  // point of slot v = this + v
            o += ln[u];
          }
        }
        uint32_t wcarry = 0;  // run of the last slot of the previous window
        for (uint32_t pw = 0; pw < P; pw += kPoolSlots) {
          const uint32_t wend = min(P, pw + kPoolSlots);
          // slot -> run map of the window [pw, wend): run indices rise with
          // the slot, so the run of a slot is the largest index of a run that
          // starts at or before it: zero the window, mark each run's first
          // slot with its index, then a prefix maximum (16 slots per lane and
          // a wave scan, carried over from the previous window).
          reinterpret_cast<uint4*>(W.map)[lane] = make_uint4(0u, 0u, 0u, 0u);
          wave_sync();
#pragma unroll
          for (int u = 0; u < 4; ++u)
            if (ln[u] && of[u] >= pw && of[u] < wend) W.map[of[u] - pw] = (uint8_t)(4 * lane + u);
          wave_sync();
          {
            uint4 m = reinterpret_cast<const uint4*>(W.map)[lane];
            uint32_t wv[4] = {m.x, m.y, m.z, m.w};
            uint32_t run_max = 0;
#pragma unroll
            for (int b = 0; b < 16; ++b) run_max = max(run_max, (wv[b >> 2] >> (8 * (b & 3))) & 0xffu);
            uint32_t incl = run_max;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
              const uint32_t t = __shfl_up(incl, o, kWave);
              if (lane >= o) incl = max(incl, t);
            }
            uint32_t cur = max(wcarry, __shfl_up(incl, 1, kWave));
            if (lane == 0) cur = wcarry;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              uint32_t o = 0;
#pragma unroll
              for (int b = 0; b < 4; ++b) {
                cur = max(cur, (wv[i] >> (8 * b)) & 0xffu);
                o |= cur << (8 * b);
              }
              wv[i] = o;
            }
            reinterpret_cast<uint4*>(W.map)[lane] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
            wcarry = max(wcarry, __shfl(incl, kWave - 1, kWave));
          }
          wave_sync();
          for (uint32_t v0 = pw; v0 < wend; v0 += (uint32_t)(kWave * kPoolUnroll)) {
            // In phases, each over every slot of the step, with no branch
            // between reads (indices masked into range), so the LDS reads,
            // then the point loads, of all slots are in flight together.
            uint32_t rid[kPoolUnroll], qq[kPoolUnroll], pi[kPoolUnroll];
            bool ok[kPoolUnroll];
#pragma unroll
            for (int u = 0; u < kPoolUnroll; ++u) {
              const uint32_t v = v0 + (uint32_t)(u * kWave + lane);
              ok[u] = v < wend;
              rid[u] = W.map[(v - pw) & (uint32_t)(kPoolSlots - 1)];
            }
#pragma unroll
            for (int u = 0; u < kPoolUnroll; ++u) {
              const uint32_t v = v0 + (uint32_t)(u * kWave + lane);
              qq[u] = (uint32_t)W.rq[rid[u]] & (uint32_t)(QW - 1);
              pi[u] = ok[u] ? W.rcell[rid[u]] + v : 0u;
            }
            float4 pv[kPoolUnroll], qd[kPoolUnroll];
#pragma unroll
            for (int u = 0; u < kPoolUnroll; ++u) {
              pv[u] = a.spts[pi[u]];
              qd[u] = W.q[qq[u]];
            }
            uint32_t at[kPoolUnroll];
            uint64_t key[kPoolUnroll];
#pragma unroll
            for (int u = 0; u < kPoolUnroll; ++u) {
              const float d2 = d2_of(qd[u].x, qd[u].y, qd[u].z, pv[u].x, pv[u].y, pv[u].z);
              const bool hit = ok[u] && (KNN ? (d2 > 0.0f && d2 <= qd[u].w) : (d2 <= qd[u].w));
              key[u] = pack_key(d2, __float_as_uint(pv[u].w));
              at[u] = hit ? atomicAdd(&W.cnt[qq[u]], 1u) : (uint32_t)CAPQ;
            }
#pragma unroll
            for (int u = 0; u < kPoolUnroll; ++u)
              if (at[u] < (uint32_t)CAPQ) W.buf[qq[u]][at[u]] = key[u];
          }
          wave_sync();
        }
      }
      wave_sync();
      const uint32_t cnt = active ? W.cnt[gq] : 0u;
      // Over-full buffers: the row is searched again by the group kernel.
      {
        const bool ov = active && l == 0 && cnt > (uint32_t)CAPQ;
        const uint64_t m = __ballot(ov);
        if (m) {
          const int leader = __ffsll((unsigned long long)m) - 1;
          uint32_t b = 0;
          if (lane == leader) b = atomicAdd(a.ovf_count, (uint32_t)__popcll(m));
          b = __shfl(b, leader, kWave);
          if (ov) a.ovf_rows[b + (uint32_t)__popcll(m & lanemask_lt64())] = qs;
        }
      }
      if (active && cnt <= (uint32_t)CAPQ)
        group_finish<G, KB>(a, W.buf[gq], cnt, lim, k, l, row,
                            a.local_index ? g.point_base : 0u, qs, R);
    }
  }
}

// --------------------------------------------------------------------------
// G lanes per query, candidates staged per wave in LDS
//
// The 64/G queries of a wave are consecutive in Morton order, so their cell
// blocks overlap.  The wave loads the points of the union of those blocks into
// LDS once (one coalesced pass, 64 loads in flight) and every group scans the
// tile with LDS reads.  A point of the union block outside a query's own block
// is tested too and fails d2 < R^2: every point within R of the query lies in
// its own block, each point appears once in the tile, so no hit is missed or
// counted twice.  A wave whose union block spans more than 64 rows or holds
// more than TP points (the tile: 256, 512 or 1024 per wave) walks its rows
// from global memory instead (the k_topk_group path).  Raster grids, compact
// group buffers.

__device__ inline int wave_min_i(int v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, kWave));
  return v;
}
__device__ inline int wave_max_i(int v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, kWave));
  return v;
}

template <int G, int KB, bool KNN, int TP>
__global__ __launch_bounds__(kTileQ) void k_topk_tiled(const SearchArgs a) {
  using C = GroupCfg<G, KB, false>;
  constexpr int kWaves = kTileQ / kWave;
  __shared__ uint64_t s_buf[C::NG][C::CAP];
  __shared__ uint32_t s_rs[C::NG][G];
  __shared__ uint32_t s_ro[C::NG][G + 1];
  __shared__ uint64_t s_thr[C::NG];
  __shared__ float4 s_tile[kWaves][TP];
  __shared__ uint32_t s_trs[kWaves][kWave];
  __shared__ uint32_t s_tro[kWaves][kWave + 1];

  const int gq = threadIdx.x / G;
  const int l = threadIdx.x % G;
  const int w = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const uint64_t gmask = G == kWave ? ~0ull : (((1ull << G) - 1ull) << (lane - l));
  uint64_t* buf = s_buf[gq];
  uint32_t* rs = s_rs[gq];
  uint32_t* ro = s_ro[gq];
  float4* tile = s_tile[w];
  uint32_t* trs = s_trs[w];
  uint32_t* tro = s_tro[w];
  const uint32_t k = a.k;

  const uint32_t nq = launch_nq(a);
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const GridDesc g = a.grids[gi];
    // Every group of the block runs the same number of rounds (the wave's
    // union and tile are collective); groups past q1 only take part in them.
    for (uint32_t qb = q0; qb < q1; qb += C::NG) {
      const uint32_t qi = qb + (uint32_t)gq;
      const bool active = qi < q1;
      uint32_t qs = 0, row = 0;
      float4 qv = make_float4(0.f, 0.f, 0.f, 0.f);
      float R = a.r, R2 = 0.0f;
      int lo0 = 0, hi0 = -1, lo1 = 0, hi1 = -1, lo2 = 0, hi2 = -1;
      if (active) {
        qs = a.qsel ? a.qsel[qi] : qi;
        qv = a.queries[qs];
        row = __float_as_uint(qv.w);
        R = query_radius(a, qi, qs);
        R2 = R * R;
        cell_range(g, qv.x, R, 0, lo0, hi0);
        cell_range(g, qv.y, R, 1, lo1, hi1);
        cell_range(g, qv.z, R, 2, lo2, hi2);
      }
      const int nx = hi0 - lo0 + 1, ny = hi1 - lo1 + 1, nz = hi2 - lo2 + 1;
      const bool mine = nx > 0 && ny > 0 && nz > 0;  // a non-empty block
      // union block of the wave's non-empty blocks (wave-uniform)
      const int ulo0 = wave_min_i(mine ? lo0 : INT_MAX), uhi0 = wave_max_i(mine ? hi0 : INT_MIN);
      const int ulo1 = wave_min_i(mine ? lo1 : INT_MAX), uhi1 = wave_max_i(mine ? hi1 : INT_MIN);
      const int ulo2 = wave_min_i(mine ? lo2 : INT_MAX), uhi2 = wave_max_i(mine ? hi2 : INT_MIN);
      const bool any = uhi0 >= ulo0;  // some block is non-empty (then all axes are)
      const int uny = any ? uhi1 - ulo1 + 1 : 0, unz = any ? uhi2 - ulo2 + 1 : 0;
      const int urows = uny * unz;
      bool tiled = urows <= kWave;
      uint32_t T = 0;
      if (tiled && urows > 0) {
        uint32_t s0 = 0, len = 0;
        if (lane < urows) {
          const int cy = ulo1 + lane % uny, cz = ulo2 + lane / uny;
          const uint32_t rowc = g.cell_base + ((uint32_t)cz * (uint32_t)g.dim[1] +
                                               (uint32_t)cy) * (uint32_t)g.dim[0];
          s0 = a.cell_start[rowc + ulo0];
          len = a.cell_start[rowc + uhi0 + 1] - s0;
        }
        const uint32_t incl = group_incl_scan<kWave>(len, lane);
        T = __shfl(incl, kWave - 1, kWave);
        tiled = T <= (uint32_t)TP;
        if (tiled) {
          trs[lane] = s0;
          tro[lane + 1] = incl;
          if (lane == 0) tro[0] = 0;
          wave_sync();
          int rp = 0;
          for (uint32_t v = (uint32_t)lane; v < T; v += kWave) {
            while (tro[rp + 1] <= v) ++rp;
            tile[v] = a.spts[trs[rp] + (v - tro[rp])];
          }
          wave_sync();
        }
      }

      uint32_t cnt = 0;
      float lim = r2_limit(R2);
      if (active && tiled) {
        for (uint32_t v0 = 0; v0 < T; v0 += G) {
          const uint32_t v = v0 + (uint32_t)l;
          const float4 p = v < T ? tile[v] : make_float4(0.f, 0.f, 0.f, 0.f);
          group_take<G, KB, C::CAP, KNN>(p, v < T, qv, gmask, buf, cnt, lim, k, l, &s_thr[gq]);
        }
      } else if (active && mine) {
        const int nrows = ny * nz;
        for (int rb = 0; rb < nrows; rb += G) {
          const int j = rb + l;
          uint32_t s0 = 0, len = 0;
          if (j < nrows) {
            const int cy = lo1 + j % ny, cz = lo2 + j / ny;
            const uint32_t rowc = g.cell_base + ((uint32_t)cz * (uint32_t)g.dim[1] +
                                                 (uint32_t)cy) * (uint32_t)g.dim[0];
            s0 = a.cell_start[rowc + lo0];
            len = a.cell_start[rowc + hi0 + 1] - s0;
          }
          const uint32_t incl = group_incl_scan<G>(len, l);
          const uint32_t total = __shfl(incl, G - 1, G);
          rs[l] = s0;
          ro[l + 1] = incl;
          if (l == 0) ro[0] = 0;
          wave_sync();
          int rp = 0;
          for (uint32_t v0 = 0; v0 < total; v0 += G) {
            const uint32_t v = v0 + (uint32_t)l;
            float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
            if (v < total) {
              while (ro[rp + 1] <= v) ++rp;
              p = a.spts[rs[rp] + (v - ro[rp])];
            }
            group_take<G, KB, C::CAP, KNN>(p, v < total, qv, gmask, buf, cnt, lim, k, l,
                                       &s_thr[gq]);
          }
          wave_sync();
        }
      }
      if (active)
        group_finish<G, KB>(a, buf, cnt, lim, k, l, row, a.local_index ? g.point_base : 0u, qs,
                            R);
      wave_sync();  // the tile and buffers are reused by the next round
    }
  }
}

// --------------------------------------------------------------------------
// dispatch

template <int KC>
static void launch_thread(const SearchArgs& a, bool knn, dim3 grid, hipStream_t st) {
  const bool hier = a.nlevels > 0;
  if (knn && hier && a.tree_lanes == 2)
    hipLaunchKernelGGL((k_topk_treeg<KC, 2>), grid, dim3(kTileQ), 0, st, a);
  else if (knn && hier && a.tree_lanes == 4)
    hipLaunchKernelGGL((k_topk_treeg<KC, 4>), grid, dim3(kTileQ), 0, st, a);
  else if (knn && hier && a.tree_lanes == 8)
    hipLaunchKernelGGL((k_topk_treeg<KC, 8>), grid, dim3(kTileQ), 0, st, a);
  else if (knn && hier)
    hipLaunchKernelGGL((k_topk_tree<KC>), grid, dim3(kTileQ), 0, st, a);
  else if (knn)
    hipLaunchKernelGGL((k_topk_thread<KC, true, false>), grid, dim3(kTileQ), 0, st, a);
  else if (hier)
    hipLaunchKernelGGL((k_topk_thread<KC, false, true>), grid, dim3(kTileQ), 0, st, a);
  else
    hipLaunchKernelGGL((k_topk_thread<KC, false, false>), grid, dim3(kTileQ), 0, st, a);
}

template <int G, int KB, int TP>
static void launch_tiled(const SearchArgs& a, bool knn, dim3 grid, hipStream_t st) {
  if (knn)
    hipLaunchKernelGGL((k_topk_tiled<G, KB, true, TP>), grid, dim3(kTileQ), 0, st, a);
  else
    hipLaunchKernelGGL((k_topk_tiled<G, KB, false, TP>), grid, dim3(kTileQ), 0, st, a);
}

template <int G, int KB>
static void launch_group(const SearchArgs& a, bool knn, bool wide, int tile, dim3 grid,
                         hipStream_t st) {
  const bool hier = a.nlevels > 0;
  if constexpr (G >= 4) {  // 2-lane groups: no room for a tile beside their buffers
    if (tile > 0 && !hier) {
      if (tile <= 256) launch_tiled<G, KB, 256>(a, knn, grid, st);
      else if (tile <= 512) launch_tiled<G, KB, 512>(a, knn, grid, st);
      else launch_tiled<G, KB, 1024>(a, knn, grid, st);
      return;
    }
  }
  if (knn && hier)
    hipLaunchKernelGGL((k_topk_group<G, KB, true, true, true>), grid, dim3(kTileQ), 0, st, a);
  else if (knn && wide)
    hipLaunchKernelGGL((k_topk_group<G, KB, true, false, true>), grid, dim3(kTileQ), 0, st, a);
  else if (knn)
    hipLaunchKernelGGL((k_topk_group<G, KB, true, false, false>), grid, dim3(kTileQ), 0, st, a);
  else if (hier)
    hipLaunchKernelGGL((k_topk_group<G, KB, false, true, true>), grid, dim3(kTileQ), 0, st, a);
  else if (wide)
    hipLaunchKernelGGL((k_topk_group<G, KB, false, false, true>), grid, dim3(kTileQ), 0, st, a);
  else
    hipLaunchKernelGGL((k_topk_group<G, KB, false, false, false>), grid, dim3(kTileQ), 0, st, a);
}

template <int QW, int KB>
static void launch_pool_kb(const SearchArgs& a, bool knn, dim3 grid, hipStream_t st) {
  if (knn)
    hipLaunchKernelGGL((k_topk_pool<QW, KB, 2 * KB, true>), grid, dim3(kTileQ), 0, st, a);
  else
    hipLaunchKernelGGL((k_topk_pool<QW, KB, 2 * KB, false>), grid, dim3(kTileQ), 0, st, a);
}

template <int QW>
static void launch_pool(const SearchArgs& a, bool knn, dim3 grid, hipStream_t st) {
  if (a.k <= 16) launch_pool_kb<QW, 16>(a, knn, grid, st);
  else if constexpr (QW <= 32) {
    if (a.k <= 32) launch_pool_kb<QW, 32>(a, knn, grid, st);
    else if constexpr (QW <= 16) launch_pool_kb<QW, 64>(a, knn, grid, st);
  }
}

uint32_t max_k_for_group(int group_lanes) {
  switch (group_lanes) {
    case 1: return 64;
    case 2: return 16;
    case 4: return 64;
    case 8: return 64;
    case 16: return 256;
    case 64: return 1024;
    default: return 0;
  }
}

hipError_t launch_search(const SearchArgs& a0, int knn_mode, int group_lanes, hipStream_t st,
                         int num_cus, bool wide, int tiled, bool deep, int pool) {
  SearchArgs a = a0;
  if (a.nq_dev) a.persistent = 1;  // tile count known only on the device
  const uint32_t k = a.k;
  if (pool) {
    // The pooled kernel, then the group kernel over the rows whose buffers
    // overflowed (their count is read on the device).
    if ((pool != 16 && pool != 32 && pool != 64) || k == 0 || k > 64 || a.nlevels > 0 ||
        a.tiles || a.qsel_hits || !a.ovf_rows || !a.ovf_count)
      return hipErrorInvalidValue;
    // Hit buffers of 2 KB keys for QW queries: at most 16 KB of LDS per wave
    // (QW * KB <= 1024), so wider rows take fewer queries per wave.
    const int kb = k <= 16 ? 16 : k <= 32 ? 32 : 64;
    pool = pool * kb > 1024 ? 1024 / kb : pool;
    a.tile_q = (uint32_t)(4 * pool);
    a.ntiles = (uint32_t)(((uint64_t)a.nq + a.tile_q - 1) / a.tile_q);
    if (a.nq_dev) a.ntiles = 0;
    if (a.ntiles || a.nq_dev) {
      uint32_t blocks = a.ntiles;
      if (a.persistent || a.nq_dev) blocks = (uint32_t)(num_cus > 0 ? num_cus : 256) * 8u;
      const dim3 grid(blocks);
      const bool knn = knn_mode != 0;
      if (pool == 16) launch_pool<16>(a, knn, grid, st);
      else if (pool == 32) launch_pool<32>(a, knn, grid, st);
      else launch_pool<64>(a, knn, grid, st);
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    SearchArgs o = a0;
    o.qsel = a0.ovf_rows;
    o.nq_dev = a0.ovf_count;
    o.ovf_rows = nullptr;
    o.ovf_count = nullptr;
    return launch_search(o, knn_mode, 4, st, num_cus, true, 0, true, 0);
  }
  if (k == 0 || k > max_k_for_group(group_lanes)) return hipErrorInvalidValue;
  if (!a.tiles) {
    // Implicit tiles: a block of the thread kernel takes kTileQ queries, a
    // block of the group kernel one query per group, so small workloads still
    // spread over many blocks.
    // the octree walk with several lanes per query takes kTileQ / lanes
    // queries per block, as a group kernel does
    const bool tree_g = group_lanes == 1 && knn_mode != 0 && a.nlevels > 0 &&
                        (a.tree_lanes == 2 || a.tree_lanes == 4 || a.tree_lanes == 8);
    a.tile_q = tree_g ? (uint32_t)(kTileQ / a.tree_lanes)
                      : group_lanes == 1 ? (uint32_t)kTileQ : (uint32_t)(kTileQ / group_lanes);
    a.ntiles = (uint32_t)(((uint64_t)a.nq + a.tile_q - 1) / a.tile_q);
  }
  if (a.ntiles == 0) return hipSuccess;
  uint32_t blocks = a.ntiles;
  if (a.persistent) {
    uint32_t cap = (uint32_t)(num_cus > 0 ? num_cus : 256) * 8u;
    if (blocks > cap) blocks = cap;
  }
  const dim3 grid(blocks);
  const bool knn = knn_mode != 0;
  if (group_lanes == 1) {
    if (k <= 1) launch_thread<1>(a, knn, grid, st);
    else if (k <= 2) launch_thread<2>(a, knn, grid, st);
    else if (k <= 4) launch_thread<4>(a, knn, grid, st);
    else if (k <= 8) launch_thread<8>(a, knn, grid, st);
    else if (k <= 16) launch_thread<16>(a, knn, grid, st);
    else if (k <= 32) launch_thread<32>(a, knn, grid, st);
    else launch_thread<64>(a, knn, grid, st);
  } else if (group_lanes == 2) {
    launch_group<2, 16>(a, knn, wide, tiled, grid, st);
  } else if (group_lanes == 4) {
    if (k <= 16 && !deep) launch_group<4, 16>(a, knn, wide, tiled, grid, st);
    else if (k <= 32) launch_group<4, 32>(a, knn, wide, tiled, grid, st);
    else launch_group<4, 64>(a, knn, wide, tiled, grid, st);
  } else if (group_lanes == 8) {
    if (k <= 16) launch_group<8, 16>(a, knn, wide, tiled, grid, st);
    else launch_group<8, 64>(a, knn, wide, tiled, grid, st);
  } else if (group_lanes == 16) {
    if (k <= 16) launch_group<16, 16>(a, knn, wide, tiled, grid, st);
    else if (k <= 64) launch_group<16, 64>(a, knn, wide, tiled, grid, st);
    else launch_group<16, 256>(a, knn, wide, tiled, grid, st);
  } else {
    if (k <= 16) launch_group<64, 16>(a, knn, wide, tiled, grid, st);
    else if (k <= 64) launch_group<64, 64>(a, knn, wide, tiled, grid, st);
    else if (k <= 256) launch_group<64, 256>(a, knn, wide, tiled, grid, st);
    else launch_group<64, 1024>(a, knn, wide, tiled, grid, st);
  }
  return hipGetLastError();
}

}  // namespace rtnn

namespace rtnn {

// --------------------------------------------------------------------------
// boundary-pair report

// Pairs (query row, point index) whose distance, evaluated in FP64 from the
// FP32 coordinates, lies within tol * r of r; these are the pairs whose
// membership a 1-ulp difference in the FP32 d2 could flip.  Count pass
// (offs == nullptr): counts[row] = pairs of that row.  Write pass: the pairs
// of a row go to pairs[2 * (offs[row] + i)] in (row, point) order, positions
// at or past cap are dropped.  Thread per query; any tile layout of
// SearchArgs (single grid or batched clouds) works.
__global__ __launch_bounds__(kTileQ) void k_boundary(const SearchArgs a, double tol,
                                                     uint32_t* __restrict__ counts,
                                                     const uint32_t* __restrict__ offs,
                                                     uint64_t cap, uint32_t* __restrict__ pairs) {
  const uint32_t nq = launch_nq(a);
  for (TileIter it(a, nq); it.t < it.end; it.t += it.step) {
    uint32_t q0, q1, gi;
    tile_of(a, nq, it.t, q0, q1, gi);
    const uint32_t qi = q0 + threadIdx.x;
    if (qi >= q1) continue;
    const uint32_t qs = a.qsel ? a.qsel[qi] : qi;
    const GridDesc g = a.grids[gi];
    const float4 qv = a.queries[qs];
    const uint32_t row = __float_as_uint(qv.w);
    const double rr = (double)a.r, lim = tol * rr;
    const float R = (float)(rr + lim);  // cell_range's padding covers the rounding
    int lo0, hi0, lo1, hi1, lo2, hi2;
    cell_range(g, qv.x, R, 0, lo0, hi0);
    cell_range(g, qv.y, R, 1, lo1, hi1);
    cell_range(g, qv.z, R, 2, lo2, hi2);
    const uint32_t base = a.local_index ? g.point_base : 0u;
    const uint64_t o = offs ? offs[row] : 0;
    uint32_t c = 0;
    for (int cz = lo2; cz <= hi2; ++cz)
      for (int cy = lo1; cy <= hi1; ++cy) {
        if (lo0 > hi0) continue;
        const uint32_t rowc = g.cell_base + ((uint32_t)cz * (uint32_t)g.dim[1] + (uint32_t)cy) *
                                                (uint32_t)g.dim[0];
        const uint32_t s = a.cell_start[rowc + lo0], e = a.cell_start[rowc + hi0 + 1];
        for (uint32_t j = s; j < e; ++j) {
          const float4 p = a.spts[j];
          const double dx = (double)qv.x - (double)p.x;
          const double dy = (double)qv.y - (double)p.y;
          const double dz = (double)qv.z - (double)p.z;
          const double d = sqrt(dx * dx + dy * dy + dz * dz);
          if (fabs(d - rr) <= lim) {
            if (offs && o + c < cap) {
              pairs[2 * (o + c)] = row;
              pairs[2 * (o + c) + 1] = __float_as_uint(p.w) - base;
            }
            ++c;
          }
        }
      }
    if (!offs) {
      counts[row] = c;
      continue;
    }
    // The row's pairs arrive in cell order; sort them by point index.
    const uint64_t end = o + c < cap ? o + c : cap;
    for (uint64_t i = o + 1; i < end; ++i) {
      const uint32_t v = pairs[2 * i + 1];
      uint64_t j = i;
      for (; j > o && pairs[2 * (j - 1) + 1] > v; --j) pairs[2 * j + 1] = pairs[2 * (j - 1) + 1];
      pairs[2 * j + 1] = v;
    }
  }
}

hipError_t launch_boundary(const SearchArgs& a0, double tol, uint32_t* counts,
                           const uint32_t* offs, uint64_t cap, uint32_t* pairs, hipStream_t st,
                           int num_cus) {
  SearchArgs a = a0;
  if (a.nlevels > 0) return hipErrorInvalidValue;  // raster grids only
  if (!a.tiles) {
    a.tile_q = kTileQ;
    a.ntiles = (uint32_t)(((uint64_t)a.nq + kTileQ - 1) / kTileQ);
  }
  if (a.ntiles == 0) return hipSuccess;
  uint32_t blocks = a.ntiles;
  if (a.persistent) {
    const uint32_t lim = (uint32_t)(num_cus > 0 ? num_cus : 256) * 8u;
    if (blocks > lim) blocks = lim;
  }
  hipLaunchKernelGGL(k_boundary, dim3(blocks), dim3(kTileQ), 0, st, a, tol, counts, offs, cap,
                     pairs);
  return hipGetLastError();
}

}  // namespace rtnn
