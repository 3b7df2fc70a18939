// Shared definitions of the gfx950 neighbor-search kernels.
//
// Every translation unit that computes d2 is compiled with -ffp-contract=off,
// so ((dx*dx + dy*dy) + dz*dz) is evaluated as five separately rounded FP32
// operations, the same as the reference's dot(O, O) with O = q - p
// (src/optixNSearch/geometry.cu:61-62,173-174; src/sutil/vec_math.h:433-436,
// 527-530) and as the CPU oracle.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rtnn {

constexpr int kWave = 64;

// One uniform grid (one per cloud in the batched path).  Cells are stored in
// raster order with x fastest, so the cells (x0..x1, y, z) of one row are
// contiguous in the cell table and their points contiguous in the sorted
// point array.
struct GridDesc {
  float mn[3];          // grid origin (bbox minimum of the cloud's points)
  float inv;            // 1 / cell edge
  float cell;           // cell edge
  int dim[3];           // cells per axis
  uint32_t cell_base;   // first cell of this grid in the global cell table
  uint32_t point_base;  // first point of this cloud in the input order
  uint32_t npoints;
  uint32_t morton;      // 1: cells numbered cell_base + Morton code (cubic grid)
};
static_assert(sizeof(GridDesc) == 48, "GridDesc layout");

// A run of queries (in processing order) that share one grid.
struct Tile {
  uint32_t q0, q1;  // [q0, q1) in the processing-order query array
  uint32_t grid;
  uint32_t pad_;
};

// Kernel argument block shared by the search kernels.
struct SearchArgs {
  const float4* spts;        // points sorted by cell: (x, y, z, bits(orig index)),
                             // followed by kSptsPadBytes of readable padding
  const uint32_t* cell_start;  // global cell table, cells + 1 entries
  const GridDesc* grids;
  const float4* queries;     // processing order: (x, y, z, bits(output row))
  const uint32_t* qsel;      // nullable: position i of this launch reads queries[qsel[i]]
  const uint32_t* nq_dev;    // nullable: query count read on the device (persistent)
  const Tile* tiles;         // nullable: implicit tiles of tile_q queries, grid 0
  uint32_t tile_q;           // queries per implicit tile (set by launch_search)
  uint32_t ntiles;
  uint32_t nq;
  float r;                   // search radius (per-query radius overrides when set)
  uint32_t k;                // output columns
  const float* qradius;      // nullable; indexed by processing position (queries[])
  float qradius_scale;       // per-query radius = min(r, qradius[.] * qradius_scale)
  int32_t* out_idx;          // [rows x k]
  float* out_d2;             // nullable
  uint32_t* short_rows;      // nullable: positions (in queries[]) with < k hits
  uint32_t* short_count;     // counter for short_rows
  uint32_t* short_hits;      // nullable: hits found by each short row (beside short_rows)
  // Fallback launch (qsel = a short list of an earlier launch at radius
  // qradius[.]): when set, query i of this launch found qsel_hits[i] < k hits
  // there, and its radius grows by hit_scale(qsel_hits[i], k) instead of
  // qradius_scale.
  const uint32_t* qsel_hits;
  float* radius_out;         // nullable: a short row stores the radius it was
                             // searched at in radius_out[queries position]
                             // (the next fallback stage grows it from there)
  int local_index;           // batched: report cloud-local point indices
  int persistent;            // grid-stride over tiles (XCD-chunked)
  // Hierarchical Morton grid (nlevels > 0): grids[l], l = 0 (finest) ..
  // nlevels-1, cubic with 2^(bits-l) cells per axis and cell 2^l * grids[0].cell;
  // tables below level_min are not built.  A query of radius R walks the
  // finest built level whose cell is >= R * level_alpha.
  int nlevels;
  int level_min;
  float level_alpha;
  // nullable: bounding box of every kChunkPts consecutive sorted points
  // (chunk c: min at [2c], max at [2c+1]); the octree walk skips the chunks of
  // a large leaf that lie beyond its limit.
  const float4* chunk_box;
  uint32_t tree_leaf;  // octree walk: cells of at most this many points are leaves
  uint32_t tree_lanes;  // octree walk: lanes per query (1: k_topk_tree; 2, 4, 8: k_topk_treeg)
  // Pooled kernel (k_topk_pool): a query with more hits than its LDS buffer
  // holds appends its processing position here (count in ovf_count); the
  // group kernel then searches those rows again.
  uint32_t* ovf_rows;
  uint32_t* ovf_count;
};

constexpr uint32_t kTreeLeafDefault = 256;

constexpr uint32_t kChunkPts = 32;

constexpr int kTileQ = 256;  // queries per tile (= threads of a KNN block)

// Sorted point buffers are allocated this much beyond their last point: the
// group kernels load a batch of candidate slots without a per-slot bounds
// select, and a slot past the end of its run reads at most 64 slots (1 KiB)
// further on (its value is not used).
constexpr uint64_t kSptsPadBytes = 4096;

// Radius growth of a row that found m < k hits within R: the k-th distance
// scales like (k / m)^(1/2) for points on surfaces and (k / m)^(1/3) in volume,
// so R * 1.15 * sqrt(k / m), kept in [1.15 R, 2 R] (2 R when nothing was
// found), usually holds the k-th neighbour at a fraction of the volume of 2 R.
// A row that still falls short goes on to the full radius, so the result does
// not depend on this estimate.
__device__ inline float hit_scale(uint32_t m, uint32_t k) {
  if (m == 0) return 2.0f;
  const float s = 1.15f * sqrtf((float)k / (float)m);
  return fminf(fmaxf(s, 1.15f), 2.0f);
}

__host__ __device__ inline float d2_of(float qx, float qy, float qz, float px, float py,
                                       float pz) {
  float dx = qx - px;
  float dy = qy - py;
  float dz = qz - pz;
  float sx = dx * dx;
  float sy = dy * dy;
  float sz = dz * dz;
  float s = sx + sy;
  return s + sz;
}

// 3D Morton (Z-order) code of 10-bit cell coordinates, the bit layout of the
// reference's MortonCode3 (helper_mortonCode.h:19-34): bit 3i of the code is
// bit i of x, 3i+1 of y, 3i+2 of z.
__host__ __device__ inline uint32_t spread3(uint32_t x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0xff0000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__host__ __device__ inline uint32_t morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (spread3(z) << 2) | (spread3(y) << 1) | spread3(x);
}

// (d2, index) packed so unsigned 64-bit order is the lexicographic order
// (d2 >= 0, so its IEEE bits order like its value).
__device__ inline uint64_t pack_key(float d2, uint32_t idx) {
  return ((uint64_t)__float_as_uint(d2) << 32) | idx;
}
__device__ inline float key_d2(uint64_t key) { return __uint_as_float((uint32_t)(key >> 32)); }
__device__ inline uint32_t key_idx(uint64_t key) { return (uint32_t)key; }

// Cell-range padding: a point with fl(d2) < fl(R*R) lies within R*(1+4u) of
// the query on each axis (u = 2^-24), and x -> floor((x - mn) * inv) is
// monotone in FP32, so padding R by 64u relative plus 2^-20 of the coordinate
// magnitude keeps every such point inside [lo, hi].
__device__ inline void cell_range(const GridDesc& g, float q, float R, int axis, int& lo,
                                  int& hi) {
  float pad = R * (1.0f + 0x1.0p-18f) + (fabsf(q) + fabsf(g.mn[axis]) + R) * 0x1.0p-20f;
  float a = floorf((q - pad - g.mn[axis]) * g.inv);
  float b = floorf((q + pad - g.mn[axis]) * g.inv);
  float dmax = (float)(g.dim[axis] - 1);
  a = fminf(fmaxf(a, 0.0f), dmax + 1.0f);  // NaN -> 0 (then b decides)
  b = fminf(fmaxf(b, -1.0f), dmax);
  lo = (int)a;
  hi = (int)b;
}

// Maps block b of n to a logical tile so that each XCD (blocks are dispatched
// round-robin over 8 XCDs) works on one contiguous chunk of tiles; the
// neighbouring queries of a chunk then share that XCD's L2.
__device__ inline uint32_t xcd_swizzle(uint32_t b, uint32_t n) {
  constexpr uint32_t kXcd = 8;
  uint32_t q = n / kXcd, rem = n % kXcd;
  uint32_t x = b % kXcd, j = b / kXcd;
  // XCD x owns chunk [start_x, start_x + q + (x < rem)).
  uint32_t start = x * q + (x < rem ? x : rem);
  uint32_t len = q + (x < rem ? 1u : 0u);
  if (j < len) return start + j;
  return b;  // not reached for b < n
}

}  // namespace rtnn
