// Host-side declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "common.h"

namespace rtnn {

// ---- sort.hip
int scan_tmp_words(uint64_t n);
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                          hipStream_t st);
uint64_t radix_tmp_words(uint64_t n);
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint64_t n, int bits,
                            uint32_t* tmp, hipStream_t st, int* passes_out);

// ---- grid.hip
constexpr int kBboxBlocks = 1024;
// part: 6 * kBboxBlocks floats; out: 6 floats (min xyz, max xyz); nonfinite: 1 u32.
hipError_t bbox(const float* xyz, uint64_t n, float* part, float* out, uint32_t* nonfinite,
                hipStream_t st);
// Per-cloud bbox: out[6 * c .. 6 * c + 6) for cloud c (offsets: nclouds + 1).
hipError_t bbox_segmented(const float* xyz, const uint64_t* offsets, uint32_t nclouds,
                          float* out, uint32_t* nonfinite, hipStream_t st);
// Per-cloud grids planned on the device from the per-cloud bboxes: edge
// cells[i] (device array, size classes) or `cell` when cells is null, every
// edge grown by the same 1.25x steps until the total cell count is <= cap
// (one cell per cloud if none fits, status[1] = 1); status[2] = total cells.
hipError_t plan_batched(const float* bboxes, const uint64_t* offsets, uint32_t nclouds,
                        float cell, const float* cells, uint64_t cap, GridDesc* grids,
                        uint32_t* status, hipStream_t st);
// keys[i] = grid cell of point i, vals[i] = i, counts[key] += 1 (counts nullable).
hipError_t cell_keys(const float* xyz, uint64_t n, const GridDesc* grids,
                     const uint64_t* offsets, uint32_t nclouds, uint32_t* keys,
                     uint32_t* vals, uint32_t* counts, hipStream_t st);
// One cloud with its descriptor by value; *gstore receives it (device).
hipError_t cell_keys_single(const float* xyz, uint64_t n, const GridDesc& g, GridDesc* gstore,
                            uint32_t* keys, uint32_t* vals, hipStream_t st);
// cell_start[0..ncells] from the sorted keys (first position with
// (key >> shift) >= c).  tmp: cell_bounds_tmp_words(ncells) words, 8-byte
// aligned.
uint64_t cell_bounds_tmp_words(uint64_t ncells);
hipError_t cell_bounds(const uint32_t* sorted_keys, uint64_t n, uint32_t ncells,
                       uint32_t* cell_start, uint32_t* tmp, hipStream_t st, int shift = 0);
// keys[i] = Morton code of point i's cell on the cubic level-0 grid g0.
hipError_t morton_keys(const float* xyz, uint64_t n, const GridDesc* g0, uint32_t* keys,
                       uint32_t* vals, hipStream_t st);
// counts[cell of spts[j]] += 1 on grid g, one atomic per run of equal cells.
hipError_t count_sorted(const float4* spts, uint64_t n, const GridDesc* g, uint32_t* counts,
                        hipStream_t st);
// spts[j] = (xyz[vals[j]], bits(vals[j])).
hipError_t gather_points(const float* xyz, const uint32_t* vals, uint64_t n, float4* spts,
                         hipStream_t st);
// Bounding box of every kChunkPts consecutive points of spts (2 float4 each).
hipError_t chunk_boxes(const float4* spts, uint64_t n, float4* box, hipStream_t st);
// Morton key (10 bits per axis) of each query's cell in grid g, shifted right
// by `shift` bits per axis (shift < 0: derived on the device from g's widest
// axis, keeping 8 bits per axis); vals[i] = i.
hipError_t query_morton_keys(const float* q, uint64_t nq, const GridDesc* g, int shift,
                             uint32_t* keys, uint32_t* vals, hipStream_t st);
hipError_t morton3_codes(const uint32_t* xyz, uint64_t n, uint32_t* codes, hipStream_t st);
hipError_t fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t st);
// out[row of qproc[i]] = min(val[i], r): per-query values back in input order.
hipError_t scatter_rows(const float4* qproc, const float* val, uint64_t n, float r, float* out,
                        hipStream_t st);
hipError_t iota_u32(uint32_t* p, uint64_t n, hipStream_t st);
// Coarse count grid: counts[cell of point] += 1 (grid g, single cloud).
hipError_t count_cells(const float* xyz, uint64_t n, const GridDesc* g, uint32_t* counts,
                       hipStream_t st);

// ---- partition.hip
// Launch radius per processing position from the point-count grid (edge
// grid->cell): min(r, sqrt(3)*(i+1)*cell) for the smallest megacell i holding
// >= k+1 points, else r.  iter_hist (nullable) counts queries per i.
hipError_t partition_radius(const float4* qproc, uint32_t nq, const GridDesc* grid,
                            const uint32_t* counts, uint32_t k, float r, float* qradius,
                            uint32_t* iter_hist, int hist_bins, hipStream_t st);

// Same rule on a hierarchical Morton grid (see partition.hip).
hipError_t partition_hier(const float4* qproc, uint32_t nq, const GridDesc* levels, int nlevels,
                          int level_min, const uint32_t* table, uint32_t k, float r,
                          float* qradius, hipStream_t st);

// Density buckets: radii >= snap_r become r (the sparse bucket, searched at
// r at once).
hipError_t bucket_snap(float* qradius, uint32_t nq, float snap_r, float r, hipStream_t st);

// ---- search.hip
// Runs the top-k search.  group_lanes = 1 selects the thread-per-query kernel
// (register-resident sorted top-k, k <= 64); 4, 16 or 64 select the group
// kernel (ballot compaction into an LDS buffer, rank-sorted output).
// wide = false selects the compact LDS buffer of the group kernel (k + 2G keys
// per query instead of 2k + G; raster grids only); tiled > 0 the group kernel
// that stages each wave's candidates in LDS, tiled points per wave (256, 512
// or 1024; k_topk_tiled: raster grids, compact buffers, G >= 4).  deep = true
// sizes the 4-lane kernel's buffer for
// k <= 16 as for k = 32 (fewer selects where rows have many hits: the
// fallback stages).  Returns hipErrorInvalidValue when no kernel variant
// covers (k, group_lanes).
// pool = 16, 32 or 64 selects the pooled kernel (k_topk_pool: the candidates
// of that many queries per wave tested as one pool; k <= 64, one raster grid,
// no tile list) followed by the 4-lane group kernel over the rows whose LDS
// buffers overflowed (a.ovf_rows / a.ovf_count: device list and counter,
// counter zero on entry).
hipError_t launch_search(const SearchArgs& a, int knn_mode, int group_lanes, hipStream_t st,
                         int num_cus, bool wide = true, int tiled = 0, bool deep = false,
                         int pool = 0);
// Every device / pinned allocation of the library (tracked for
// rtnn_mem_stats; logged with RTNN_MEMSTAT set).
hipError_t dev_alloc(void** p, size_t bytes);
void dev_free(void* p, size_t bytes);
hipError_t pinned_alloc(void** p, size_t bytes);
void pinned_free(void* p, size_t bytes);
uint64_t mem_dev_cur();
uint64_t mem_dev_peak();
uint64_t mem_pin_cur();
uint64_t mem_pin_peak();
// Largest k supported for a group width (0 = unsupported width).
uint32_t max_k_for_group(int group_lanes);
// Boundary-pair report over a raster grid (a.r = r): count pass when offs is
// null (counts[row]), else write pass into pairs[2 * (offs[row] + i)] (first
// cap pairs), each row sorted by point index.
hipError_t launch_boundary(const SearchArgs& a, double tol, uint32_t* counts,
                           const uint32_t* offs, uint64_t cap, uint32_t* pairs, hipStream_t st,
                           int num_cus);

}  // namespace rtnn
