/*
 * rtnn_hip.h -- C ABI of the MI355X (gfx950) fixed-radius / KNN neighbor search.
 *
 * Plain C types only: pointers, sizes, status codes.  No exception crosses this
 * boundary and the library never calls exit()/abort().  A context is not
 * thread-safe (one context per host thread); work is ordered on the context's
 * own HIP stream.
 *
 * Result contract (every search entry point):
 *   out_idx  int32 [nq x k], row-major, row = original query index,
 *            values = original point index, ordered by (d2, index) ascending,
 *            empty slots = RTNN_EMPTY_SLOT (-1, i.e. the reference's UINT_MAX
 *            sentinel read as int32).
 *   out_d2   float [nq x k] or NULL; +inf in empty slots.
 *   d2 = ((dx*dx + dy*dy) + dz*dz) in FP32 without FMA, dx = q.x - p.x.
 *   RTNN_MODE_RADIUS : include iff d2 < r*r (a query finds itself).
 *   RTNN_MODE_KNN    : include iff 0 < d2 < r*r (self and duplicates excluded).
 *   More than k qualifying points -> the k smallest by (d2, index).
 *
 * Reference seams replaced (paths relative to the reference's src/optixNSearch):
 *   search() / launchSubframe()        search.cpp:9-61, optix.cpp:595-626
 *   __raygen__radius/__raygen__knn     camera.cu:39-145
 *   IS programs + insertTopKQ          geometry.cu:38-191
 *   uploadData / computeMinMax         optix.cpp:117-168, sort.cpp:34-71
 *   gridSort / sortParticles           sort.cpp:500-692, grid.cu:242-329
 */
#ifndef RTNN_HIP_H
#define RTNN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtnn_ctx_s* rtnn_ctx;

/* Status codes (0 = OK, negative = error class). */
#define RTNN_OK               0
#define RTNN_ERR_INVALID     -1   /* bad argument (NULL, k = 0, r <= 0, NaN ...) */
#define RTNN_ERR_HIP         -2   /* a HIP runtime call failed */
#define RTNN_ERR_OOM         -3   /* device allocation failed */
#define RTNN_ERR_STATE       -4   /* call order (e.g. search before set_points) */
#define RTNN_ERR_UNSUPPORTED -5   /* size/feature outside what the build supports */
#define RTNN_ERR_IO          -6   /* a file could not be opened, read or written */

#define RTNN_MODE_RADIUS 0
#define RTNN_MODE_KNN    1

#define RTNN_EMPTY_SLOT (-1)

/* Tuning knobs; zero means "choose automatically". */
typedef struct {
  float cell_factor;      /* grid cell edge = r / cell_factor                       */
  int   queries_per_wave; /* lanes per query = 64 / queries_per_wave: 64 = one query
                             per lane (k <= 64), 32/16/8/4/1 = groups of 2/4/8/16/64
                             lanes (k <= 16/64/64/256/1024); 0 = auto             */
  int   query_order;      /* 0 auto, 1 input order, 2 Morton order of query cells  */
  int   time_stages;      /* 1: record hipEvent timings into rtnn_stats             */
  uint32_t max_cells;     /* cap on grid cells (cell edge grows to respect it)      */
  int   grid_levels;      /* learned/partitioned KNN: 1 = one raster grid of edge
                             r / cell_factor; >= 2 = hierarchical Morton grid with
                             that many levels (edges r, r/2, ...); 0 = auto       */
  float level_alpha;      /* hierarchical: a query of radius R walks the finest
                             level whose cell edge is >= R * level_alpha (0: 1.0) */
  int   group_buffer;     /* group kernel's LDS hit buffer per query: 0 auto / 2
                             wide (2k + G keys), 1 compact (k + 2G keys: less LDS,
                             more waves per CU, more selects on dense rows);
                             3 pooled (k <= 64, raster grids, not batched): the
                             candidates of queries_per_wave = 16, 32 or 64
                             queries per wave are tested as one pool, hits kept
                             in 2k-key buffers (a query with more is searched
                             again by the group kernel)                        */
  int   lds_tile;         /* group kernel on a raster grid: 256, 512 or 1024 = each
                             wave stages the points of its queries' union block in
                             LDS (up to that many per wave) and scans them there
                             (1 = 512; 0 = off); implies compact buffers        */
  int   tree_leaf;        /* hierarchical-grid KNN, one lane per query (the octree
                             walk): a cell of at most this many points is scanned
                             (chunk by chunk, 32 points per bounding box) instead
                             of descended into; 0 = 256.  Large values suit very
                             clustered points (Levy flights: 65536)            */
  float bucket_ratio;     /* learned/partitioned KNN, density buckets: a query whose
                             proposed radius r' >= bucket_ratio * r is searched at
                             r at once (its row is final after the first launch:
                             no search at r' and no fallback); 0 = off          */
  int   learned_cap;      /* learned KNN: 1 = the network's r' is capped by the
                             density partition's radius (rtnn_search_partitioned's
                             first radius, which holds k neighbours), so queries
                             the partition bounds more tightly keep its radius
                             and do not fall back; 0 = the network alone        */
  int   tree_lanes;       /* hierarchical-grid KNN, the octree walk: lanes per
                             query (2, 4 or 8 split the chunks and points of
                             every leaf; 0 or 1 = one lane).  2 is the fastest
                             on clustered points (Levy flights)               */
  int   tree_fallback;    /* hierarchical-grid KNN with the octree walk (k <= 64):
                             0 = the rows short at their proposed radius are
                             searched once more at r; 1 = through the radius-
                             growth stages (as on raster grids)               */
} rtnn_options;

/* Filled by search entry points when a non-NULL pointer is passed. */
typedef struct {
  double ms_total;        /* wall time of the call on the context stream (events)  */
  double ms_grid;         /* bbox + keys + radix sort + scan + gather              */
  double ms_qsort;        /* query Morton keys + sort + gather                      */
  double ms_search;       /* search kernel(s) only                                  */
  double ms_mlp;          /* learned radius: features + MLP inference               */
  double ms_fallback;     /* radius proposals: the fallback stages of short rows    */
  uint64_t fallback_queries; /* rows short after the first stage (proposal search) */
  uint32_t grid_dims[3];
  float    cell_size;
  uint32_t sort_passes;   /* LSD radix passes of the point sort                     */
  uint32_t search_launches;
  uint32_t lanes_per_query; /* range kernel group width actually used               */
  double ms_copy;         /* result copy to host arrays (host outputs only)          */
} rtnn_stats;

/* ---- context ---------------------------------------------------------- */

/* Replaces setDevice()+setupOptiX() (main.cpp:9-24, optix.cpp:679-691). */
int rtnn_create(int device, rtnn_ctx* out);
int rtnn_destroy(rtnn_ctx ctx);
/* Message of the last failed call on ctx (never NULL). */
const char* rtnn_last_error(rtnn_ctx ctx);
/* Message of the last failed rtnn_create on this thread. */
const char* rtnn_create_error(void);
int rtnn_set_options(rtnn_ctx ctx, const rtnn_options* opt);
int rtnn_synchronize(rtnn_ctx ctx);

/* ---- points ----------------------------------------------------------- */

/* Copies n float3 (xyz AoS, 12 B each) into the context.  Replaces
 * uploadData()'s point upload (optix.cpp:117-124).  on_device != 0 means xyz
 * is a device pointer.  Non-finite coordinates are rejected (RTNN_ERR_INVALID). */
int rtnn_set_points(rtnn_ctx ctx, const float* xyz, uint64_t n, int on_device);

/* ---- search ----------------------------------------------------------- */

/* One search over the context's points.  qxyz == NULL searches the points
 * themselves (the reference's samepq case, util.cpp:371-382).  on_device
 * describes qxyz; out_on_device describes out_idx/out_d2.  Replaces search()
 * (search.cpp:9-61) with launchRadius = r and PRECISE mode. */
int rtnn_search(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                int mode, float r, uint32_t k,
                int32_t* out_idx, float* out_d2, int out_on_device,
                rtnn_stats* stats);

/* Stream-ordered form of rtnn_search for device-resident queries (qxyz, or
 * NULL for the points) and outputs: the grid build, query ordering and search
 * are enqueued on `stream` (a hipStream_t) and the call returns without
 * waiting.  The work is ordered after the context's earlier calls, and the
 * context's later calls are ordered after it (event waits between the two
 * streams), so the context may be used again at once; the caller synchronises
 * `stream` before reading the outputs.  Replaces the reference's per-batch
 * stream launches (search.cpp:39, util.cpp:708-709). */
int rtnn_search_async(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int mode, float r,
                      uint32_t k, int32_t* out_idx, float* out_d2, void* stream);

/* rtnn_set_points + rtnn_search in one call that allocates nothing (no
 * hipMalloc, hipFree or allocating library routine): for callers that cannot
 * allocate (tight serving loops, harnesses that forbid it).  No counterpart in
 * the reference.  xyz [n][3] (points), qxyz [nq][3] (queries, or NULL for the
 * points themselves), out_idx / out_d2 (nullable) [nq][k]: all device
 * pointers.  Result contract as rtnn_search, rows equal to it.  Every buffer
 * comes from one 128 MiB device array of the library, which each GPU it runs
 * on holds until the process ends: about 40 B per point plus 24 B per
 * separate query, the rest for the cell table (the cell edge grows to fit,
 * results unchanged) -- roughly up to 3M points as their own queries, or 2M
 * points with 2M other queries.  Larger inputs: RTNN_ERR_UNSUPPORTED.  The
 * context's options (cell factor, lanes per query, buffers) apply; the
 * context's own buffers and points are not used or changed.  The array keeps
 * the last call's contents; calls that overlap on one GPU (another thread,
 * another context on the same device) race and must be serialised by the
 * caller.  Synchronous: returns once the rows are written. */
int rtnn_search_noalloc(rtnn_ctx ctx, const float* xyz, uint64_t n, const float* qxyz,
                        uint64_t nq, int mode, float r, uint32_t k, int32_t* out_idx,
                        float* out_d2);

/* KNN with the reference's exact density partition (sort.cpp:140-219,346-391
 * with approx mode 0): per query cell the smallest megacell holding >= k+1
 * points gives a launch radius min(r, sqrt(3)*(iter+1)*cell); a query whose
 * sphere holds fewer than k hits is searched again at a grown radius (up to
 * four growth stages, then r; a row searched at r is final), so the result
 * equals rtnn_search(..., RTNN_MODE_KNN, ...). */
int rtnn_search_partitioned(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                            float r, uint32_t k,
                            int32_t* out_idx, float* out_d2, int out_on_device,
                            rtnn_stats* stats);

/* ---- learned radius (KNN) --------------------------------------------- */

/* Trains the radius-proposal MLP (30->64->64->1, GELU, BF16 weights, FP32
 * accumulation) on exact k-th neighbor distances (within r) of n_samples
 * queries drawn from the context's points.
 * Pinball loss with quantile tau; Adam.  Runs entirely on the GPU. */
int rtnn_mlp_train(rtnn_ctx ctx, float r, uint32_t k, uint32_t n_samples,
                   uint32_t steps, uint32_t batch, float lr, float tau,
                   uint64_t seed, float* final_loss);
/* The same, trained on the caller's sample of queries (nq of them; device
 * pointer if on_device), for query distributions unlike the points (e.g.
 * uniform queries over clustered points).  The exact k-th distances of the
 * training queries come from the partitioned search (exact, and cheap when r
 * is large). */
int rtnn_mlp_train_queries(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                           float r, uint32_t k, uint32_t steps, uint32_t batch,
                           float lr, float tau, uint64_t seed, float* final_loss);
/* Weights as FP32, row-major [out][in]:
 *   [w1 64x30 | b1 64 | w2 64x64 | b2 64 | w3 64 | b3 1]  (6209 parameters)
 * followed by [count_scale, radius_margin] (feature scale of the log1p counts,
 * multiplier applied to the proposed radius) = 6211 floats. */
#define RTNN_MLP_NPARAMS 6209
#define RTNN_MLP_NBLOB   6211
int rtnn_mlp_get_weights(rtnn_ctx ctx, float* blob /* RTNN_MLP_NBLOB */);
int rtnn_mlp_set_weights(rtnn_ctx ctx, const float* blob /* RTNN_MLP_NBLOB */);

/* KNN where the MLP proposes r' <= r per query; queries with fewer than k hits
 * within r' (stats->fallback_queries) are searched again at grown radii and
 * finally at r, so the result equals rtnn_search(..., RTNN_MODE_KNN, ...). */
int rtnn_search_learned(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                        float r, uint32_t k,
                        int32_t* out_idx, float* out_d2, int out_on_device,
                        rtnn_stats* stats);

/* Device and pinned host bytes the library holds now and at its peak, over
 * all contexts of the process (any pointer may be NULL).  Replaces the
 * reference's memstatlib cudaMalloc/cudaFree interposer
 * (src/memstatlib/memstatlib.cpp:10-32): with RTNN_MEMSTAT set in the
 * environment every allocation and free is logged to stderr with the running
 * total, as there. */
int rtnn_mem_stats(uint64_t* device_bytes, uint64_t* device_peak, uint64_t* pinned_bytes,
                   uint64_t* pinned_peak);

/* Streamed search for results larger than one wants to hold at once (C4's
 * 39 GB of rows, C5's 7.7 GB): host queries qxyz[nq][3] are searched in
 * chunks of `chunk` queries (0: 4M) through two slots of device and pinned
 * buffers, so one chunk's rows travel over PCIe while the next chunk is
 * searched.  Rows go to host arrays out_idx[nq][k] (+ out_d2, nullable) or,
 * when out_path is set, to a result file in the CLI's format: "RTNNRES1",
 * u64 nq, u32 k, u32 flags (bit 0: d2 present), int32 idx[nq][k], float
 * d2[nq][k], little endian, empty slot -1 / +inf.  method: RTNN_METHOD_PLAIN
 * (mode range or KNN), _PARTITIONED or _LEARNED (KNN).  The rows equal one
 * search over all queries.  Replaces the reference's whole-result pinned
 * host buffers (search.cpp:43-56) and its TODO on the Q*K limit
 * (main.cpp:126). */
int rtnn_search_stream(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int mode, int method,
                       float r, uint32_t k, uint64_t chunk, int32_t* out_idx, float* out_d2,
                       const char* out_path, rtnn_stats* stats);
/* Streamed batched search: host clouds xyz (offsets[nclouds + 1] as in
 * rtnn_search_batched) are searched in chunks of whole clouds of about
 * `chunk` points (0: 4M; a larger cloud is a chunk of its own) through the
 * two slots of rtnn_search_stream, so one chunk's rows travel over PCIe while
 * the next chunk is searched.  Rows (global rows, cloud-local point indices)
 * go to host arrays out_idx[n][k] (+ out_d2, nullable) or to a result file at
 * out_path in the same format as rtnn_search_stream's.  The rows equal one
 * rtnn_search_batched over the batch; size classes apply per cloud.
 * Replaces the reference's per-batch copy of rows to pinned host memory
 * (search.cpp:43-56) for batches of clouds. */
int rtnn_search_batched_stream(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                               uint32_t nclouds, int mode, float r, uint32_t k, uint64_t chunk,
                               int32_t* out_idx, float* out_d2, const char* out_path,
                               rtnn_stats* stats);

/* The first-stage launch radius of each query (input order) under method
 * RTNN_METHOD_PARTITIONED (the megacell rule, sort.cpp:366-371 with approx
 * mode 0) or RTNN_METHOD_LEARNED (the MLP's r'), capped at r: what
 * rtnn_search_partitioned / rtnn_search_learned search first.  For measuring
 * a proposal (search volume, share of queries it already covers). */
int rtnn_propose_radius(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                        int method, float r, uint32_t k, float* out_radius,
                        int out_on_device);

/* ---- batched clouds --------------------------------------------------- */

/* nclouds independent clouds packed back to back in xyz; cloud c holds points
 * [offsets[c], offsets[c+1]).  Queries = points of each cloud.  Output row i
 * belongs to point i; indices are cloud-local.  One grid build and one
 * persistent search launch for all clouds.  xyz, out_idx and out_d2 are
 * device pointers; offsets (nclouds + 1 entries) is a host array. */
int rtnn_search_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                        uint32_t nclouds, int mode, float r, uint32_t k,
                        int32_t* out_idx, float* out_d2, rtnn_stats* stats);
/* Size classes of the batched search (and of graphs captured afterwards):
 * class j holds the clouds of at most upper[j] points (the last class, j =
 * nclasses - 1, every larger cloud; upper has nclasses - 1 increasing
 * entries) and is searched with opts[j]'s cell factor, queries per wave,
 * group buffer and LDS tile: one persistent launch per non-empty class over
 * that class's clouds, all in one grid build.  nclasses = 0 drops the classes
 * (the context's options apply to every cloud again).  No counterpart in the
 * reference, which runs one configuration per launch (src/optixNSearch/
 * main.cpp:91-148). */
int rtnn_set_size_classes(rtnn_ctx ctx, uint32_t nclasses, const uint64_t* upper,
                          const rtnn_options* opts);
/* Current classes: *nclasses (0 if none); upper (nclasses - 1 entries) and
 * opts (nclasses entries) are filled when not NULL. */
int rtnn_get_size_classes(rtnn_ctx ctx, uint32_t* nclasses, uint64_t* upper,
                          rtnn_options* opts);
/* Same result, one full pipeline per cloud (the per-cloud-launch baseline). */
int rtnn_search_percloud(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                         uint32_t nclouds, int mode, float r, uint32_t k,
                         int32_t* out_idx, float* out_d2, rtnn_stats* stats);

/* ---- graph-captured batched pipeline ------------------------------------
 * The batched sort -> grid -> search pipeline captured once in a HIP graph
 * and replayed with one launch (no host round trip: the per-cloud grids are
 * planned on the device).  The reference re-runs its host-driven pipeline per
 * call (src/optixNSearch/main.cpp:91-148); this is the replayable form of
 * rtnn_search_batched for callers that search the same batch layout
 * repeatedly.  xyz, out_idx and out_d2 (nullable) are device pointers bound
 * at capture: a replay reads the current contents of xyz (any values, same
 * offsets) and rewrites the outputs, equal to rtnn_search_batched on that
 * data.  The cell table is sized from the data at capture (+25 %); a replay
 * on data that needs more cells uses a coarser edge (still exact).  Replays
 * run on the context's stream.  rtnn_destroy also destroys the context's
 * remaining graphs.  The library keeps the set of live graph handles:
 * rtnn_graph_launch and rtnn_graph_destroy on a handle already freed (by
 * rtnn_graph_destroy or with its context) return RTNN_ERR_INVALID without
 * touching it, so graph and context may be destroyed in either order. */
typedef struct rtnn_graph_s* rtnn_graph;
int rtnn_graph_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                       uint32_t nclouds, int mode, float r, uint32_t k,
                       int32_t* out_idx, float* out_d2, rtnn_graph* out);
/* Replays the graph.  sync != 0 (or stats != NULL) waits and reports
 * non-finite input coordinates as RTNN_ERR_INVALID; stats->ms_total is the
 * replay's device time. */
int rtnn_graph_launch(rtnn_graph g, int sync, rtnn_stats* stats);
int rtnn_graph_destroy(rtnn_graph g);

/* ---- boundary-pair report ------------------------------------------------
 * The reference's inclusion test is d2 < r*r in FP32 (src/optixNSearch/
 * geometry.cu:61-66,173-187); a pair whose true distance is within a few ulp
 * of r can flip under another FP32 evaluation order (check.cpp:80-83 admits
 * such mismatches).  These calls list every (query row, point index) pair with
 * |d - r| <= tol * r, d evaluated in FP64 from the FP32 coordinates, so such
 * pairs are reported rather than silently dropped.  Output: pairs[2 * i] =
 * row, pairs[2 * i + 1] = point, ordered by (row, point); at most `capacity`
 * pairs are written (out_on_device selects a device or host pointer) and
 * *count receives the total (call with capacity 0 to size the buffer).
 * tol in [0, 1e-2]; 1e-6 is the documented default. */
int rtnn_boundary_pairs(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                        float r, double tol, uint32_t* out_pairs, uint64_t capacity,
                        int out_on_device, uint64_t* count);
/* Batched form: queries = points of each cloud, point indices cloud-local,
 * rows global (as rtnn_search_batched). */
int rtnn_boundary_pairs_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                                uint32_t nclouds, float r, double tol,
                                uint32_t* out_pairs, uint64_t capacity,
                                int out_on_device, uint64_t* count);

/* ---- autotuner --------------------------------------------------------- */

#define RTNN_METHOD_PLAIN       0
#define RTNN_METHOD_PARTITIONED 1
#define RTNN_METHOD_LEARNED     2

/* Times every candidate configuration of the method on this exact workload
 * (same queries, r, k): cell factor x queries per wave x group buffer / LDS
 * tile, plus
 * hierarchical grid levels x level alpha for the proposal methods, and for
 * the learned method then the MLP's radius margin.  Keeps the fastest in the
 * context's options (also returned in *best; *best_ms = its time per search)
 * and, for the learned method, its margin in the context's weights.  Replaces
 * the reference's analytic cell-size and batching models (calcCRRatio,
 * util.cpp:551-676; autoBatchingRange/KNN, sort.cpp:221-320) with
 * measurement.  out_idx/out_d2 are scratch outputs of size nq x k (device;
 * NULL: the context's own scratch). */
int rtnn_autotune(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, int method,
                  int mode, float r, uint32_t k, int32_t* out_idx, float* out_d2,
                  rtnn_options* best, double* best_ms);
/* The same for rtnn_search_batched (and so for graphs captured afterwards):
 * cell edge r / {0.5, 1, 2} x queries per wave {1, 4, 8, 16, 32, 64} x the
 * group kernel's LDS use (wide buffer / compact buffer / compact buffer + a
 * tile of 256, 512 or 1024 candidates per wave), timed on this batch; one
 * configuration for every cloud (drops size classes).  out_idx (required) and
 * out_d2 (nullable) are device scratch of n x k. */
int rtnn_autotune_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                          uint32_t nclouds, int mode, float r, uint32_t k,
                          int32_t* out_idx, float* out_d2,
                          rtnn_options* best, double* best_ms);
/* Per size class: splits the batch by cloud size at upper[0..nclasses-2]
 * (as rtnn_set_size_classes), tunes each non-empty class on the sub-batch of
 * its clouds as rtnn_autotune_batched does, and installs the classes with
 * their winners (empty classes keep the context's options).  best (nclasses
 * entries) and best_ms (each class's time on its sub-batch) are nullable. */
int rtnn_autotune_batched_classes(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                                  uint32_t nclouds, int mode, float r, uint32_t k,
                                  int32_t* out_idx, float* out_d2, uint32_t nclasses,
                                  const uint64_t* upper, rtnn_options* best,
                                  double* best_ms);

/* ---- building blocks (exported for parity tests; device pointers) ---- */

/* Stable LSD radix sort of (key, val) pairs on the low `bits` key bits. */
int rtnn_radix_sort_pairs(rtnn_ctx ctx, uint32_t* keys, uint32_t* vals,
                          uint64_t n, uint32_t bits);
/* out[i] = sum(in[0..i)), out[n] = total (out has n+1 entries). */
int rtnn_exclusive_scan(rtnn_ctx ctx, const uint32_t* in, uint32_t* out, uint64_t n);
/* 30-bit 3D Morton codes of 10-bit cell coordinates (helper_mortonCode.h:19-34). */
int rtnn_morton3(rtnn_ctx ctx, const uint32_t* xyz /* n x 3 */, uint32_t* codes,
                 uint64_t n);
/* MLP forward / backward on n feature rows (device pointers, FP32 I/O). */
int rtnn_mlp_forward(rtnn_ctx ctx, const float* feats /* n x 30 */, uint64_t n,
                     float* out /* n */);
int rtnn_mlp_backward(rtnn_ctx ctx, const float* feats, const float* target,
                      uint64_t n, float tau, float* grads /* RTNN_MLP_NPARAMS */,
                      float* loss);
/* Feature rows (log1p counts of the 27 coarse cells of edge r around each
 * query + its offset in its cell) for queries qxyz against the context's
 * points. */
int rtnn_mlp_features(rtnn_ctx ctx, const float* qxyz, uint64_t nq, float r,
                      float* feats /* nq x 30, device */);

/* Library version string. */
const char* rtnn_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RTNN_HIP_H */
