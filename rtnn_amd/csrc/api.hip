// C ABI of the gfx950 neighbor search: context, grid pipeline, entry points.
//
// Pipeline of one search (the reference's uploadData -> sortParticles ->
// createGeometry -> search, main.cpp:84-143, with the BVH replaced by the cell
// list):
//   points bbox (once per point set) -> grid plan (host) -> cell keys + counts
//   -> exclusive scan -> stable radix sort -> gather -> query Morton order ->
//   search kernel -> rows in original query order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/rtnn_hip.h"
#include "common.h"
#include "internal.h"
#include "mlp.h"

// Allocation statistics of the library (all contexts of the process): the
// reference's memstatlib (src/memstatlib/memstatlib.cpp:10-32) interposes
// cudaMalloc/cudaFree to log each allocation and the running total; here
// every device and pinned allocation of the library goes through these
// functions, which keep the current and peak bytes and, with RTNN_MEMSTAT set
// in the environment, log the same lines.
namespace rtnn {
namespace {
std::atomic<uint64_t> g_dev_cur{0}, g_dev_peak{0}, g_pin_cur{0}, g_pin_peak{0};
bool memstat_log() {
  static const bool on = std::getenv("RTNN_MEMSTAT") != nullptr;
  return on;
}
void bump_peak(std::atomic<uint64_t>& peak, uint64_t v) {
  uint64_t p = peak.load();
  while (v > p && !peak.compare_exchange_weak(p, v)) {
  }
}
}  // namespace

// Simulated allocation failure, forced from the host for tests of the OOM
// paths: with RTNN_FAIL_ALLOC_BYTES=b in the environment every device or
// pinned allocation of at least b bytes fails as out of memory without
// reaching the runtime.  Read at each allocation, so a caller may set and
// clear it between calls.
static bool forced_alloc_failure(size_t bytes) {
  const char* v = std::getenv("RTNN_FAIL_ALLOC_BYTES");
  if (!v || !*v) return false;
  char* end = nullptr;
  const unsigned long long b = std::strtoull(v, &end, 10);
  return end != v && bytes >= b;
}

hipError_t dev_alloc(void** p, size_t bytes) {
  const hipError_t e = forced_alloc_failure(bytes) ? hipErrorOutOfMemory : hipMalloc(p, bytes);
  if (e != hipSuccess) {
    if (memstat_log())
      std::fprintf(stderr, "[MEM_STATS] hipMalloc of %.3f MB failed (%s), in use %.3f MB\n",
                   bytes / 1048576.0, hipGetErrorString(e), g_dev_cur.load() / 1048576.0);
    return e;
  }
  const uint64_t cur = g_dev_cur += bytes;
  bump_peak(g_dev_peak, cur);
  if (memstat_log())
    std::fprintf(stderr, "[MEM_STATS] hipMalloc (%p): %.3f MB\n[MEM_STATS] %.3f MB\n", *p,
                 bytes / 1048576.0, cur / 1048576.0);
  return e;
}
void dev_free(void* p, size_t bytes) {
  (void)hipFree(p);
  const uint64_t cur = g_dev_cur -= bytes;
  if (memstat_log())
    std::fprintf(stderr, "[MEM_STATS] hipFree (%p): %.3f MB\n[MEM_STATS] %.3f MB\n", p,
                 bytes / 1048576.0, cur / 1048576.0);
}
hipError_t pinned_alloc(void** p, size_t bytes) {
  const hipError_t e = forced_alloc_failure(bytes) ? hipErrorOutOfMemory
                                                   : hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) return e;
  bump_peak(g_pin_peak, g_pin_cur += bytes);
  return e;
}
void pinned_free(void* p, size_t bytes) {
  (void)hipHostFree(p);
  g_pin_cur -= bytes;
}
uint64_t mem_dev_cur() { return g_dev_cur.load(); }
uint64_t mem_dev_peak() { return g_dev_peak.load(); }
uint64_t mem_pin_cur() { return g_pin_cur.load(); }
uint64_t mem_pin_peak() { return g_pin_peak.load(); }
}  // namespace rtnn

using namespace rtnn;

namespace {

thread_local std::string g_create_err;


struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      // Queued work may still read the old buffer.
      (void)hipDeviceSynchronize();
      rtnn::dev_free(p, cap);
      p = nullptr;
      cap = 0;
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = rtnn::dev_alloc(&p, want);
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    cap = want;
    return hipSuccess;
  }
  void release() {
    if (p) rtnn::dev_free(p, cap);
    p = nullptr;
    cap = 0;
  }
  template <class T>
  T* as() const { return reinterpret_cast<T*>(p); }
};

}  // namespace

// Pinned host staging memory (grown on demand, freed with the context).
struct PinBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      (void)hipDeviceSynchronize();  // copies may still read or write it
      rtnn::pinned_free(p, cap);
      p = nullptr;
      cap = 0;
    }
    hipError_t e = rtnn::pinned_alloc(&p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
      p = nullptr;
      return e;
    }
    cap = std::max<size_t>(bytes, 256);
    return hipSuccess;
  }
  void release() {
    if (p) rtnn::pinned_free(p, cap);
    p = nullptr;
    cap = 0;
  }
};

// Scratch memory of rtnn_search_noalloc: one file-scope device array, which
// every GPU the library runs on holds (its own copy) until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is taken once per device with hipGetSymbolAddress (with
// that device current) and only that address is used.
constexpr size_t kScratchBytes = 128ull << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_addr[kMaxDevices] = {};

// Device address of `scratch` on `device` (current), looked up once per device.
hipError_t scratch_address(int device, void** out) {
  *out = nullptr;
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return hipSuccess;
}
}  // namespace

struct rtnn_ctx_s {
  int device = 0;
  hipStream_t st = nullptr;
  int num_cus = 256;
  std::string err;
  rtnn_options opt{};

  // points
  DevBuf pts;
  uint64_t n = 0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};

  // grid
  DevBuf keys, vals, counts, cell_start, spts, tmp, grids;
  DevBuf cboxes;  // hierarchical grids: bounding box per kChunkPts sorted points
  GridDesc hgrid{};
  uint64_t ncells = 0;
  int sort_passes = 0;
  bool spts_current = false;  // spts/cell_start describe ctx->pts
  int hlevels = 0;            // > 0: grids/cell_start hold a hierarchical Morton grid
  int hlevel_min = 0;

  // queries / outputs
  DevBuf qin, qkeys, qvals, qsorted, oidx, od2, shortrows, misc, tiles, offs, bboxes;
  hipEvent_t ev[8] = {};
  hipEvent_t ev_join = nullptr;  // orders rtnn_search_async work with the context stream
  // The query ordering runs on st2 while the point grid is sorted on st:
  // ev_desc (grid descriptors uploaded) starts it, ev_q joins it back.
  hipStream_t st2 = nullptr;
  hipEvent_t ev_desc = nullptr, ev_q = nullptr;
  DevBuf tmp2;  // radix-sort scratch of the query ordering

  // radius proposals (learned / partitioned)
  DevBuf pgrid, pcounts, qradius;
  DevBuf ovf;  // pooled kernel: counter + rows whose hit buffers overflowed
  GridDesc hpgrid{};
  DevBuf tr_q, tr_idx, tr_d2, tr_t, tr_f, tr_out;
  MlpState mlp;

  // graphs captured on this context (destroyed with it)
  std::vector<rtnn_graph_s*> graphs;

  // boundary-pair report
  DevBuf bcnt, boff, bpairs;
  uint32_t batch_tiles = 0;  // tiles of the last batched grid

  // batched size classes (rtnn_set_size_classes): class j holds the clouds of
  // at most class_upper[j] points (the last class: any size) and is searched
  // with class_opt[j]; class_tiles[j] .. [j+1] are its tiles in the last
  // batched grid (tiles are class-major)
  std::vector<uint64_t> class_upper;
  std::vector<rtnn_options> class_opt;
  std::vector<uint32_t> class_tiles;

  // the device address of this device's copy of `scratch` (rtnn_search_noalloc)
  void* scratch = nullptr;
  hipError_t scratch_err = hipSuccess;

  // streamed search (rtnn_search_stream): two slots of device and pinned
  // buffers, a copy stream and its events
  DevBuf sq[2], sidx[2], sd2[2];
  PinBuf pq[2], pidx[2], pd2[2];
  hipStream_t st3 = nullptr;
  hipEvent_t ev_up[2] = {}, ev_done[2] = {}, ev_down[2] = {};

  ~rtnn_ctx_s() {
    DevBuf* all[] = {&pts, &keys, &vals, &counts, &cell_start, &spts, &tmp, &grids,
                     &qin, &qkeys, &qvals, &qsorted, &oidx, &od2, &shortrows, &misc,
                     &tiles, &offs, &bboxes, &pgrid, &pcounts, &qradius, &tr_q, &tr_idx,
                     &tr_d2, &tr_t, &tr_f, &tr_out, &bcnt, &boff, &bpairs, &cboxes, &ovf};
    for (DevBuf* b : all) b->release();
    if (st3) (void)hipStreamSynchronize(st3);
    for (int i = 0; i < 2; ++i) {
      sq[i].release();
      sidx[i].release();
      sd2[i].release();
      pq[i].release();
      pidx[i].release();
      pd2[i].release();
      for (hipEvent_t e : {ev_up[i], ev_done[i], ev_down[i]})
        if (e) (void)hipEventDestroy(e);
    }
    if (st3) (void)hipStreamDestroy(st3);
    mlp_release(mlp);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (ev_desc) (void)hipEventDestroy(ev_desc);
    if (ev_q) (void)hipEventDestroy(ev_q);
    tmp2.release();
    if (st2) (void)hipStreamDestroy(st2);
    if (st) (void)hipStreamDestroy(st);
  }
};

// Destroys the graphs still alive on ctx (defined with rtnn_graph_s).
static void destroy_graphs(rtnn_ctx ctx);

namespace {

int fail(rtnn_ctx c, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int hip_fail(rtnn_ctx c, hipError_t e, const char* what) {
  int code = e == hipErrorOutOfMemory ? RTNN_ERR_OOM : RTNN_ERR_HIP;
  return fail(c, code, "%s: %s", what, hipGetErrorString(e));
}

#define CK(expr)                                           \
  do {                                                     \
    hipError_t e_ = (expr);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

int bits_for(uint64_t maxval) {
  int b = 0;
  while (b < 64 && (maxval >> b) != 0) ++b;
  return b;
}

// Grid plan for one cloud: dims = floor((max - min) / cell) + 1, computed in
// FP32 exactly as the device computes cell coordinates, so every point's cell
// lies inside the grid.
void plan_grid(const float* bb, float cell, GridDesc& g) {
  std::memset(&g, 0, sizeof g);
  g.cell = cell;
  g.inv = 1.0f / cell;
  for (int a = 0; a < 3; ++a) {
    g.mn[a] = bb[a];
    volatile float ext = bb[a + 3] - bb[a];
    volatile float e = ext * g.inv;
    float d = std::floor((float)e) + 1.0f;
    g.dim[a] = (int)std::min(d, 2.0e9f);
  }
}

// Cell count, saturated at 1e15 (dims up to 2e9 per axis would wrap 64 bits).
uint64_t grid_cells(const GridDesc& g) {
  const double c = (double)g.dim[0] * (double)g.dim[1] * (double)g.dim[2];
  if (c > 1e15) return (uint64_t)1e15;
  return (uint64_t)g.dim[0] * (uint64_t)g.dim[1] * (uint64_t)g.dim[2];
}

uint64_t max_cells_for(rtnn_ctx ctx, uint64_t npts) {
  if (ctx->opt.max_cells) return ctx->opt.max_cells;
  uint64_t m = std::max<uint64_t>(1ull << 22, 32ull * npts);
  return std::min<uint64_t>(m, 1ull << 30);
}

float default_cell_factor(int mode) { return mode == RTNN_MODE_KNN ? 1.0f : 1.0f; }

int default_group_lanes(int mode, uint32_t k) {
  if (mode == RTNN_MODE_KNN && k <= 64) return 1;
  if (k <= 64) return 16;
  if (k <= 256) return 16;
  return 64;
}

// Group kernel LDS buffer: compact (k + 2G keys per query) only on request.
bool wide_buffers(rtnn_ctx ctx) { return ctx->opt.group_buffer != 1; }
// Pooled kernel (group_buffer 3): queries_per_wave 16, 32 or 64 queries per
// wave share one candidate pool; k <= 64 and a raster grid (else the group
// kernel of queries_per_wave runs, with wide buffers).
bool pool_requested(rtnn_ctx ctx, uint32_t k) {
  const int q = ctx->opt.queries_per_wave;
  return ctx->opt.group_buffer == 3 && k <= 64 && (q == 16 || q == 32 || q == 64);
}
// Points per wave of the LDS candidate tile (0: no tile; option 1 = 512).
int tile_points(rtnn_ctx ctx) { return ctx->opt.lds_tile == 1 ? 512 : ctx->opt.lds_tile; }

int group_lanes_from_opt(rtnn_ctx ctx, int mode, uint32_t k) {
  int qpw = ctx->opt.queries_per_wave;
  if (qpw <= 0) return default_group_lanes(mode, k);
  int gl = 64 / qpw;
  if (gl != 1 && gl != 2 && gl != 4 && gl != 8 && gl != 16 && gl != 64) return -1;
  return gl;
}

// Builds the single-cloud grid over ctx->pts with edge `cell` (grown until the
// cell count fits the budget).
int build_grid(rtnn_ctx ctx, float cell) {
  GridDesc g;
  const uint64_t cap = max_cells_for(ctx, ctx->n);
  for (int tries = 0;; ++tries) {
    plan_grid(ctx->bbox, cell, g);
    if (grid_cells(g) <= cap && grid_cells(g) < 0xfffffff0ull) break;
    if (tries > 200) return fail(ctx, RTNN_ERR_UNSUPPORTED, "grid planning failed");
    cell *= 1.25f;
  }
  g.cell_base = 0;
  g.point_base = 0;
  g.npoints = (uint32_t)ctx->n;
  ctx->hgrid = g;
  ctx->ncells = grid_cells(g);
  const uint64_t n = ctx->n;
  CK(ctx->grids.ensure(sizeof(GridDesc)));
  CK(ctx->keys.ensure(4 * n));
  CK(ctx->vals.ensure(4 * n));
  CK(ctx->cell_start.ensure(4 * (ctx->ncells + 1)));
  CK(ctx->spts.ensure(16 * std::max<uint64_t>(n, 1) + kSptsPadBytes));
  CK(ctx->tmp.ensure(4 * std::max(radix_tmp_words(n), cell_bounds_tmp_words(ctx->ncells))));
  // The descriptor goes to the device as the key kernel's argument (no
  // pageable copy, which costs the host more than a launch).
  CK(cell_keys_single(ctx->pts.as<float>(), n, ctx->hgrid, ctx->grids.as<GridDesc>(),
                      ctx->keys.as<uint32_t>(), ctx->vals.as<uint32_t>(), ctx->st));
  CK(hipEventRecord(ctx->ev_desc, ctx->st));
  CK(radix_sort_pairs(ctx->keys.as<uint32_t>(), ctx->vals.as<uint32_t>(), n,
                      bits_for(ctx->ncells - 1), ctx->tmp.as<uint32_t>(), ctx->st,
                      &ctx->sort_passes));
  CK(cell_bounds(ctx->keys.as<uint32_t>(), n, (uint32_t)ctx->ncells,
                 ctx->cell_start.as<uint32_t>(), ctx->tmp.as<uint32_t>(), ctx->st));
  CK(gather_points(ctx->pts.as<float>(), ctx->vals.as<uint32_t>(), n, ctx->spts.as<float4>(),
                   ctx->st));
  ctx->spts_current = true;
  ctx->hlevels = 0;
  return RTNN_OK;
}

// Hierarchical Morton grid over ctx->pts: level l has cell edge r * 2^(l-L+1)
// (the top level's edge is exactly r) on a cubic grid of 2^(b-l) cells per
// axis.  Points are sorted once by their level-0 Morton key; a level-l cell is
// the key range sharing key >> 3l, so every level's table is built from the
// same sorted keys.  Level tables larger than 2^27 cells are skipped (queries
// use the finest built level).
int build_hgrid(rtnn_ctx ctx, float r, int want_levels) {
  int L = std::max(2, std::min(want_levels, 8));
  float c0 = 0, inv0 = 0;
  int b = 0;
  for (;;) {
    c0 = std::ldexp(r, -(L - 1));
    inv0 = 1.0f / c0;
    int dmax = 1;
    for (int a = 0; a < 3; ++a) {
      volatile float ext = ctx->bbox[a + 3] - ctx->bbox[a];
      volatile float e = ext * inv0;
      dmax = std::max(dmax, (int)std::min(std::floor((float)e) + 1.0f, 2.0e9f));
    }
    b = std::max(bits_for((uint64_t)(dmax - 1)), L - 1);
    if (b <= 10) break;
    if (--L < 1) return fail(ctx, RTNN_ERR_UNSUPPORTED, "radius too small for a Morton grid");
  }
  const uint64_t kMaxTable = 1ull << 27;
  int lmin = 0;
  while (lmin < L - 1 && (1ull << (3 * (b - lmin))) > kMaxTable) ++lmin;
  std::vector<GridDesc> lv(L);
  uint64_t total = 0;
  for (int l = 0; l < L; ++l) {
    GridDesc& g = lv[l];
    std::memset(&g, 0, sizeof g);
    for (int a = 0; a < 3; ++a) {
      g.mn[a] = ctx->bbox[a];
      g.dim[a] = 1 << (b - l);
    }
    g.cell = std::ldexp(c0, l);
    g.inv = std::ldexp(inv0, -l);  // exactly 2^-l * inv0: same cells as key >> 3l
    g.npoints = (uint32_t)ctx->n;
    g.morton = 1;
    if (l >= lmin) {
      g.cell_base = (uint32_t)total;
      total += (1ull << (3 * (b - l))) + 1;
    }
  }
  if (total >= 0xfffffff0ull) return fail(ctx, RTNN_ERR_UNSUPPORTED, "Morton tables too large");
  const uint64_t n = ctx->n;
  ctx->hgrid = lv[0];
  CK(ctx->grids.ensure(sizeof(GridDesc) * L));
  CK(hipMemcpyAsync(ctx->grids.p, lv.data(), sizeof(GridDesc) * L, hipMemcpyHostToDevice,
                    ctx->st));
  CK(hipEventRecord(ctx->ev_desc, ctx->st));
  CK(ctx->keys.ensure(4 * n));
  CK(ctx->vals.ensure(4 * n));
  CK(ctx->cell_start.ensure(4 * total));
  CK(ctx->spts.ensure(16 * std::max<uint64_t>(n, 1) + kSptsPadBytes));
  CK(ctx->tmp.ensure(
      4 * std::max(radix_tmp_words(n), cell_bounds_tmp_words(1ull << (3 * (b - lmin))))));
  CK(morton_keys(ctx->pts.as<float>(), n, ctx->grids.as<GridDesc>(), ctx->keys.as<uint32_t>(),
                 ctx->vals.as<uint32_t>(), ctx->st));
  CK(radix_sort_pairs(ctx->keys.as<uint32_t>(), ctx->vals.as<uint32_t>(), n, 3 * b,
                      ctx->tmp.as<uint32_t>(), ctx->st, &ctx->sort_passes));
  for (int l = lmin; l < L; ++l)
    CK(cell_bounds(ctx->keys.as<uint32_t>(), n, (uint32_t)(1ull << (3 * (b - l))),
                   ctx->cell_start.as<uint32_t>() + lv[l].cell_base, ctx->tmp.as<uint32_t>(),
                   ctx->st, 3 * l));
  CK(gather_points(ctx->pts.as<float>(), ctx->vals.as<uint32_t>(), n, ctx->spts.as<float4>(),
                   ctx->st));
  CK(ctx->cboxes.ensure(32 * ((n + kChunkPts - 1) / kChunkPts) + 32));
  CK(chunk_boxes(ctx->spts.as<float4>(), n, ctx->cboxes.as<float4>(), ctx->st));
  ctx->ncells = total;
  ctx->spts_current = true;
  ctx->hlevels = L;
  ctx->hlevel_min = lmin;
  return RTNN_OK;
}

constexpr int kQueryKeyBits = 8;  // Morton bits per axis of the query order

// Prepares queries in processing order (float4 with the output row in w).
// Returns the processing-order array through *qproc.  The ordering runs on
// ctx->st2, started by ctx->ev_desc (recorded by the grid build once its
// descriptors are on the device), so it overlaps the point sort on ctx->st;
// ctx->st waits for it before returning work that reads *qproc.
int prepare_queries(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                    const float4** qproc) {
  if (!qxyz) {  // the points themselves, already in cell order
    *qproc = ctx->spts.as<float4>();
    return RTNN_OK;
  }
  CK(ctx->qkeys.ensure(4 * nq));
  CK(ctx->qvals.ensure(4 * nq));
  CK(ctx->qsorted.ensure(16 * nq));
  CK(ctx->tmp2.ensure(4 * radix_tmp_words(nq)));
  if (!on_device) CK(ctx->qin.ensure(12 * nq));
  hipStream_t s2 = ctx->st2;
  CK(hipStreamWaitEvent(s2, ctx->ev_desc, 0));
  const float* dq = qxyz;
  if (!on_device) {
    CK(hipMemcpyAsync(ctx->qin.p, qxyz, 12 * nq, hipMemcpyHostToDevice, s2));
    dq = ctx->qin.as<float>();
  }
  if (ctx->opt.query_order == 1) {
    CK(iota_u32(ctx->qvals.as<uint32_t>(), nq, s2));
  } else {
    // Morton order of the query cells, 8 bits per axis (cells merged in
    // aligned 2^shift blocks on finer grids): 3 radix passes, and a wave's
    // queries still share their neighbourhood.
    int maxdim = std::max(ctx->hgrid.dim[0], std::max(ctx->hgrid.dim[1], ctx->hgrid.dim[2]));
    const int bits = bits_for((uint64_t)(maxdim - 1));
    const int shift = std::max(0, bits - kQueryKeyBits);
    CK(query_morton_keys(dq, nq, ctx->grids.as<GridDesc>(), shift, ctx->qkeys.as<uint32_t>(),
                         ctx->qvals.as<uint32_t>(), s2));
    CK(radix_sort_pairs(ctx->qkeys.as<uint32_t>(), ctx->qvals.as<uint32_t>(), nq,
                        3 * std::min(bits, kQueryKeyBits), ctx->tmp2.as<uint32_t>(), s2,
                        nullptr));
  }
  CK(gather_points(dq, ctx->qvals.as<uint32_t>(), nq, ctx->qsorted.as<float4>(), s2));
  CK(hipEventRecord(ctx->ev_q, s2));
  CK(hipStreamWaitEvent(ctx->st, ctx->ev_q, 0));
  *qproc = ctx->qsorted.as<float4>();
  return RTNN_OK;
}

double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return ms;
}

int check_search_args(rtnn_ctx ctx, int mode, float r, uint32_t k, const void* out_idx) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (mode != RTNN_MODE_RADIUS && mode != RTNN_MODE_KNN)
    return fail(ctx, RTNN_ERR_INVALID, "mode must be RTNN_MODE_RADIUS or RTNN_MODE_KNN");
  if (!(r > 0.0f) || !std::isfinite(r)) return fail(ctx, RTNN_ERR_INVALID, "radius must be > 0");
  if (k == 0) return fail(ctx, RTNN_ERR_INVALID, "k must be >= 1");
  if (!out_idx) return fail(ctx, RTNN_ERR_INVALID, "out_idx is NULL");
  return RTNN_OK;
}

struct OutBufs {
  int32_t* idx;
  float* d2;
};

int output_targets(rtnn_ctx ctx, uint64_t nq, uint32_t k, int32_t* out_idx, float* out_d2,
                   int out_on_device, OutBufs& o) {
  if (out_on_device) {
    o.idx = out_idx;
    o.d2 = out_d2;
    return RTNN_OK;
  }
  CK(ctx->oidx.ensure(4 * nq * k));
  o.idx = ctx->oidx.as<int32_t>();
  o.d2 = nullptr;
  if (out_d2) {
    CK(ctx->od2.ensure(4 * nq * k));
    o.d2 = ctx->od2.as<float>();
  }
  return RTNN_OK;
}

int copy_out(rtnn_ctx ctx, uint64_t nq, uint32_t k, int32_t* out_idx, float* out_d2,
             int out_on_device, const OutBufs& o) {
  if (out_on_device) return RTNN_OK;
  CK(hipMemcpyAsync(out_idx, o.idx, 4 * nq * k, hipMemcpyDeviceToHost, ctx->st));
  if (out_d2) CK(hipMemcpyAsync(out_d2, o.d2, 4 * nq * k, hipMemcpyDeviceToHost, ctx->st));
  return RTNN_OK;
}

SearchArgs base_args(rtnn_ctx ctx, const float4* qproc, uint64_t nq, float r, uint32_t k,
                     const OutBufs& o) {
  SearchArgs a;
  std::memset(&a, 0, sizeof a);
  a.spts = ctx->spts.as<float4>();
  a.cell_start = ctx->cell_start.as<uint32_t>();
  a.grids = ctx->grids.as<GridDesc>();
  a.queries = qproc;
  a.tiles = nullptr;
  a.nq = (uint32_t)nq;
  a.ntiles = (uint32_t)((nq + kTileQ - 1) / kTileQ);
  a.r = r;
  a.qradius_scale = 1.0f;
  a.k = k;
  a.out_idx = o.idx;
  a.out_d2 = o.d2;
  if (ctx->hlevels > 0) {
    a.nlevels = ctx->hlevels;
    a.level_min = ctx->hlevel_min;
    a.level_alpha = ctx->opt.level_alpha > 0 ? ctx->opt.level_alpha : 1.0f;
    a.chunk_box = ctx->cboxes.as<float4>();
    a.tree_leaf = ctx->opt.tree_leaf > 0 ? (uint32_t)ctx->opt.tree_leaf : kTreeLeafDefault;
    a.tree_lanes = ctx->opt.tree_lanes > 1 ? (uint32_t)ctx->opt.tree_lanes : 1u;
  }
  return a;
}

// The pooled kernel's overflow list (nq rows) and its counter, cleared.
int arm_overflow(rtnn_ctx ctx, uint64_t nq, SearchArgs& a) {
  CK(ctx->ovf.ensure(4 * (nq + 64)));
  a.ovf_count = ctx->ovf.as<uint32_t>();
  a.ovf_rows = ctx->ovf.as<uint32_t>() + 64;
  CK(fill_u32(a.ovf_count, 1, 0, ctx->st));
  return RTNN_OK;
}

void fill_stats(rtnn_ctx ctx, rtnn_stats* s) {
  if (!s) return;
  for (int a = 0; a < 3; ++a) s->grid_dims[a] = (uint32_t)ctx->hgrid.dim[a];
  s->cell_size = ctx->hgrid.cell;
  s->sort_passes = (uint32_t)ctx->sort_passes;
}

}  // namespace

// ===================================================================== ABI

extern "C" {

const char* rtnn_version(void) { return "rtnn-mi355x 0.1 (gfx950)"; }

int rtnn_mem_stats(uint64_t* device_bytes, uint64_t* device_peak, uint64_t* pinned_bytes,
                   uint64_t* pinned_peak) {
  if (device_bytes) *device_bytes = rtnn::mem_dev_cur();
  if (device_peak) *device_peak = rtnn::mem_dev_peak();
  if (pinned_bytes) *pinned_bytes = rtnn::mem_pin_cur();
  if (pinned_peak) *pinned_peak = rtnn::mem_pin_peak();
  return RTNN_OK;
}

const char* rtnn_create_error(void) { return g_create_err.c_str(); }

int rtnn_create(int device, rtnn_ctx* out) {
  if (!out) return RTNN_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    g_create_err = std::string("no HIP device: ") + hipGetErrorString(e);
    return RTNN_ERR_HIP;
  }
  if (device < 0 || device >= count) {
    g_create_err = "device id out of range";
    return RTNN_ERR_INVALID;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return RTNN_ERR_HIP;
  }
  rtnn_ctx c = new (std::nothrow) rtnn_ctx_s();
  if (!c) return RTNN_ERR_OOM;
  c->device = device;
  e = hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    delete c;
    return RTNN_ERR_HIP;
  }
  e = hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking);
  if (e != hipSuccess) {
    g_create_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    delete c;
    return RTNN_ERR_HIP;
  }
  for (auto& ev : c->ev) (void)hipEventCreate(&ev);
  (void)hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
  (void)hipEventCreateWithFlags(&c->ev_desc, hipEventDisableTiming);
  (void)hipEventCreateWithFlags(&c->ev_q, hipEventDisableTiming);
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      cus > 0)
    c->num_cus = cus;
  // Set-up of the allocation-free search: this device's scratch address.  A
  // failure is kept and reported by rtnn_search_noalloc only.
  c->scratch_err = scratch_address(device, &c->scratch);
  if (c->scratch_err != hipSuccess) (void)hipGetLastError();
  *out = c;
  return RTNN_OK;
}

int rtnn_destroy(rtnn_ctx ctx) {
  if (!ctx) return RTNN_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->st);
  destroy_graphs(ctx);
  delete ctx;
  return RTNN_OK;
}

const char* rtnn_last_error(rtnn_ctx ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"

namespace {
int check_options(rtnn_ctx ctx, const rtnn_options* opt);
}  // namespace

extern "C" {

int rtnn_set_options(rtnn_ctx ctx, const rtnn_options* opt) {
  if (!ctx || !opt) return RTNN_ERR_INVALID;
  const int rc = check_options(ctx, opt);
  if (rc) return rc;
  ctx->opt = *opt;
  return RTNN_OK;
}

int rtnn_set_size_classes(rtnn_ctx ctx, uint32_t nclasses, const uint64_t* upper,
                          const rtnn_options* opts) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (nclasses == 0) {
    ctx->class_upper.clear();
    ctx->class_opt.clear();
    return RTNN_OK;
  }
  if (nclasses > 16 || !opts || (nclasses > 1 && !upper))
    return fail(ctx, RTNN_ERR_INVALID, "1..16 classes with %s", "bounds and options");
  for (uint32_t j = 0; j + 1 < nclasses; ++j)
    if (j > 0 && upper[j] <= upper[j - 1])
      return fail(ctx, RTNN_ERR_INVALID, "class bounds must increase");
  for (uint32_t j = 0; j < nclasses; ++j) {
    const int rc = check_options(ctx, &opts[j]);
    if (rc) return rc;
  }
  ctx->class_upper.assign(upper, upper + (nclasses - 1));
  ctx->class_opt.assign(opts, opts + nclasses);
  return RTNN_OK;
}

int rtnn_get_size_classes(rtnn_ctx ctx, uint32_t* nclasses, uint64_t* upper,
                          rtnn_options* opts) {
  if (!ctx || !nclasses) return RTNN_ERR_INVALID;
  const uint32_t n = (uint32_t)ctx->class_opt.size();
  if (upper && n > 1) std::memcpy(upper, ctx->class_upper.data(), 8 * (n - 1));
  if (opts && n) std::memcpy(opts, ctx->class_opt.data(), sizeof(rtnn_options) * n);
  *nclasses = n;
  return RTNN_OK;
}

}  // extern "C"

namespace {

int check_options(rtnn_ctx ctx, const rtnn_options* opt) {
  if (opt->cell_factor < 0 || !std::isfinite(opt->cell_factor))
    return fail(ctx, RTNN_ERR_INVALID, "cell_factor must be >= 0");
  const int qpw = opt->queries_per_wave;
  if (qpw != 0 && qpw != 1 && qpw != 4 && qpw != 8 && qpw != 16 && qpw != 32 && qpw != 64)
    return fail(ctx, RTNN_ERR_INVALID, "queries_per_wave must be 0, 1, 4, 8, 16, 32 or 64");
  if (opt->query_order < 0 || opt->query_order > 2)
    return fail(ctx, RTNN_ERR_INVALID, "query_order must be 0, 1 or 2");
  if (opt->grid_levels < 0 || opt->grid_levels > 8)
    return fail(ctx, RTNN_ERR_INVALID, "grid_levels must be in [0, 8]");
  if (opt->group_buffer < 0 || opt->group_buffer > 3)
    return fail(ctx, RTNN_ERR_INVALID, "group_buffer must be 0, 1, 2 or 3");
  if (opt->lds_tile != 0 && opt->lds_tile != 1 && opt->lds_tile != 256 && opt->lds_tile != 512 &&
      opt->lds_tile != 1024)
    return fail(ctx, RTNN_ERR_INVALID, "lds_tile must be 0, 1 (= 512), 256, 512 or 1024");
  if (!(opt->level_alpha >= 0.0f) || !std::isfinite(opt->level_alpha))
    return fail(ctx, RTNN_ERR_INVALID, "level_alpha must be >= 0");
  if (opt->tree_leaf < 0) return fail(ctx, RTNN_ERR_INVALID, "tree_leaf must be >= 0");
  if (!(opt->bucket_ratio >= 0.0f && opt->bucket_ratio <= 1.0f))
    return fail(ctx, RTNN_ERR_INVALID, "bucket_ratio must be in [0, 1]");
  if (opt->learned_cap != 0 && opt->learned_cap != 1)
    return fail(ctx, RTNN_ERR_INVALID, "learned_cap must be 0 or 1");
  if (opt->tree_fallback != 0 && opt->tree_fallback != 1)
    return fail(ctx, RTNN_ERR_INVALID, "tree_fallback must be 0 or 1");
  if (opt->tree_lanes < 0 || opt->tree_lanes > 8 || (opt->tree_lanes & (opt->tree_lanes - 1)))
    return fail(ctx, RTNN_ERR_INVALID, "tree_lanes must be 0, 1, 2, 4 or 8");
  return RTNN_OK;
}

}  // namespace

extern "C" {

int rtnn_synchronize(rtnn_ctx ctx) {
  if (!ctx) return RTNN_ERR_INVALID;
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

int rtnn_set_points(rtnn_ctx ctx, const float* xyz, uint64_t n, int on_device) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (n == 0) return fail(ctx, RTNN_ERR_INVALID, "empty point set");
  if (!xyz) return fail(ctx, RTNN_ERR_INVALID, "xyz is NULL");
  if (n >= 0xffffffffull) return fail(ctx, RTNN_ERR_UNSUPPORTED, "more than 2^32-2 points");
  CK(hipSetDevice(ctx->device));
  CK(ctx->pts.ensure(12 * n));
  CK(hipMemcpyAsync(ctx->pts.p, xyz, 12 * n,
                    on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->st));
  ctx->n = n;
  ctx->spts_current = false;
  CK(ctx->misc.ensure(4096 + 6 * 4 * kBboxBlocks));
  float* bb = ctx->misc.as<float>();
  uint32_t* bad = reinterpret_cast<uint32_t*>(bb + 8);
  float* part = bb + 16;
  CK(fill_u32(bad, 1, 0, ctx->st));
  CK(bbox(ctx->pts.as<float>(), n, part, bb, bad, ctx->st));
  float hb[16];
  CK(hipMemcpyAsync(hb, bb, 64, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  uint32_t nbad;
  std::memcpy(&nbad, &hb[8], 4);
  if (nbad) {
    ctx->n = 0;
    return fail(ctx, RTNN_ERR_INVALID, "%u non-finite point coordinates", nbad);
  }
  std::memcpy(ctx->bbox, hb, sizeof ctx->bbox);
  return RTNN_OK;
}

static int search_impl(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, int mode,
                       float r, uint32_t k, int32_t* out_idx, float* out_d2, int out_on_device,
                       rtnn_stats* stats, bool wait = true) {
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  if (!qxyz) nq = ctx->n;
  if (nq == 0) return RTNN_OK;
  if (nq >= 0xffffffffull) return fail(ctx, RTNN_ERR_UNSUPPORTED, "more than 2^32-2 queries");
  const int gl = group_lanes_from_opt(ctx, mode, k);
  const bool pooled = pool_requested(ctx, k);
  if (!pooled && (gl < 0 || k > max_k_for_group(gl))) {
    if (ctx->opt.queries_per_wave > 0 && k <= 1024)
      return fail(ctx, RTNN_ERR_UNSUPPORTED, "k=%u is too large for queries_per_wave=%d", k,
                  ctx->opt.queries_per_wave);
    return fail(ctx, RTNN_ERR_UNSUPPORTED, "k=%u exceeds the supported maximum (1024)", k);
  }
  CK(hipSetDevice(ctx->device));
  const bool timing = wait && (ctx->opt.time_stages != 0 || stats != nullptr);
  if (timing) CK(hipEventRecord(ctx->ev[0], ctx->st));
  float f = ctx->opt.cell_factor > 0 ? ctx->opt.cell_factor : default_cell_factor(mode);
  rc = build_grid(ctx, r / f);
  if (rc) return rc;
  if (timing) CK(hipEventRecord(ctx->ev[1], ctx->st));
  const float4* qproc = nullptr;
  rc = prepare_queries(ctx, qxyz, nq, on_device, &qproc);
  if (rc) return rc;
  if (timing) CK(hipEventRecord(ctx->ev[2], ctx->st));
  OutBufs o;
  rc = output_targets(ctx, nq, k, out_idx, out_d2, out_on_device, o);
  if (rc) return rc;
  SearchArgs a = base_args(ctx, qproc, nq, r, k, o);
  if (pooled) {
    rc = arm_overflow(ctx, nq, a);
    if (rc) return rc;
  }
  CK(launch_search(a, mode == RTNN_MODE_KNN, gl, ctx->st, ctx->num_cus, wide_buffers(ctx),
                   tile_points(ctx), false, pooled ? ctx->opt.queries_per_wave : 0));
  if (timing) CK(hipEventRecord(ctx->ev[3], ctx->st));
  rc = copy_out(ctx, nq, k, out_idx, out_d2, out_on_device, o);
  if (rc) return rc;
  if (!wait) return RTNN_OK;
  if (timing) CK(hipEventRecord(ctx->ev[4], ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->ms_grid = ev_ms(ctx->ev[0], ctx->ev[1]);
    stats->ms_qsort = ev_ms(ctx->ev[1], ctx->ev[2]);
    stats->ms_search = ev_ms(ctx->ev[2], ctx->ev[3]);
    stats->ms_copy = out_on_device ? 0.0 : ev_ms(ctx->ev[3], ctx->ev[4]);
    stats->ms_total = ev_ms(ctx->ev[0], ctx->ev[4]);
    stats->search_launches = pooled ? 2 : 1;
    stats->lanes_per_query = pooled ? (uint32_t)(64 / ctx->opt.queries_per_wave) : (uint32_t)gl;
    fill_stats(ctx, stats);
  }
  return RTNN_OK;
}

int rtnn_search(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, int mode, float r,
                uint32_t k, int32_t* out_idx, float* out_d2, int out_on_device,
                rtnn_stats* stats) {
  return search_impl(ctx, qxyz, nq, on_device, mode, r, k, out_idx, out_d2, out_on_device,
                     stats);
}

int rtnn_search_async(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int mode, float r,
                      uint32_t k, int32_t* out_idx, float* out_d2, void* stream) {
  if (!ctx) return RTNN_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  hipStream_t us = static_cast<hipStream_t>(stream);
  hipStream_t own = ctx->st;
  if (us == own) return search_impl(ctx, qxyz, nq, 1, mode, r, k, out_idx, out_d2, 1, nullptr,
                                    false);
  // The caller's stream waits for the context's earlier work; the context's
  // stream waits for this work, so later calls on the context are ordered
  // after it without a host synchronisation.
  CK(hipEventRecord(ctx->ev_join, own));
  CK(hipStreamWaitEvent(us, ctx->ev_join, 0));
  ctx->st = us;
  const int rc = search_impl(ctx, qxyz, nq, 1, mode, r, k, out_idx, out_d2, 1, nullptr, false);
  ctx->st = own;
  CK(hipEventRecord(ctx->ev_join, us));
  CK(hipStreamWaitEvent(own, ctx->ev_join, 0));
  return rc;
}

// The plain search of rtnn_set_points + rtnn_search with every buffer carved
// from this device's `scratch` (no allocation): the grid is planned on the
// device (bbox -> plan -> cell keys -> radix sort -> cell table -> gather),
// the queries are put in Morton order of their cells, then the search kernel
// writes the caller's rows.  Only the 16-byte status word is read back.
int rtnn_search_noalloc(rtnn_ctx ctx, const float* xyz, uint64_t n, const float* qxyz,
                        uint64_t nq, int mode, float r, uint32_t k, int32_t* out_idx,
                        float* out_d2) {
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  if (!xyz || n == 0) return fail(ctx, RTNN_ERR_INVALID, "xyz must be set and n >= 1");
  if (!qxyz) nq = n;
  if (nq == 0) return RTNN_OK;
  if (n >= (kScratchBytes >> 4) || nq >= (kScratchBytes >> 4))
    return fail(ctx, RTNN_ERR_UNSUPPORTED,
                "%llu points / %llu queries exceed the %llu MiB scratch of rtnn_search_noalloc",
                (unsigned long long)n, (unsigned long long)nq,
                (unsigned long long)(kScratchBytes >> 20));
  const int gl = group_lanes_from_opt(ctx, mode, k);
  const bool pooled = pool_requested(ctx, k);
  if (!pooled && (gl < 0 || k > max_k_for_group(gl)))
    return fail(ctx, RTNN_ERR_UNSUPPORTED, "k=%u is not supported by the selected kernel", k);
  if (ctx->scratch_err != hipSuccess || !ctx->scratch)
    return hip_fail(ctx, ctx->scratch_err != hipSuccess ? ctx->scratch_err : hipErrorUnknown,
                    "hipGetSymbolAddress(scratch)");
  // Carve the scratch (256-byte aligned pieces); sizes beyond it are refused.
  const bool own_q = qxyz != nullptr;
  const uint64_t sort_n = std::max(n, own_q ? nq : 0);
  uint64_t used = 0;
  auto take = [&](uint64_t bytes) -> uint64_t {
    const uint64_t at = used;
    used += (bytes + 255) & ~255ull;
    return at;
  };
  const uint64_t o_status = take(16), o_offs = take(16), o_part = take(4 * 6 * kBboxBlocks);
  const uint64_t o_bb = take(4 * 6), o_grid = take(sizeof(GridDesc));
  const uint64_t o_keys = take(4 * n), o_vals = take(4 * n);
  const uint64_t o_spts = take(16 * n + kSptsPadBytes);
  const uint64_t o_tmp = take(4 * std::max(radix_tmp_words(sort_n),
                                           cell_bounds_tmp_words(kScratchBytes / 4)));
  const uint64_t o_qkeys = own_q ? take(4 * nq) : 0, o_qvals = own_q ? take(4 * nq) : 0;
  const uint64_t o_qs = own_q ? take(16 * nq) : 0;
  const uint64_t o_ovf = pooled ? take(4 * (nq + 64)) : 0;
  const uint64_t o_cells = take(0);
  constexpr uint64_t kMinCells = 1ull << 16;
  if (o_cells + 4 * (kMinCells + 1) > kScratchBytes)
    return fail(ctx, RTNN_ERR_UNSUPPORTED,
                "%llu points and %llu queries exceed the %llu MiB scratch of "
                "rtnn_search_noalloc (about 40 B per point, 24 B per separate query, and "
                "the cell table)",
                (unsigned long long)n, (unsigned long long)(own_q ? nq : 0),
                (unsigned long long)(kScratchBytes >> 20));
  // The cell table takes the rest, up to the cell budget of rtnn_search.
  const uint64_t cap = std::min<uint64_t>(std::min<uint64_t>((kScratchBytes - o_cells) / 4 - 1,
                                                             max_cells_for(ctx, n)),
                                          0xfffffff0ull);
  unsigned char* base = static_cast<unsigned char*>(ctx->scratch);
  uint32_t* status = reinterpret_cast<uint32_t*>(base + o_status);
  uint64_t* offs = reinterpret_cast<uint64_t*>(base + o_offs);
  float* bb = reinterpret_cast<float*>(base + o_bb);
  GridDesc* grid = reinterpret_cast<GridDesc*>(base + o_grid);
  uint32_t* keys = reinterpret_cast<uint32_t*>(base + o_keys);
  uint32_t* vals = reinterpret_cast<uint32_t*>(base + o_vals);
  float4* spts = reinterpret_cast<float4*>(base + o_spts);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(base + o_tmp);
  uint32_t* cell_start = reinterpret_cast<uint32_t*>(base + o_cells);
  CK(hipSetDevice(ctx->device));
  hipStream_t st = ctx->st;
  // status: [0] non-finite coordinates, [1] planning fell back, [2] cells;
  // offsets of the one cloud: {0, n}
  CK(fill_u32(status, 4, 0, st));
  CK(fill_u32(reinterpret_cast<uint32_t*>(offs), 4, 0, st));
  CK(fill_u32(reinterpret_cast<uint32_t*>(offs) + 2, 1, (uint32_t)n, st));
  CK(bbox(xyz, n, reinterpret_cast<float*>(base + o_part), bb, &status[0], st));
  const float f = ctx->opt.cell_factor > 0 ? ctx->opt.cell_factor : default_cell_factor(mode);
  CK(plan_batched(bb, offs, 1, r / f, nullptr, cap, grid, status, st));
  CK(cell_keys(xyz, n, grid, offs, 1, keys, vals, nullptr, st));
  CK(radix_sort_pairs(keys, vals, n, bits_for(cap - 1), tmp, st, nullptr));
  CK(cell_bounds(keys, n, (uint32_t)cap, cell_start, tmp, st));
  CK(gather_points(xyz, vals, n, spts, st));
  const float4* qproc = spts;
  if (own_q) {
    uint32_t* qkeys = reinterpret_cast<uint32_t*>(base + o_qkeys);
    uint32_t* qvals = reinterpret_cast<uint32_t*>(base + o_qvals);
    float4* qs = reinterpret_cast<float4*>(base + o_qs);
    if (ctx->opt.query_order == 1) {
      CK(iota_u32(qvals, nq, st));
    } else {
      CK(query_morton_keys(qxyz, nq, grid, -1, qkeys, qvals, st));
      CK(radix_sort_pairs(qkeys, qvals, nq, 3 * kQueryKeyBits, tmp, st, nullptr));
    }
    CK(gather_points(qxyz, qvals, nq, qs, st));
    qproc = qs;
  }
  SearchArgs a;
  std::memset(&a, 0, sizeof a);
  a.spts = spts;
  a.cell_start = cell_start;
  a.grids = grid;
  a.queries = qproc;
  a.nq = (uint32_t)nq;
  a.r = r;
  a.qradius_scale = 1.0f;
  a.k = k;
  a.out_idx = out_idx;
  a.out_d2 = out_d2;
  if (pooled) {
    a.ovf_count = reinterpret_cast<uint32_t*>(base + o_ovf);
    a.ovf_rows = a.ovf_count + 64;
    CK(fill_u32(a.ovf_count, 1, 0, st));
  }
  CK(launch_search(a, mode == RTNN_MODE_KNN, gl, st, ctx->num_cus, wide_buffers(ctx),
                   tile_points(ctx), false, pooled ? ctx->opt.queries_per_wave : 0));
  uint32_t hs[4] = {0, 0, 0, 0};
  CK(hipMemcpyAsync(hs, status, 16, hipMemcpyDeviceToHost, st));
  CK(hipStreamSynchronize(st));
  if (hs[0])
    return fail(ctx, RTNN_ERR_INVALID, "%u non-finite point coordinates (rows undefined)", hs[0]);
  return RTNN_OK;
}

}  // extern "C"

namespace {

// Point-count grid over ctx->pts with edge `cell` (origin = points' bbox min).
// grow = false: the edge is kept (error if the grid would exceed the budget).
int build_count_grid(rtnn_ctx ctx, float cell, bool grow) {
  GridDesc g;
  const uint64_t cap = max_cells_for(ctx, ctx->n);
  for (int tries = 0;; ++tries) {
    plan_grid(ctx->bbox, cell, g);
    if (grid_cells(g) <= cap && grid_cells(g) < 0xfffffff0ull) break;
    if (!grow || tries > 200)
      return fail(ctx, RTNN_ERR_UNSUPPORTED, "count grid of edge %g needs too many cells", cell);
    cell *= 1.25f;
  }
  g.npoints = (uint32_t)ctx->n;
  ctx->hpgrid = g;
  const uint64_t nc = grid_cells(g);
  CK(ctx->pgrid.ensure(sizeof(GridDesc)));
  CK(hipMemcpyAsync(ctx->pgrid.p, &ctx->hpgrid, sizeof(GridDesc), hipMemcpyHostToDevice,
                    ctx->st));
  CK(ctx->pcounts.ensure(4 * nc));
  CK(fill_u32(ctx->pcounts.as<uint32_t>(), nc, 0, ctx->st));
  if (ctx->spts_current)  // points grouped by cell: one atomic per run
    CK(count_sorted(ctx->spts.as<float4>(), ctx->n, ctx->pgrid.as<GridDesc>(),
                    ctx->pcounts.as<uint32_t>(), ctx->st));
  else
    CK(count_cells(ctx->pts.as<float>(), ctx->n, ctx->pgrid.as<GridDesc>(),
                   ctx->pcounts.as<uint32_t>(), ctx->st));
  return RTNN_OK;
}

enum class Proposal { kLearned, kPartition };

// Grid for the proposal method (hierarchical Morton grid or one raster grid),
// queries in processing order (*qproc) and their launch radius in
// ctx->qradius; ctx->ev[1] is recorded after the grid and query sort.
// The density partition's launch radius per processing position into
// ctx->qradius (its grid already built: hierarchical or raster).
int partition_proposals(rtnn_ctx ctx, bool hier, const float4* qproc, uint64_t nq, float r,
                        uint32_t k) {
  if (hier) {
    CK(partition_hier(qproc, (uint32_t)nq, ctx->grids.as<GridDesc>(), ctx->hlevels,
                      ctx->hlevel_min, ctx->cell_start.as<uint32_t>(), k, r,
                      ctx->qradius.as<float>(), ctx->st));
    return RTNN_OK;
  }
  const int rc = build_count_grid(ctx, r / 8.0f, true);
  if (rc) return rc;
  CK(partition_radius(qproc, (uint32_t)nq, ctx->pgrid.as<GridDesc>(),
                      ctx->pcounts.as<uint32_t>(), k, r, ctx->qradius.as<float>(), nullptr, 0,
                      ctx->st));
  return RTNN_OK;
}

int compute_proposals(rtnn_ctx ctx, Proposal kind, const float* qxyz, uint64_t nq, int on_device,
                      float r, uint32_t k, const float4** qproc_out) {
  const float f = ctx->opt.cell_factor > 0 ? ctx->opt.cell_factor
                                           : (kind == Proposal::kLearned ? 2.0f : 4.0f);
  const int levels = ctx->opt.grid_levels > 0 ? ctx->opt.grid_levels
                                              : (kind == Proposal::kLearned ? 1 : 5);
  bool hier = levels >= 2;
  int rc = hier ? build_hgrid(ctx, r, levels) : build_grid(ctx, r / f);
  if (rc == RTNN_ERR_UNSUPPORTED && hier) {
    // The Morton grid needs the box within 1024 top-level cells of edge r per
    // axis; a wider box (or a smaller r) takes the raster grid, which grows
    // its cell edge to fit the cell budget.
    hier = false;
    rc = build_grid(ctx, r / f);
  }
  if (rc) return rc;
  const float4* qproc = nullptr;
  rc = prepare_queries(ctx, qxyz, nq, on_device, &qproc);
  if (rc) return rc;
  CK(hipEventRecord(ctx->ev[1], ctx->st));
  CK(ctx->qradius.ensure(4 * nq));
  if (kind == Proposal::kLearned) {
    // learned_cap: the density partition's radius first; the MLP's proposal
    // is then capped by it (that radius holds k neighbours, so a query it
    // bounds more tightly than the network never falls back).
    const bool cap = ctx->opt.learned_cap != 0;
    if (cap) {
      rc = partition_proposals(ctx, hier, qproc, nq, r, k);
      if (rc) return rc;
    }
    // Features: point counts of the edge-r cells.  On the hierarchical grid
    // they are the top level's table (end - start); otherwise a count grid.
    const GridDesc* coarse;
    const uint32_t* counts;
    if (hier) {
      coarse = ctx->grids.as<GridDesc>() + (ctx->hlevels - 1);
      counts = ctx->cell_start.as<uint32_t>();
    } else {
      rc = build_count_grid(ctx, r, false);
      if (rc) return rc;
      coarse = ctx->pgrid.as<GridDesc>();
      counts = ctx->pcounts.as<uint32_t>();
    }
    CK(mlp_propose(ctx->mlp, qproc, (uint32_t)nq, coarse, hier, counts, r, cap,
                   ctx->qradius.as<float>(), ctx->st, ctx->num_cus));
  } else {
    rc = partition_proposals(ctx, hier, qproc, nq, r, k);
    if (rc) return rc;
  }
  // Density buckets: proposals >= bucket_ratio * r are searched at r at once.
  if (ctx->opt.bucket_ratio > 0.0f)
    CK(bucket_snap(ctx->qradius.as<float>(), (uint32_t)nq, ctx->opt.bucket_ratio * r, r,
                   ctx->st));
  *qproc_out = qproc;
  return RTNN_OK;
}

// KNN with a proposed radius r' <= r per query, then an exact fallback: the
// rows with fewer than k hits within r' are searched again at r.  If a row
// has >= k hits with d2 < r'^2 <= r^2, any other point has d2 >= r'^2, so
// those k are the k nearest within r.
constexpr int kFallbackLanes = 4;  // lanes per query of the fallback stages
constexpr int kGrowStages = 4;     // radius-growth stages before the last one at r

int proposal_search(rtnn_ctx ctx, Proposal kind, const float* qxyz, uint64_t nq, int on_device,
                    float r, uint32_t k, int32_t* out_idx, float* out_d2, int out_on_device,
                    rtnn_stats* stats) {
  int rc = check_search_args(ctx, RTNN_MODE_KNN, r, k, out_idx);
  if (rc) return rc;
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  if (kind == Proposal::kLearned && !ctx->mlp.ready)
    return fail(ctx, RTNN_ERR_STATE, "the radius MLP has no weights (train or load it)");
  if (k > 1024) return fail(ctx, RTNN_ERR_UNSUPPORTED, "k=%u exceeds the supported maximum (1024)", k);
  if (!qxyz) nq = ctx->n;
  if (nq == 0) return RTNN_OK;
  if (nq >= 0xffffffffull) return fail(ctx, RTNN_ERR_UNSUPPORTED, "more than 2^32-2 queries");
  CK(hipSetDevice(ctx->device));
  CK(hipEventRecord(ctx->ev[0], ctx->st));
  const float4* qproc = nullptr;
  rc = compute_proposals(ctx, kind, qxyz, nq, on_device, r, k, &qproc);
  if (rc) return rc;
  CK(hipEventRecord(ctx->ev[2], ctx->st));
  OutBufs o;
  rc = output_targets(ctx, nq, k, out_idx, out_d2, out_on_device, o);
  if (rc) return rc;
  // Two short-row lists (positions into qproc) used in turn, the hits each
  // entry found, and one counter per list-producing stage.
  CK(ctx->shortrows.ensure(4 * (4 * nq + 64)));
  uint32_t* lists[2] = {ctx->shortrows.as<uint32_t>(), ctx->shortrows.as<uint32_t>() + nq};
  uint32_t* hits[2] = {lists[0] + 2 * nq, lists[0] + 3 * nq};
  uint32_t* counters = lists[0] + 4 * nq;
  CK(fill_u32(counters, kGrowStages + 1, 0, ctx->st));
  const int gl = group_lanes_from_opt(ctx, RTNN_MODE_KNN, k);
  const bool pooled = pool_requested(ctx, k) && ctx->hlevels == 0;
  if (!pooled && (gl < 0 || k > max_k_for_group(gl)))
    return fail(ctx, RTNN_ERR_UNSUPPORTED, "k=%u not supported by the selected kernel", k);
  // Stage 1: every query at its proposed radius r'.
  SearchArgs a = base_args(ctx, qproc, nq, r, k, o);
  a.qradius = ctx->qradius.as<float>();
  a.short_rows = lists[0];
  a.short_count = &counters[0];
  a.short_hits = hits[0];
  if (pooled) {
    rc = arm_overflow(ctx, nq, a);
    if (rc) return rc;
  }
  CK(launch_search(a, 1, gl, ctx->st, ctx->num_cus, wide_buffers(ctx), tile_points(ctx), false,
                   pooled ? ctx->opt.queries_per_wave : 0));
  CK(hipEventRecord(ctx->ev[3], ctx->st));
  // Growth stages: a row short at radius R (m < k hits) is searched again at
  // min(r, R * hit_scale(m, k)), hit_scale in [1.15, 2]; the radius it used is
  // stored for the next stage.  A row searched at r is final and leaves the
  // lists, and the last stage takes every remaining row at r.  These rows
  // are few and their radii large (many candidates each): four lanes per
  // query with wide, deep buffers (k <= 64 here; room for 2 max(k, 32) keys)
  // measured faster than one lane (C3: 1.25 -> 0.84 ms), two lanes, or the
  // k <= 16 buffer.
  // The octree walk (hierarchical grid, one lane per query) costs about as
  // much at a grown radius as at r -- its limit, not its radius, bounds the
  // walk -- so its short rows go straight to the last stage: on the
  // Levy-flight KNN 151 short rows took 27 ms through five stages, since
  // each stage waits for its slowest walk.  Where many rows fall back at
  // radii well below r the stages can still pay (option tree_fallback = 1,
  // which the autotuner tries).
  const bool tree_fallback = ctx->hlevels > 0 && k <= 64 && ctx->opt.tree_fallback == 0;
  for (int j = tree_fallback ? kGrowStages : 0; j <= kGrowStages; ++j) {
    SearchArgs s = base_args(ctx, qproc, nq, r, k, o);
    s.qsel = tree_fallback ? lists[0] : lists[j & 1];
    s.nq_dev = tree_fallback ? &counters[0] : &counters[j];
    if (j < kGrowStages) {
      s.qsel_hits = hits[j & 1];
      s.qradius = ctx->qradius.as<float>();
      s.radius_out = ctx->qradius.as<float>();
      s.short_rows = lists[(j + 1) & 1];
      s.short_hits = hits[(j + 1) & 1];
      s.short_count = &counters[j + 1];
    }
    // On the hierarchical grid the octree walk (one lane per query) prunes
    // by its shrinking limit, which pays most for these wide searches.
    // Beyond k = 64 (the register top-k and the 4-lane buffers) the rows go
    // to 16- or 64-lane groups.
    const int fl = k > 256 ? 64 : k > 64 ? 16 : ctx->hlevels > 0 ? 1 : kFallbackLanes;
    CK(launch_search(s, 1, fl, ctx->st, ctx->num_cus, true, false, true));
  }
  CK(hipEventRecord(ctx->ev[4], ctx->st));
  rc = copy_out(ctx, nq, k, out_idx, out_d2, out_on_device, o);
  if (rc) return rc;
  uint32_t nshort = 0;
  CK(hipMemcpyAsync(&nshort, &counters[0], 4, hipMemcpyDeviceToHost, ctx->st));
  CK(hipEventRecord(ctx->ev[5], ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->ms_grid = ev_ms(ctx->ev[0], ctx->ev[1]);
    stats->ms_mlp = ev_ms(ctx->ev[1], ctx->ev[2]);
    stats->ms_search = ev_ms(ctx->ev[2], ctx->ev[3]);
    stats->ms_fallback = ev_ms(ctx->ev[3], ctx->ev[4]);
    stats->ms_copy = out_on_device ? 0.0 : ev_ms(ctx->ev[4], ctx->ev[5]);
    stats->ms_total = ev_ms(ctx->ev[0], ctx->ev[5]);
    stats->fallback_queries = nshort;
    stats->search_launches = 2 + (tree_fallback ? 0 : kGrowStages) + (pooled ? 1 : 0);
    stats->lanes_per_query = pooled ? (uint32_t)(64 / ctx->opt.queries_per_wave) : (uint32_t)gl;
    fill_stats(ctx, stats);
  }
  return RTNN_OK;
}

}  // namespace

extern "C" {

int rtnn_search_partitioned(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device,
                            float r, uint32_t k, int32_t* out_idx, float* out_d2,
                            int out_on_device, rtnn_stats* stats) {
  return proposal_search(ctx, Proposal::kPartition, qxyz, nq, on_device, r, k, out_idx, out_d2,
                         out_on_device, stats);
}

int rtnn_search_learned(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, float r,
                        uint32_t k, int32_t* out_idx, float* out_d2, int out_on_device,
                        rtnn_stats* stats) {
  return proposal_search(ctx, Proposal::kLearned, qxyz, nq, on_device, r, k, out_idx, out_d2,
                         out_on_device, stats);
}

int rtnn_propose_radius(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, int method,
                        float r, uint32_t k, float* out_radius, int out_on_device) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (method != RTNN_METHOD_PARTITIONED && method != RTNN_METHOD_LEARNED)
    return fail(ctx, RTNN_ERR_INVALID, "method must be partitioned or learned");
  if (!(r > 0.0f) || !std::isfinite(r) || k == 0 || !out_radius)
    return fail(ctx, RTNN_ERR_INVALID, "invalid radius, k or output");
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  const Proposal kind = method == RTNN_METHOD_LEARNED ? Proposal::kLearned : Proposal::kPartition;
  if (kind == Proposal::kLearned && !ctx->mlp.ready)
    return fail(ctx, RTNN_ERR_STATE, "the radius MLP has no weights (train or load it)");
  if (!qxyz) nq = ctx->n;
  if (nq == 0) return RTNN_OK;
  if (nq >= 0xffffffffull) return fail(ctx, RTNN_ERR_UNSUPPORTED, "more than 2^32-2 queries");
  CK(hipSetDevice(ctx->device));
  const float4* qproc = nullptr;
  int rc = compute_proposals(ctx, kind, qxyz, nq, on_device, r, k, &qproc);
  if (rc) return rc;
  float* dst = out_radius;
  if (!out_on_device) {
    CK(ctx->od2.ensure(4 * nq));
    dst = ctx->od2.as<float>();
  }
  CK(scatter_rows(qproc, ctx->qradius.as<float>(), nq, r, dst, ctx->st));
  if (!out_on_device)
    CK(hipMemcpyAsync(out_radius, dst, 4 * nq, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

// ----------------------------------------------------- streamed search --

}  // extern "C"

namespace {
// Defined with the batched search below.
int check_offsets(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets, uint32_t nclouds);
int check_batched_lanes(rtnn_ctx ctx, int mode, uint32_t k);
int batched_impl(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets, uint32_t nclouds,
                 int mode, float r, uint32_t k, int32_t* out_idx, float* out_d2,
                 rtnn_stats* stats);
}  // namespace

extern "C" {

// Host queries in, host results (or a result file) out, in chunks of
// `chunk` queries through two slots: while chunk i is searched, chunk i-1's
// rows travel to pinned memory on a copy stream and chunk i-2's are copied
// from there to the caller's arrays or written to the file.  Each chunk is
// an ordinary search of its queries (rows = positions in the chunk), so the
// concatenated rows equal one search over all queries.
int rtnn_search_stream(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int mode, int method,
                       float r, uint32_t k, uint64_t chunk, int32_t* out_idx, float* out_d2,
                       const char* out_path, rtnn_stats* stats) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (!qxyz && nq > 0) return fail(ctx, RTNN_ERR_INVALID, "qxyz is NULL");
  if (!out_path && !out_idx) return fail(ctx, RTNN_ERR_INVALID, "no output (out_idx or out_path)");
  if (method != RTNN_METHOD_PLAIN && method != RTNN_METHOD_PARTITIONED &&
      method != RTNN_METHOD_LEARNED)
    return fail(ctx, RTNN_ERR_INVALID, "unknown method %d", method);
  if (method != RTNN_METHOD_PLAIN && mode != RTNN_MODE_KNN)
    return fail(ctx, RTNN_ERR_INVALID, "the partitioned and learned methods are KNN only");
  int rc = check_search_args(ctx, mode, r, k, out_path ? reinterpret_cast<const void*>(out_path)
                                                      : out_idx);
  if (rc) return rc;
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  if (chunk == 0) chunk = 1ull << 22;
  chunk = std::min<uint64_t>(chunk, std::max<uint64_t>(nq, 1));
  CK(hipSetDevice(ctx->device));
  if (!ctx->st3) {
    CK(hipStreamCreateWithFlags(&ctx->st3, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
      CK(hipEventCreateWithFlags(&ctx->ev_up[i], hipEventDisableTiming));
      CK(hipEventCreateWithFlags(&ctx->ev_done[i], hipEventDisableTiming));
      CK(hipEventCreateWithFlags(&ctx->ev_down[i], hipEventDisableTiming));
    }
  }
  // The result file is closed on every return (an early HIP error included).
  std::unique_ptr<FILE, int (*)(FILE*)> fguard(nullptr, &std::fclose);
  FILE* f = nullptr;
  const uint64_t head = 8 + 8 + 4 + 4;
  if (out_path) {
    f = std::fopen(out_path, "wb");
    if (!f) return fail(ctx, RTNN_ERR_IO, "cannot open %s for writing", out_path);
    fguard.reset(f);
    const uint32_t flags = 1;  // d2 present
    std::fwrite("RTNNRES1", 1, 8, f);
    std::fwrite(&nq, 8, 1, f);
    std::fwrite(&k, 4, 1, f);
    std::fwrite(&flags, 4, 1, f);
  }
  const bool want_d2 = out_path || out_d2;
  const uint64_t rowb = 4ull * k;
  for (int i = 0; i < 2; ++i) {
    rc = RTNN_OK;
    if (ctx->sq[i].ensure(12 * chunk) != hipSuccess || ctx->sidx[i].ensure(rowb * chunk) ||
        (want_d2 && ctx->sd2[i].ensure(rowb * chunk)) || ctx->pq[i].ensure(12 * chunk) ||
        ctx->pidx[i].ensure(rowb * chunk) || (want_d2 && ctx->pd2[i].ensure(rowb * chunk))) {
      return fail(ctx, RTNN_ERR_OOM, "streaming buffers of %llu queries",
                  (unsigned long long)chunk);
    }
  }
  const uint64_t nchunks = (nq + chunk - 1) / chunk;
  rtnn_stats sum{};
  bool io_err = false;
  // Rows of chunk c (slot c & 1) from pinned memory to the caller / file.
  auto flush = [&](uint64_t c) -> hipError_t {
    const int s = (int)(c & 1);
    hipError_t e = hipEventSynchronize(ctx->ev_down[s]);
    if (e != hipSuccess) return e;
    const uint64_t c0 = c * chunk, n = std::min(chunk, nq - c0);
    if (f) {
      io_err |= std::fseek(f, (long)(head + c0 * rowb), SEEK_SET) != 0 ||
                std::fwrite(ctx->pidx[s].p, 1, n * rowb, f) != n * rowb;
      io_err |= std::fseek(f, (long)(head + nq * rowb + c0 * rowb), SEEK_SET) != 0 ||
                std::fwrite(ctx->pd2[s].p, 1, n * rowb, f) != n * rowb;
    } else {
      std::memcpy(out_idx + c0 * k, ctx->pidx[s].p, n * rowb);
      if (out_d2) std::memcpy(out_d2 + c0 * k, ctx->pd2[s].p, n * rowb);
    }
    return hipSuccess;
  };
  for (uint64_t c = 0; c < nchunks && rc == RTNN_OK; ++c) {
    const int s = (int)(c & 1);
    const uint64_t c0 = c * chunk, n = std::min(chunk, nq - c0);
    if (c >= 2) CK(flush(c - 2));  // frees slot s's pinned rows
    // queries: caller -> pinned (once its last upload is done) -> device
    CK(hipEventSynchronize(ctx->ev_up[s]));
    std::memcpy(ctx->pq[s].p, qxyz + 3 * c0, 12 * n);
    CK(hipStreamWaitEvent(ctx->st, ctx->ev_down[s], 0));  // slot s's device rows are free
    CK(hipMemcpyAsync(ctx->sq[s].p, ctx->pq[s].p, 12 * n, hipMemcpyHostToDevice, ctx->st));
    CK(hipEventRecord(ctx->ev_up[s], ctx->st));
    rtnn_stats st{};
    int32_t* di = ctx->sidx[s].as<int32_t>();
    float* dd = want_d2 ? ctx->sd2[s].as<float>() : nullptr;
    if (method == RTNN_METHOD_PLAIN)
      rc = search_impl(ctx, ctx->sq[s].as<float>(), n, 1, mode, r, k, di, dd, 1,
                       stats ? &st : nullptr);
    else
      rc = proposal_search(ctx, method == RTNN_METHOD_LEARNED ? Proposal::kLearned
                                                              : Proposal::kPartition,
                           ctx->sq[s].as<float>(), n, 1, r, k, di, dd, 1,
                           stats ? &st : nullptr);
    if (rc) break;
    CK(hipEventRecord(ctx->ev_done[s], ctx->st));
    CK(hipStreamWaitEvent(ctx->st3, ctx->ev_done[s], 0));
    CK(hipMemcpyAsync(ctx->pidx[s].p, di, n * rowb, hipMemcpyDeviceToHost, ctx->st3));
    if (want_d2)
      CK(hipMemcpyAsync(ctx->pd2[s].p, dd, n * rowb, hipMemcpyDeviceToHost, ctx->st3));
    CK(hipEventRecord(ctx->ev_down[s], ctx->st3));
    sum.ms_total += st.ms_total;
    sum.ms_grid += st.ms_grid;
    sum.ms_qsort += st.ms_qsort;
    sum.ms_search += st.ms_search;
    sum.ms_mlp += st.ms_mlp;
    sum.ms_fallback += st.ms_fallback;
    sum.fallback_queries += st.fallback_queries;
    sum.search_launches += st.search_launches;
    sum.lanes_per_query = st.lanes_per_query;
  }
  for (uint64_t c = nchunks >= 2 ? nchunks - 2 : 0; c < nchunks && rc == RTNN_OK; ++c)
    CK(flush(c));
  CK(hipStreamSynchronize(ctx->st3));
  if (f && std::fclose(fguard.release()) != 0) io_err = true;
  if (rc) return rc;
  if (io_err) return fail(ctx, RTNN_ERR_IO, "writing %s failed", out_path);
  if (stats) {
    *stats = sum;
    fill_stats(ctx, stats);
  }
  return RTNN_OK;
}

// Host clouds in, host rows (or a result file) out, in chunks of whole
// clouds of about `chunk` points through the two slots of the streamed
// search: while chunk i is searched, chunk i-1's rows travel to pinned
// memory on the copy stream and chunk i-2's are copied from there to the
// caller's arrays or written to the file.  Each chunk is an ordinary
// batched search of its clouds (rows = its points, cloud-local indices), so
// the concatenated rows equal one batched search over all clouds.
int rtnn_search_batched_stream(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                               uint32_t nclouds, int mode, float r, uint32_t k, uint64_t chunk,
                               int32_t* out_idx, float* out_d2, const char* out_path,
                               rtnn_stats* stats) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (!out_path && !out_idx) return fail(ctx, RTNN_ERR_INVALID, "no output (out_idx or out_path)");
  int rc = check_search_args(ctx, mode, r, k, out_path ? reinterpret_cast<const void*>(out_path)
                                                      : out_idx);
  if (rc) return rc;
  rc = check_offsets(ctx, xyz, offsets, nclouds);
  if (rc) return rc;
  rc = check_batched_lanes(ctx, mode, k);
  if (rc) return rc;
  const uint64_t n = offsets[nclouds];
  if (chunk == 0) chunk = 1ull << 22;
  // Chunks of whole clouds: cloud c0 .. c1-1 of about `chunk` points (a
  // cloud larger than that is a chunk of its own).
  std::vector<uint32_t> cuts{0};
  uint64_t maxpts = 0;
  for (uint32_t c = 0; c < nclouds; ++c) {
    const uint64_t start = offsets[cuts.back()];
    if (c > cuts.back() && offsets[c + 1] - start > chunk) {
      maxpts = std::max(maxpts, offsets[c] - start);
      cuts.push_back(c);
    }
  }
  maxpts = std::max<uint64_t>({maxpts, offsets[nclouds] - offsets[cuts.back()], 1});
  cuts.push_back(nclouds);
  const uint64_t nchunks = cuts.size() - 1;
  CK(hipSetDevice(ctx->device));
  if (!ctx->st3) {
    CK(hipStreamCreateWithFlags(&ctx->st3, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
      CK(hipEventCreateWithFlags(&ctx->ev_up[i], hipEventDisableTiming));
      CK(hipEventCreateWithFlags(&ctx->ev_done[i], hipEventDisableTiming));
      CK(hipEventCreateWithFlags(&ctx->ev_down[i], hipEventDisableTiming));
    }
  }
  std::unique_ptr<FILE, int (*)(FILE*)> fguard(nullptr, &std::fclose);
  FILE* f = nullptr;
  const uint64_t head = 8 + 8 + 4 + 4;
  if (out_path) {
    f = std::fopen(out_path, "wb");
    if (!f) return fail(ctx, RTNN_ERR_IO, "cannot open %s for writing", out_path);
    fguard.reset(f);
    const uint32_t flags = 1;  // d2 present
    std::fwrite("RTNNRES1", 1, 8, f);
    std::fwrite(&n, 8, 1, f);
    std::fwrite(&k, 4, 1, f);
    std::fwrite(&flags, 4, 1, f);
  }
  const bool want_d2 = out_path || out_d2;
  const uint64_t rowb = 4ull * k;
  for (int i = 0; i < 2; ++i) {
    if (ctx->sq[i].ensure(12 * maxpts) != hipSuccess || ctx->sidx[i].ensure(rowb * maxpts) ||
        (want_d2 && ctx->sd2[i].ensure(rowb * maxpts)) || ctx->pq[i].ensure(12 * maxpts) ||
        ctx->pidx[i].ensure(rowb * maxpts) || (want_d2 && ctx->pd2[i].ensure(rowb * maxpts)))
      return fail(ctx, RTNN_ERR_OOM, "streaming buffers of %llu points",
                  (unsigned long long)maxpts);
  }
  rtnn_stats sum{};
  bool io_err = false;
  auto flush = [&](uint64_t c) -> hipError_t {
    const int s = (int)(c & 1);
    hipError_t e = hipEventSynchronize(ctx->ev_down[s]);
    if (e != hipSuccess) return e;
    const uint64_t p0 = offsets[cuts[c]], np = offsets[cuts[c + 1]] - p0;
    if (f) {
      io_err |= std::fseek(f, (long)(head + p0 * rowb), SEEK_SET) != 0 ||
                std::fwrite(ctx->pidx[s].p, 1, np * rowb, f) != np * rowb;
      io_err |= std::fseek(f, (long)(head + n * rowb + p0 * rowb), SEEK_SET) != 0 ||
                std::fwrite(ctx->pd2[s].p, 1, np * rowb, f) != np * rowb;
    } else {
      std::memcpy(out_idx + p0 * k, ctx->pidx[s].p, np * rowb);
      if (out_d2) std::memcpy(out_d2 + p0 * k, ctx->pd2[s].p, np * rowb);
    }
    return hipSuccess;
  };
  std::vector<uint64_t> sub;
  for (uint64_t c = 0; c < nchunks && rc == RTNN_OK; ++c) {
    const int s = (int)(c & 1);
    const uint32_t c0 = cuts[c], nc = cuts[c + 1] - c0;
    const uint64_t p0 = offsets[c0], np = offsets[c0 + nc] - p0;
    if (c >= 2) CK(flush(c - 2));  // frees slot s's pinned rows
    if (np == 0) {  // clouds without points: no rows
      CK(hipEventRecord(ctx->ev_down[s], ctx->st3));
      continue;
    }
    sub.assign(offsets + c0, offsets + c0 + nc + 1);
    for (uint64_t& o : sub) o -= p0;
    CK(hipEventSynchronize(ctx->ev_up[s]));
    std::memcpy(ctx->pq[s].p, xyz + 3 * p0, 12 * np);
    CK(hipStreamWaitEvent(ctx->st, ctx->ev_down[s], 0));  // slot s's device rows are free
    CK(hipMemcpyAsync(ctx->sq[s].p, ctx->pq[s].p, 12 * np, hipMemcpyHostToDevice, ctx->st));
    CK(hipEventRecord(ctx->ev_up[s], ctx->st));
    rtnn_stats st{};
    int32_t* di = ctx->sidx[s].as<int32_t>();
    float* dd = want_d2 ? ctx->sd2[s].as<float>() : nullptr;
    rc = batched_impl(ctx, ctx->sq[s].as<float>(), sub.data(), nc, mode, r, k, di, dd,
                      stats ? &st : nullptr);
    if (rc) break;
    CK(hipEventRecord(ctx->ev_done[s], ctx->st));
    CK(hipStreamWaitEvent(ctx->st3, ctx->ev_done[s], 0));
    CK(hipMemcpyAsync(ctx->pidx[s].p, di, np * rowb, hipMemcpyDeviceToHost, ctx->st3));
    if (want_d2)
      CK(hipMemcpyAsync(ctx->pd2[s].p, dd, np * rowb, hipMemcpyDeviceToHost, ctx->st3));
    CK(hipEventRecord(ctx->ev_down[s], ctx->st3));
    sum.ms_total += st.ms_total;
    sum.ms_grid += st.ms_grid;
    sum.ms_search += st.ms_search;
    sum.search_launches += st.search_launches;
    sum.lanes_per_query = st.lanes_per_query;
  }
  for (uint64_t c = nchunks >= 2 ? nchunks - 2 : 0; c < nchunks && rc == RTNN_OK; ++c)
    CK(flush(c));
  CK(hipStreamSynchronize(ctx->st3));
  if (f && std::fclose(fguard.release()) != 0) io_err = true;
  if (rc) return rc;
  if (io_err) return fail(ctx, RTNN_ERR_IO, "writing %s failed", out_path);
  if (stats) {
    *stats = sum;
    fill_stats(ctx, stats);
  }
  return RTNN_OK;
}

// ------------------------------------------------------------------ MLP --

}  // extern "C"

namespace {

// Trains on the ns training queries already in ctx->tr_q: their exact k-th
// distances (the partitioned search: exact, and far cheaper than the plain
// search when r is large), features on the edge-r count grid, Adam on the
// pinball loss, loss on the first 65536 samples.
int train_impl(rtnn_ctx ctx, uint64_t ns, float r, uint32_t k, uint32_t steps, uint32_t batch,
               float lr, float tau, uint64_t seed, float* final_loss) {
  CK(ctx->tr_idx.ensure(4 * ns * k));
  CK(ctx->tr_d2.ensure(4 * ns * k));
  rtnn_options saved = ctx->opt;
  ctx->opt.query_order = 0;
  int rc = proposal_search(ctx, Proposal::kPartition, ctx->tr_q.as<float>(), ns, 1, r, k,
                           ctx->tr_idx.as<int32_t>(), ctx->tr_d2.as<float>(), 1, nullptr);
  ctx->opt = saved;
  if (rc) return rc;
  CK(ctx->tr_t.ensure(4 * ns));
  CK(mlp_targets(ctx->tr_d2.as<float>(), k, ns, r, ctx->tr_t.as<float>(), ctx->st));
  // features on the edge-r count grid
  rc = build_count_grid(ctx, r, false);
  if (rc) return rc;
  CK(ctx->tr_f.ensure(4ull * kMlpFP * ns));
  CK(mlp_features(ctx->tr_q.as<float>(), ns, ctx->pgrid.as<GridDesc>(),
                  ctx->pcounts.as<uint32_t>(), ctx->mlp.count_scale, ctx->tr_f.as<float>(),
                  ctx->st));
  // Adam on the pinball loss
  const uint32_t nb = (batch + kGradBlock - 1) / kGradBlock;
  const uint32_t eval_n = (uint32_t)std::min<uint64_t>(ns, 1u << 16);
  const uint32_t eval_nb = (eval_n + kGradBlock - 1) / kGradBlock;
  CK(mlp_alloc(ctx->mlp, std::max(nb, eval_nb)));
  std::vector<float> init(kMlpNP);
  mlp_init_host(init.data(), seed ^ 0x5eedull);
  CK(hipMemcpyAsync(ctx->mlp.params, init.data(), 4 * kMlpNP, hipMemcpyHostToDevice, ctx->st));
  CK(fill_u32(reinterpret_cast<uint32_t*>(ctx->mlp.adam_m), kMlpNP, 0, ctx->st));
  CK(fill_u32(reinterpret_cast<uint32_t*>(ctx->mlp.adam_v), kMlpNP, 0, ctx->st));
  for (uint32_t s = 1; s <= steps; ++s) {
    uint64_t bseed = (seed * 0x9e3779b97f4a7c15ull + s) | 1ull;
    CK(mlp_grad(ctx->mlp, ctx->tr_f.as<float>(), ctx->tr_t.as<float>(), ns, batch, bseed, tau,
                1.0f / (float)batch, ctx->mlp.partial, ctx->st));
    CK(mlp_adam(ctx->mlp, nb, lr, (int)s, ctx->st));
  }
  // loss on the first eval_n samples
  CK(ctx->tr_out.ensure(4 * (kMlpNP + 8)));
  CK(mlp_grad(ctx->mlp, ctx->tr_f.as<float>(), ctx->tr_t.as<float>(), ns, eval_n, 0, tau,
              1.0f / (float)eval_n, ctx->mlp.partial, ctx->st));
  CK(mlp_reduce(ctx->mlp.partial, eval_nb, ctx->tr_out.as<float>(),
                ctx->tr_out.as<float>() + kMlpNP, ctx->st));
  float loss = 0;
  CK(hipMemcpyAsync(&loss, ctx->tr_out.as<float>() + kMlpNP, 4, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (final_loss) *final_loss = loss;
  ctx->mlp.ready = true;
  return RTNN_OK;
}

int check_train_args(rtnn_ctx ctx, float r, uint32_t k, uint64_t n_samples, uint32_t batch,
                     float lr, float tau) {
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  if (!(r > 0.0f) || k == 0 || k > 64 || n_samples == 0 || n_samples >= 0xffffffffull ||
      batch == 0 || batch > (1u << 20) || !(lr > 0.0f) || !(tau > 0.0f && tau < 1.0f))
    return fail(ctx, RTNN_ERR_INVALID, "invalid training arguments");
  return RTNN_OK;
}

}  // namespace

extern "C" {

int rtnn_mlp_train(rtnn_ctx ctx, float r, uint32_t k, uint32_t n_samples, uint32_t steps,
                   uint32_t batch, float lr, float tau, uint64_t seed, float* final_loss) {
  if (!ctx) return RTNN_ERR_INVALID;
  int rc = check_train_args(ctx, r, k, n_samples, batch, lr, tau);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device));
  const uint64_t ns = n_samples;
  // training queries drawn from the points
  CK(ctx->tr_q.ensure(12 * ns));
  CK(mlp_sample_points(ctx->pts.as<float>(), ctx->n, ns, seed, ctx->tr_q.as<float>(), ctx->st));
  return train_impl(ctx, ns, r, k, steps, batch, lr, tau, seed, final_loss);
}

int rtnn_mlp_train_queries(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, float r,
                           uint32_t k, uint32_t steps, uint32_t batch, float lr, float tau,
                           uint64_t seed, float* final_loss) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (!qxyz) return fail(ctx, RTNN_ERR_INVALID, "qxyz is NULL");
  int rc = check_train_args(ctx, r, k, nq, batch, lr, tau);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device));
  CK(ctx->tr_q.ensure(12 * nq));
  CK(hipMemcpyAsync(ctx->tr_q.p, qxyz, 12 * nq,
                    on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->st));
  return train_impl(ctx, nq, r, k, steps, batch, lr, tau, seed, final_loss);
}

int rtnn_mlp_get_weights(rtnn_ctx ctx, float* blob) {
  if (!ctx || !blob) return RTNN_ERR_INVALID;
  if (!ctx->mlp.ready) return fail(ctx, RTNN_ERR_STATE, "the radius MLP has no weights");
  CK(hipSetDevice(ctx->device));
  CK(hipMemcpyAsync(blob, ctx->mlp.params, 4 * kMlpNP, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  blob[kMlpNP] = ctx->mlp.count_scale;
  blob[kMlpNP + 1] = ctx->mlp.margin;
  return RTNN_OK;
}

int rtnn_mlp_set_weights(rtnn_ctx ctx, const float* blob) {
  if (!ctx || !blob) return RTNN_ERR_INVALID;
  for (int i = 0; i < kMlpNBlob; ++i)
    if (!std::isfinite(blob[i])) return fail(ctx, RTNN_ERR_INVALID, "non-finite weight %d", i);
  if (!(blob[kMlpNP] > 0.0f) || !(blob[kMlpNP + 1] > 0.0f))
    return fail(ctx, RTNN_ERR_INVALID, "count_scale and radius_margin must be > 0");
  CK(hipSetDevice(ctx->device));
  CK(mlp_alloc(ctx->mlp, 1));
  CK(hipMemcpyAsync(ctx->mlp.params, blob, 4 * kMlpNP, hipMemcpyHostToDevice, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  ctx->mlp.count_scale = blob[kMlpNP];
  ctx->mlp.margin = blob[kMlpNP + 1];
  ctx->mlp.ready = true;
  return RTNN_OK;
}

int rtnn_mlp_forward(rtnn_ctx ctx, const float* feats, uint64_t n, float* out) {
  if (!ctx || ((!feats || !out) && n)) return RTNN_ERR_INVALID;
  if (!ctx->mlp.ready) return fail(ctx, RTNN_ERR_STATE, "the radius MLP has no weights");
  CK(hipSetDevice(ctx->device));
  CK(mlp_forward(ctx->mlp, feats, n, out, ctx->st, ctx->num_cus));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

int rtnn_mlp_backward(rtnn_ctx ctx, const float* feats, const float* target, uint64_t n,
                      float tau, float* grads, float* loss) {
  if (!ctx || !feats || !target || !grads || n == 0 || n > (1u << 24)) return RTNN_ERR_INVALID;
  if (!ctx->mlp.ready) return fail(ctx, RTNN_ERR_STATE, "the radius MLP has no weights");
  CK(hipSetDevice(ctx->device));
  const uint32_t nb = (uint32_t)((n + kGradBlock - 1) / kGradBlock);
  CK(mlp_alloc(ctx->mlp, nb));
  CK(ctx->tr_out.ensure(4 * (kMlpNP + 8)));
  CK(mlp_grad(ctx->mlp, feats, target, n, (uint32_t)n, 0, tau, 1.0f / (float)n,
              ctx->mlp.partial, ctx->st));
  CK(mlp_reduce(ctx->mlp.partial, nb, grads, ctx->tr_out.as<float>() + kMlpNP, ctx->st));
  float l = 0;
  CK(hipMemcpyAsync(&l, ctx->tr_out.as<float>() + kMlpNP, 4, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (loss) *loss = l;
  return RTNN_OK;
}

int rtnn_mlp_features(rtnn_ctx ctx, const float* qxyz, uint64_t nq, float r, float* feats) {
  if (!ctx || ((!qxyz || !feats) && nq) || !(r > 0.0f)) return RTNN_ERR_INVALID;
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  CK(hipSetDevice(ctx->device));
  int rc = build_count_grid(ctx, r, false);
  if (rc) return rc;
  CK(mlp_features(qxyz, nq, ctx->pgrid.as<GridDesc>(), ctx->pcounts.as<uint32_t>(),
                  ctx->mlp.count_scale, feats, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

}  // extern "C"

// ------------------------------------------------------------- batched ----

namespace {

int check_offsets(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets, uint32_t nclouds) {
  if (!xyz || !offsets || nclouds == 0)
    return fail(ctx, RTNN_ERR_INVALID, "xyz/offsets must be set and nclouds >= 1");
  if (offsets[0] != 0) return fail(ctx, RTNN_ERR_INVALID, "offsets[0] must be 0");
  for (uint32_t c = 0; c < nclouds; ++c)
    if (offsets[c + 1] < offsets[c])
      return fail(ctx, RTNN_ERR_INVALID, "offsets must be non-decreasing");
  if (offsets[nclouds] >= 0xffffffffull)
    return fail(ctx, RTNN_ERR_UNSUPPORTED, "more than 2^32-2 points");
  return RTNN_OK;
}

// Builds one grid per cloud (edge `cell`, grown until the total cell count
// fits the budget) over the concatenated clouds, the sorted points and the
// query tiles (runs of <= kTileQ points of one cloud); ctx->batch_tiles is set.
// Size class of a cloud of np points (rtnn_set_size_classes).
int cloud_class(rtnn_ctx ctx, uint64_t np) {
  int j = 0;
  while (j < (int)ctx->class_upper.size() && np > ctx->class_upper[j]) ++j;
  return j;
}

// Cell edge per cloud: its class's cell factor over r with size classes set,
// else `cell` for every cloud.
std::vector<float> batched_cells(rtnn_ctx ctx, const uint64_t* offsets, uint32_t nclouds,
                                 int mode, float r, float cell) {
  std::vector<float> v(nclouds, cell);
  if (ctx->class_opt.empty()) return v;
  for (uint32_t c = 0; c < nclouds; ++c) {
    const rtnn_options& o = ctx->class_opt[cloud_class(ctx, offsets[c + 1] - offsets[c])];
    v[c] = r / (o.cell_factor > 0 ? o.cell_factor : default_cell_factor(mode));
  }
  return v;
}

int build_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets, uint32_t nclouds,
                  const std::vector<float>& cells) {
  const uint64_t n = offsets[nclouds];
  // Per-cloud bbox on the device, plan on the host.
  CK(ctx->offs.ensure(8 * (nclouds + 1)));
  CK(hipMemcpyAsync(ctx->offs.p, offsets, 8 * (nclouds + 1), hipMemcpyHostToDevice, ctx->st));
  CK(ctx->bboxes.ensure(24 * nclouds + 64 + (nclouds == 1 ? 4 * 6 * kBboxBlocks : 0)));
  uint32_t* bad = reinterpret_cast<uint32_t*>(ctx->bboxes.as<float>() + 6 * nclouds);
  CK(fill_u32(bad, 1, 0, ctx->st));
  if (nclouds == 1)  // one cloud: the multi-block reduction (partials after the box)
    CK(bbox(xyz, n, ctx->bboxes.as<float>() + 16, ctx->bboxes.as<float>(), bad, ctx->st));
  else
    CK(bbox_segmented(xyz, ctx->offs.as<uint64_t>(), nclouds, ctx->bboxes.as<float>(), bad,
                      ctx->st));
  std::vector<float> hb(6 * nclouds + 1);
  CK(hipMemcpyAsync(hb.data(), ctx->bboxes.p, 4 * (6 * nclouds + 1), hipMemcpyDeviceToHost,
                    ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  uint32_t nbad;
  std::memcpy(&nbad, &hb[6 * nclouds], 4);
  if (nbad) return fail(ctx, RTNN_ERR_INVALID, "%u non-finite point coordinates", nbad);

  std::vector<GridDesc> gd(nclouds);
  uint64_t total = 0;
  const uint64_t cap = max_cells_for(ctx, n);
  float gf = 1.0f;  // common growth of every cloud's edge
  for (int tries = 0;; ++tries) {
    total = 0;
    for (uint32_t c = 0; c < nclouds; ++c) {
      plan_grid(&hb[6 * c], cells[c] * gf, gd[c]);
      if (offsets[c + 1] == offsets[c]) gd[c].dim[0] = gd[c].dim[1] = gd[c].dim[2] = 1;
      gd[c].cell_base = (uint32_t)total;
      gd[c].point_base = (uint32_t)offsets[c];
      gd[c].npoints = (uint32_t)(offsets[c + 1] - offsets[c]);
      total += grid_cells(gd[c]);
    }
    if (total <= cap && total < 0xfffffff0ull) break;
    if (tries > 200) return fail(ctx, RTNN_ERR_UNSUPPORTED, "grid planning failed");
    gf *= 1.25f;
  }
  ctx->hgrid = gd[0];
  ctx->ncells = total;
  CK(ctx->grids.ensure(sizeof(GridDesc) * nclouds));
  CK(hipMemcpyAsync(ctx->grids.p, gd.data(), sizeof(GridDesc) * nclouds, hipMemcpyHostToDevice,
                    ctx->st));
  CK(ctx->keys.ensure(4 * n));
  CK(ctx->vals.ensure(4 * n));
  CK(ctx->cell_start.ensure(4 * (total + 1)));
  CK(ctx->spts.ensure(16 * n + kSptsPadBytes));
  CK(ctx->tmp.ensure(4 * std::max(radix_tmp_words(n), cell_bounds_tmp_words(total))));
  ctx->spts_current = false;  // spts now holds the clouds, not ctx->pts
  ctx->hlevels = 0;           // and grids/cell_start the clouds' raster grids
  CK(cell_keys(xyz, n, ctx->grids.as<GridDesc>(), ctx->offs.as<uint64_t>(), nclouds,
               ctx->keys.as<uint32_t>(), ctx->vals.as<uint32_t>(), nullptr, ctx->st));
  CK(radix_sort_pairs(ctx->keys.as<uint32_t>(), ctx->vals.as<uint32_t>(), n,
                      bits_for(total - 1), ctx->tmp.as<uint32_t>(), ctx->st,
                      &ctx->sort_passes));
  CK(cell_bounds(ctx->keys.as<uint32_t>(), n, (uint32_t)total, ctx->cell_start.as<uint32_t>(),
                 ctx->tmp.as<uint32_t>(), ctx->st));
  CK(gather_points(xyz, ctx->vals.as<uint32_t>(), n, ctx->spts.as<float4>(), ctx->st));

  // Tiles: runs of <= kTileQ sorted queries inside one cloud.  Cell keys are
  // cloud-major, so cloud c's points occupy sorted positions
  // [offsets[c], offsets[c+1]).
  // With size classes the tiles are class-major (class j's tiles are
  // class_tiles[j] .. class_tiles[j+1]), one launch per class.
  std::vector<Tile> tl;
  tl.reserve(n / kTileQ + nclouds);
  const int ncls = std::max<int>(1, (int)ctx->class_opt.size());
  ctx->class_tiles.assign(1, 0);
  for (int j = 0; j < ncls; ++j) {
    for (uint32_t c = 0; c < nclouds; ++c) {
      if (ncls > 1 && cloud_class(ctx, offsets[c + 1] - offsets[c]) != j) continue;
      for (uint64_t q0 = offsets[c]; q0 < offsets[c + 1]; q0 += kTileQ)
        tl.push_back(Tile{(uint32_t)q0,
                          (uint32_t)std::min<uint64_t>(q0 + kTileQ, offsets[c + 1]), c, 0});
    }
    ctx->class_tiles.push_back((uint32_t)tl.size());
  }
  CK(ctx->tiles.ensure(sizeof(Tile) * tl.size()));
  CK(hipMemcpyAsync(ctx->tiles.p, tl.data(), sizeof(Tile) * tl.size(), hipMemcpyHostToDevice,
                    ctx->st));
  ctx->batch_tiles = (uint32_t)tl.size();
  return RTNN_OK;
}

// Lanes per query of options o for (mode, k).
int opt_lanes(const rtnn_options& o, int mode, uint32_t k) {
  const int qpw = o.queries_per_wave;
  return qpw > 0 ? 64 / qpw : default_group_lanes(mode, k);
}

// The batched search over a's tiles: one persistent launch with options o, or
// with size classes one per class over its tile range with its options.
hipError_t launch_batched(const SearchArgs& a, int mode, uint32_t k, int num_cus,
                          const rtnn_options& o, const std::vector<uint32_t>& ctiles,
                          const std::vector<rtnn_options>& copt, hipStream_t st,
                          bool one_cloud = false) {
  const bool knn = mode == RTNN_MODE_KNN;
  if (copt.empty() && one_cloud) {
    // One cloud: the plain launch (implicit tiles over all its points, as
    // many blocks as the queries fill) instead of the persistent walk of the
    // batch's 256-query tile list, which leaves CUs idle on small batches.
    SearchArgs b = a;
    b.tiles = nullptr;
    b.ntiles = 0;
    b.persistent = 0;
    return launch_search(b, knn, opt_lanes(o, mode, k), st, num_cus, o.group_buffer != 1,
                         o.lds_tile == 1 ? 512 : o.lds_tile);
  }
  if (copt.empty())
    return launch_search(a, knn, opt_lanes(o, mode, k), st, num_cus, o.group_buffer != 1,
                         o.lds_tile == 1 ? 512 : o.lds_tile);
  for (size_t j = 0; j < copt.size(); ++j) {
    const uint32_t t0 = ctiles[j], t1 = ctiles[j + 1];
    if (t0 == t1) continue;
    SearchArgs b = a;
    b.tiles = a.tiles + t0;
    b.ntiles = t1 - t0;
    const rtnn_options& c = copt[j];
    const hipError_t e = launch_search(b, knn, opt_lanes(c, mode, k), st, num_cus,
                                       c.group_buffer != 1, c.lds_tile == 1 ? 512 : c.lds_tile);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Fails unless every configuration the batched search would launch covers k.
int check_batched_lanes(rtnn_ctx ctx, int mode, uint32_t k) {
  std::vector<rtnn_options> all = ctx->class_opt;
  if (all.empty()) all.push_back(ctx->opt);
  for (const rtnn_options& o : all) {
    const int gl = opt_lanes(o, mode, k);
    if (gl < 0 || k > max_k_for_group(gl))
      return fail(ctx, RTNN_ERR_UNSUPPORTED, "k=%u not supported by queries_per_wave=%d", k,
                  o.queries_per_wave);
  }
  return RTNN_OK;
}

// SearchArgs over the batched grid: queries = the sorted points, one tile
// list, cloud-local indices, persistent launch.
SearchArgs batched_args(rtnn_ctx ctx, uint64_t n, float r, uint32_t k, const OutBufs& o) {
  SearchArgs a = base_args(ctx, ctx->spts.as<float4>(), n, r, k, o);
  a.tiles = ctx->tiles.as<Tile>();
  a.ntiles = ctx->batch_tiles;
  a.local_index = 1;
  a.persistent = 1;
  return a;
}

// Boundary-pair report over the grid and queries described by `a` (rows =
// output rows): count pass, scan, total to the host, write pass.
int boundary_run(rtnn_ctx ctx, const SearchArgs& a, uint64_t rows, double tol,
                 uint32_t* out_pairs, uint64_t capacity, int out_on_device, uint64_t* count) {
  CK(ctx->bcnt.ensure(4 * (rows + 1)));
  CK(ctx->boff.ensure(4 * (rows + 1)));
  CK(ctx->tmp.ensure(4 * (uint64_t)scan_tmp_words(rows + 1)));
  uint32_t* cnt = ctx->bcnt.as<uint32_t>();
  uint32_t* off = ctx->boff.as<uint32_t>();
  CK(fill_u32(cnt, rows + 1, 0, ctx->st));
  CK(launch_boundary(a, tol, cnt, nullptr, 0, nullptr, ctx->st, ctx->num_cus));
  CK(exclusive_scan(cnt, off, rows + 1, ctx->tmp.as<uint32_t>(), ctx->st));
  uint32_t total = 0;
  CK(hipMemcpyAsync(&total, off + rows, 4, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  *count = total;
  const uint64_t w = std::min<uint64_t>(total, capacity);
  if (w == 0) return RTNN_OK;
  uint32_t* dst = out_pairs;
  if (!out_on_device) {
    CK(ctx->bpairs.ensure(8 * w));
    dst = ctx->bpairs.as<uint32_t>();
  }
  CK(launch_boundary(a, tol, cnt, off, w, dst, ctx->st, ctx->num_cus));
  if (!out_on_device)
    CK(hipMemcpyAsync(out_pairs, dst, 8 * w, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

int check_boundary_args(rtnn_ctx ctx, float r, double tol, const void* out_pairs,
                        uint64_t capacity, const uint64_t* count) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (!(r > 0.0f) || !std::isfinite(r)) return fail(ctx, RTNN_ERR_INVALID, "radius must be > 0");
  if (!(tol >= 0.0 && tol <= 1e-2)) return fail(ctx, RTNN_ERR_INVALID, "tol must be in [0, 1e-2]");
  if (!count) return fail(ctx, RTNN_ERR_INVALID, "count is NULL");
  if (capacity > 0 && !out_pairs) return fail(ctx, RTNN_ERR_INVALID, "out_pairs is NULL");
  return RTNN_OK;
}

}  // namespace

namespace {

// With RTNN_AUTOTUNE_LOG set in the environment the autotuners print one
// stderr line per candidate as it is timed (a sweep over a large workload
// then shows its progress).  Read at each call, so a caller may set it late.
void tune_log(const char* fmt, ...) {
  const char* v = std::getenv("RTNN_AUTOTUNE_LOG");
  if (!v || !*v || std::strcmp(v, "0") == 0) return;
  va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "[rtnn] ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, "\n");
  std::fflush(stderr);
  va_end(ap);
}

const char* method_name(int method) {
  return method == RTNN_METHOD_LEARNED ? "learned"
         : method == RTNN_METHOD_PARTITIONED ? "partitioned" : "plain";
}

// Times one autotuner candidate: a first run (sizes buffers, loads code),
// then enough runs for about 4 ms of work (1 to 8); the fastest counts, so a
// short workload's choice is not decided by one noisy sample.
template <class Run>
int time_candidate(Run run, double& best_ms) {
  rtnn_stats st{};
  int rc = run(&st);
  if (rc != RTNN_OK) return rc;
  // the fastest of about 12 ms of runs, at least 3 for candidates under
  // 50 ms (one 4 ms sample left near-equal configurations to chance), one
  // for longer ones
  const double t = std::max(st.ms_total, 1e-3);
  const int reps = t < 50.0 ? std::max(3, std::min(8, (int)(12.0 / t))) : 1;
  best_ms = 1e300;
  for (int i = 0; i < reps; ++i) {
    rc = run(&st);
    if (rc != RTNN_OK) return rc;
    best_ms = std::min(best_ms, st.ms_total);
  }
  return RTNN_OK;
}

int batched_impl(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets, uint32_t nclouds,
                 int mode, float r, uint32_t k, int32_t* out_idx, float* out_d2,
                 rtnn_stats* stats);

}  // namespace

extern "C" {

int rtnn_search_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                        uint32_t nclouds, int mode, float r, uint32_t k, int32_t* out_idx,
                        float* out_d2, rtnn_stats* stats) {
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  rc = check_offsets(ctx, xyz, offsets, nclouds);
  if (rc) return rc;
  if (offsets[nclouds] == 0) return RTNN_OK;
  rc = check_batched_lanes(ctx, mode, k);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device));
  return batched_impl(ctx, xyz, offsets, nclouds, mode, r, k, out_idx, out_d2, stats);
}

}  // extern "C"

namespace {

// The batched search of rtnn_search_batched after its argument checks
// (device points and outputs, host offsets, at least one point).
int batched_impl(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets, uint32_t nclouds,
                 int mode, float r, uint32_t k, int32_t* out_idx, float* out_d2,
                 rtnn_stats* stats) {
  const uint64_t n = offsets[nclouds];
  int rc;
  const bool timing = ctx->opt.time_stages != 0 || stats != nullptr;
  if (timing) CK(hipEventRecord(ctx->ev[0], ctx->st));
  const float f = ctx->opt.cell_factor > 0 ? ctx->opt.cell_factor : default_cell_factor(mode);
  rc = build_batched(ctx, xyz, offsets, nclouds,
                     batched_cells(ctx, offsets, nclouds, mode, r, r / f));
  if (rc) return rc;
  if (timing) CK(hipEventRecord(ctx->ev[1], ctx->st));

  OutBufs o{out_idx, out_d2};
  SearchArgs a = batched_args(ctx, n, r, k, o);
  CK(launch_batched(a, mode, k, ctx->num_cus, ctx->opt, ctx->class_tiles, ctx->class_opt,
                    ctx->st, nclouds == 1));
  if (timing) CK(hipEventRecord(ctx->ev[2], ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->ms_grid = ev_ms(ctx->ev[0], ctx->ev[1]);
    stats->ms_search = ev_ms(ctx->ev[1], ctx->ev[2]);
    stats->ms_total = ev_ms(ctx->ev[0], ctx->ev[2]);
    stats->search_launches = (uint32_t)std::max<size_t>(1, ctx->class_opt.size());
    stats->lanes_per_query = (uint32_t)opt_lanes(ctx->opt, mode, k);
    fill_stats(ctx, stats);
  }
  return RTNN_OK;
}

}  // namespace

extern "C" {

int rtnn_search_percloud(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                         uint32_t nclouds, int mode, float r, uint32_t k, int32_t* out_idx,
                         float* out_d2, rtnn_stats* stats) {
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  if (!xyz || !offsets || nclouds == 0)
    return fail(ctx, RTNN_ERR_INVALID, "xyz/offsets must be set and nclouds >= 1");
  CK(hipSetDevice(ctx->device));
  CK(hipEventRecord(ctx->ev[6], ctx->st));
  uint32_t launches = 0;
  for (uint32_t c = 0; c < nclouds; ++c) {
    uint64_t n = offsets[c + 1] - offsets[c];
    if (n == 0) continue;
    rc = rtnn_set_points(ctx, xyz + 3 * offsets[c], n, 1);
    if (rc) return rc;
    rc = search_impl(ctx, nullptr, n, 1, mode, r, k, out_idx + offsets[c] * k,
                     out_d2 ? out_d2 + offsets[c] * k : nullptr, 1, nullptr);
    if (rc) return rc;
    ++launches;
  }
  CK(hipEventRecord(ctx->ev[7], ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->ms_total = ev_ms(ctx->ev[6], ctx->ev[7]);
    stats->search_launches = launches;
  }
  return RTNN_OK;
}

// ------------------------------------------------------ boundary pairs --

int rtnn_boundary_pairs(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, float r,
                        double tol, uint32_t* out_pairs, uint64_t capacity, int out_on_device,
                        uint64_t* count) {
  int rc = check_boundary_args(ctx, r, tol, out_pairs, capacity, count);
  if (rc) return rc;
  *count = 0;
  if (ctx->n == 0) return fail(ctx, RTNN_ERR_STATE, "rtnn_set_points has not been called");
  if (!qxyz) nq = ctx->n;
  if (nq == 0) return RTNN_OK;
  if (nq >= 0xffffffffull) return fail(ctx, RTNN_ERR_UNSUPPORTED, "more than 2^32-2 queries");
  CK(hipSetDevice(ctx->device));
  rc = build_grid(ctx, r);
  if (rc) return rc;
  const float4* qproc = nullptr;
  rc = prepare_queries(ctx, qxyz, nq, on_device, &qproc);
  if (rc) return rc;
  OutBufs o{nullptr, nullptr};
  SearchArgs a = base_args(ctx, qproc, nq, r, 1, o);
  return boundary_run(ctx, a, nq, tol, out_pairs, capacity, out_on_device, count);
}

int rtnn_boundary_pairs_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                                uint32_t nclouds, float r, double tol, uint32_t* out_pairs,
                                uint64_t capacity, int out_on_device, uint64_t* count) {
  int rc = check_boundary_args(ctx, r, tol, out_pairs, capacity, count);
  if (rc) return rc;
  *count = 0;
  rc = check_offsets(ctx, xyz, offsets, nclouds);
  if (rc) return rc;
  const uint64_t n = offsets[nclouds];
  if (n == 0) return RTNN_OK;
  CK(hipSetDevice(ctx->device));
  rc = build_batched(ctx, xyz, offsets, nclouds, std::vector<float>(nclouds, r));
  if (rc) return rc;
  OutBufs o{nullptr, nullptr};
  SearchArgs a = batched_args(ctx, n, r, 1, o);
  return boundary_run(ctx, a, n, tol, out_pairs, capacity, out_on_device, count);
}

// ------------------------------------------------------------ autotune --

int rtnn_autotune(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int on_device, int method,
                  int mode, float r, uint32_t k, int32_t* out_idx, float* out_d2,
                  rtnn_options* best, double* best_ms) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (method < RTNN_METHOD_PLAIN || method > RTNN_METHOD_LEARNED)
    return fail(ctx, RTNN_ERR_INVALID, "unknown method %d", method);
  if (!out_idx) {  // scratch outputs owned by the context
    const uint64_t rows = qxyz ? nq : ctx->n;
    CK(hipSetDevice(ctx->device));
    CK(ctx->oidx.ensure(4 * rows * std::max<uint32_t>(k, 1)));
    CK(ctx->od2.ensure(4 * rows * std::max<uint32_t>(k, 1)));
    out_idx = ctx->oidx.as<int32_t>();
    out_d2 = ctx->od2.as<float>();
  }
  const rtnn_options base = ctx->opt;
  static const float kPlainCf[] = {1.0f, 2.0f, 3.0f};
  static const float kPropCf[] = {1.5f, 2.0f, 2.5f, 3.0f, 4.0f};
  static const int kQpwKnn[] = {0, 32, 16, 8, 4};
  static const int kQpwRange[] = {64, 32, 16, 8, 4, 1};
  const bool plain = method == RTNN_METHOD_PLAIN;
  const float* cfs = plain ? kPlainCf : kPropCf;
  const int ncf = plain ? 3 : 5;
  const bool range = plain && mode == RTNN_MODE_RADIUS;
  const int* qpws = range ? kQpwRange : kQpwKnn;
  const int nqpw = range ? 6 : 5;
  double best_t = 1e300;
  rtnn_options best_opt = base;
  // Candidates: raster grids (cell factor x lanes per query); for the radius
  // proposal methods also hierarchical Morton grids (levels x level alpha).
  // Group kernels on raster grids are also tried with the compact LDS buffer.
  struct Cand {
    float cf;
    int qpw, levels;
    float alpha;
    int buffer, tile, leaf;
    int tlanes = 0;
  };
  std::vector<Cand> cands;
  for (int i = 0; i < ncf; ++i)
    for (int j = 0; j < nqpw; ++j) {
      const bool group = qpws[j] > 0 && qpws[j] < 64;
      cands.push_back({cfs[i], qpws[j], plain ? 0 : 1, 0.0f, group ? 2 : 0, 0, 0});
      if (group) {
        cands.push_back({cfs[i], qpws[j], plain ? 0 : 1, 0.0f, 1, 0, 0});
        if (qpws[j] != 32)  // 2-lane groups have no tiled kernel
          for (int tp : {256, 512, 1024})
            cands.push_back({cfs[i], qpws[j], plain ? 0 : 1, 0.0f, 1, tp, 0});
      }
    }
  // The pooled kernel (k <= 64): 16, 32 or 64 queries per wave share one
  // candidate pool (raster grids).
  if (k <= 64)
    for (int i = 0; i < ncf; ++i)
      for (int qpw : {16, 32, 64}) cands.push_back({cfs[i], qpw, plain ? 0 : 1, 0.0f, 3, 0, 0});
  // Hierarchical grids: the octree walk (one lane per query) with leaves of
  // 256 or 65536 points, and 4-lane groups on the level rule.
  if (!plain)
    for (int lv : {3, 5})
      for (float al : {1.0f, 0.5f}) {
        for (int leaf : {0, 65536})
          for (int tl : {1, 2}) cands.push_back({0.0f, 0, lv, al, 0, 0, leaf, tl});
        cands.push_back({0.0f, 16, lv, al, 0, 0, 0});
      }
  for (const Cand& cand : cands) {
    {
      const int qpw = cand.qpw;
      if (cand.buffer != 3 && qpw > 0 && k > max_k_for_group(64 / qpw)) continue;
      if (qpw == 0 && k > 64) continue;
      ctx->opt = base;
      ctx->opt.cell_factor = cand.cf;
      ctx->opt.queries_per_wave = qpw;
      ctx->opt.grid_levels = cand.levels;
      ctx->opt.level_alpha = cand.alpha;
      ctx->opt.group_buffer = cand.buffer;
      ctx->opt.lds_tile = cand.tile;
      ctx->opt.tree_leaf = cand.leaf;
      ctx->opt.tree_lanes = cand.tlanes;
      double ms = 0.0;
      const int rc = time_candidate(
          [&](rtnn_stats* st) {
            if (plain)
              return search_impl(ctx, qxyz, nq, on_device, mode, r, k, out_idx, out_d2, 1, st);
            return proposal_search(ctx, method == RTNN_METHOD_LEARNED ? Proposal::kLearned
                                                                       : Proposal::kPartition,
                                   qxyz, nq, on_device, r, k, out_idx, out_d2, 1, st);
          },
          ms);
      if (rc == RTNN_ERR_UNSUPPORTED) {  // a grid or kernel this workload cannot use
        tune_log("autotune %s: candidate %zu/%zu unsupported (%s)", method_name(method),
                 (size_t)(&cand - cands.data()) + 1, cands.size(), ctx->err.c_str());
        continue;
      }
      if (rc != RTNN_OK) {
        ctx->opt = base;
        return rc;
      }
      tune_log("autotune %s: candidate %zu/%zu cell_factor %g qpw %d levels %d alpha %g "
               "buffer %d tile %d leaf %d tree lanes %d: %.4f ms", method_name(method),
               (size_t)(&cand - cands.data()) + 1, cands.size(), cand.cf, cand.qpw, cand.levels,
               cand.alpha, cand.buffer, cand.tile, cand.leaf, cand.tlanes, ms);
      if (ms < best_t) {
        best_t = ms;
        best_opt = ctx->opt;
      }
    }
  }
  if (best_t == 1e300) {
    ctx->opt = base;
    return fail(ctx, RTNN_ERR_UNSUPPORTED, "no configuration supports this workload");
  }
  ctx->opt = best_opt;
  if (method == RTNN_METHOD_LEARNED) {
    // The MLP's radius margin trades first-stage volume against fallbacks;
    // measured here with the chosen configuration (results are exact for any).
    const float m0 = ctx->mlp.margin;
    float best_m = m0;
    for (float f : {0.85f, 1.15f, 1.3f, 1.5f, 2.0f}) {
      ctx->mlp.margin = m0 * f;
      double ms = 0.0;
      const int rc = time_candidate(
          [&](rtnn_stats* st) {
            return proposal_search(ctx, Proposal::kLearned, qxyz, nq, on_device, r, k, out_idx,
                                   out_d2, 1, st);
          },
          ms);
      if (rc != RTNN_OK) {
        ctx->mlp.margin = m0;
        return rc;
      }
      tune_log("autotune learned: radius margin %g: %.4f ms", ctx->mlp.margin, ms);
      if (ms < best_t) {
        best_t = ms;
        best_m = ctx->mlp.margin;
      }
    }
    ctx->mlp.margin = best_m;
  }
  if (!plain) {
    // Density buckets: which proposed radii are close enough to r that the
    // query is searched at r at once (measured with the chosen configuration).
    const float b0 = best_opt.bucket_ratio;
    for (float b : {0.95f, 0.9f, 0.8f, 0.7f, 0.6f, 0.5f}) {
      if (b == b0) continue;
      ctx->opt.bucket_ratio = b;
      double ms = 0.0;
      const int rc = time_candidate(
          [&](rtnn_stats* st) {
            return proposal_search(ctx, method == RTNN_METHOD_LEARNED ? Proposal::kLearned
                                                                       : Proposal::kPartition,
                                   qxyz, nq, on_device, r, k, out_idx, out_d2, 1, st);
          },
          ms);
      if (rc != RTNN_OK) {
        ctx->opt = base;
        return rc;
      }
      tune_log("autotune %s: bucket ratio %g: %.4f ms", method_name(method), b, ms);
      if (ms < best_t) {
        best_t = ms;
        best_opt.bucket_ratio = b;
      }
    }
    ctx->opt = best_opt;
    if (method == RTNN_METHOD_LEARNED) {
      // The network's proposal capped by the density partition's radius.
      ctx->opt.learned_cap = 1 - best_opt.learned_cap;
      double ms = 0.0;
      const int rc = time_candidate(
          [&](rtnn_stats* st) {
            return proposal_search(ctx, Proposal::kLearned, qxyz, nq, on_device, r, k, out_idx,
                                   out_d2, 1, st);
          },
          ms);
      if (rc != RTNN_OK) {
        ctx->opt = base;
        return rc;
      }
      tune_log("autotune learned: learned_cap %d: %.4f ms", ctx->opt.learned_cap, ms);
      if (ms < best_t) {
        best_t = ms;
        best_opt.learned_cap = ctx->opt.learned_cap;
      }
      ctx->opt = best_opt;
    }
    if (best_opt.grid_levels >= 2 && best_opt.queries_per_wave == 0 && k <= 64) {
      // The octree walk's short rows: at r at once, or through the growth stages.
      ctx->opt.tree_fallback = 1 - best_opt.tree_fallback;
      double ms = 0.0;
      const int rc = time_candidate(
          [&](rtnn_stats* st) {
            return proposal_search(ctx, method == RTNN_METHOD_LEARNED ? Proposal::kLearned
                                                                       : Proposal::kPartition,
                                   qxyz, nq, on_device, r, k, out_idx, out_d2, 1, st);
          },
          ms);
      if (rc != RTNN_OK) {
        ctx->opt = base;
        return rc;
      }
      tune_log("autotune %s: tree_fallback %d: %.4f ms", method_name(method),
               ctx->opt.tree_fallback, ms);
      if (ms < best_t) {
        best_t = ms;
        best_opt.tree_fallback = ctx->opt.tree_fallback;
      }
      ctx->opt = best_opt;
    }
  }
  if (best) *best = best_opt;
  if (best_ms) *best_ms = best_t;
  return RTNN_OK;
}

int rtnn_autotune_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                          uint32_t nclouds, int mode, float r, uint32_t k, int32_t* out_idx,
                          float* out_d2, rtnn_options* best, double* best_ms) {
  if (!ctx) return RTNN_ERR_INVALID;
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  // One configuration for the whole batch: size classes are dropped.
  ctx->class_upper.clear();
  ctx->class_opt.clear();
  const rtnn_options base = ctx->opt;
  double best_t = 1e300;
  rtnn_options best_opt = base;
  // Cell edge r / {0.5, 1, 2} x queries per wave {1, 4, 8, 16, 32, 64} x the
  // group kernel's LDS use: wide buffer, compact buffer, or compact buffer +
  // a wave tile of 256, 512 or 1024 candidate points.
  for (float cf : {0.5f, 1.0f, 2.0f})
    for (int qpw : {1, 4, 8, 16, 32, 64})
      for (int lds : {0, 1, 256, 512, 1024}) {
        const int gl = 64 / qpw;
        if (k > max_k_for_group(gl) || (gl == 1 && lds > 0) || (gl == 2 && lds > 1)) continue;
        ctx->opt = base;
        ctx->opt.cell_factor = cf;
        ctx->opt.queries_per_wave = qpw;
        ctx->opt.group_buffer = gl == 1 ? 0 : (lds == 0 ? 2 : 1);
        ctx->opt.lds_tile = lds > 1 ? lds : 0;
        double ms = 0.0;
        rc = time_candidate(
            [&](rtnn_stats* st) {
              return rtnn_search_batched(ctx, xyz, offsets, nclouds, mode, r, k, out_idx, out_d2,
                                         st);
            },
            ms);
        if (rc != RTNN_OK) {
          ctx->opt = base;
          return rc;
        }
        tune_log("autotune batched: cell_factor %g qpw %d lds %d: %.4f ms", cf, qpw, lds, ms);
        if (ms < best_t) {
          best_t = ms;
          best_opt = ctx->opt;
        }
      }
  ctx->opt = best_opt;
  if (best) *best = best_opt;
  if (best_ms) *best_ms = best_t;
  return RTNN_OK;
}

int rtnn_autotune_batched_classes(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                                  uint32_t nclouds, int mode, float r, uint32_t k,
                                  int32_t* out_idx, float* out_d2, uint32_t nclasses,
                                  const uint64_t* upper, rtnn_options* best, double* best_ms) {
  if (!ctx) return RTNN_ERR_INVALID;
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  rc = check_offsets(ctx, xyz, offsets, nclouds);
  if (rc) return rc;
  if (nclasses == 0 || nclasses > 16 || (nclasses > 1 && !upper))
    return fail(ctx, RTNN_ERR_INVALID, "1..16 classes with %s", "bounds");
  for (uint32_t j = 1; j + 1 < nclasses; ++j)
    if (upper[j] <= upper[j - 1]) return fail(ctx, RTNN_ERR_INVALID, "class bounds must increase");
  CK(hipSetDevice(ctx->device));
  // Every class is tuned on the sub-batch of its clouds (gathered into one
  // buffer, in batch order) with a single configuration, then all classes
  // are installed together.
  ctx->class_upper.assign(upper, upper + (nclasses - 1));
  ctx->class_opt.clear();
  std::vector<int> cls(nclouds);
  for (uint32_t c = 0; c < nclouds; ++c) cls[c] = cloud_class(ctx, offsets[c + 1] - offsets[c]);
  const rtnn_options base = ctx->opt;
  std::vector<rtnn_options> opts(nclasses, base);
  std::vector<double> times(nclasses, 0.0);
  DevBuf sub;
  int out = RTNN_OK;
  for (uint32_t j = 0; j < nclasses && out == RTNN_OK; ++j) {
    std::vector<uint64_t> so(1, 0);
    for (uint32_t c = 0; c < nclouds; ++c)
      if (cls[c] == (int)j) so.push_back(so.back() + (offsets[c + 1] - offsets[c]));
    const uint32_t nc = (uint32_t)so.size() - 1;
    if (nc == 0 || so.back() == 0) continue;
    const hipError_t he = sub.ensure(12 * so.back());
    if (he != hipSuccess) {
      out = hip_fail(ctx, he, "allocating a class's sub-batch");
      break;
    }
    // Copy the class's clouds, merging runs of adjacent clouds.
    uint64_t dst = 0;
    for (uint32_t c = 0; c < nclouds;) {
      if (cls[c] != (int)j) {
        ++c;
        continue;
      }
      uint32_t e = c + 1;
      while (e < nclouds && cls[e] == (int)j) ++e;
      const uint64_t np = offsets[e] - offsets[c];
      const hipError_t ce = np ? hipMemcpyAsync(sub.as<float>() + 3 * dst, xyz + 3 * offsets[c],
                                                12 * np, hipMemcpyDeviceToDevice, ctx->st)
                               : hipSuccess;
      if (ce != hipSuccess) {
        out = hip_fail(ctx, ce, "gathering a class's clouds");
        break;
      }
      dst += np;
      c = e;
    }
    if (out != RTNN_OK) break;
    ctx->opt = base;
    out = rtnn_autotune_batched(ctx, sub.as<float>(), so.data(), nc, mode, r, k, out_idx, out_d2,
                                &opts[j], &times[j]);
  }
  sub.release();
  ctx->opt = base;
  if (out != RTNN_OK) return out;
  ctx->class_upper.assign(upper, upper + (nclasses - 1));
  ctx->class_opt = opts;
  if (best) std::memcpy(best, opts.data(), sizeof(rtnn_options) * nclasses);
  if (best_ms) std::memcpy(best_ms, times.data(), sizeof(double) * nclasses);
  return RTNN_OK;
}

// ------------------------------------------------------ building blocks ---

int rtnn_radix_sort_pairs(rtnn_ctx ctx, uint32_t* keys, uint32_t* vals, uint64_t n,
                          uint32_t bits) {
  if (!ctx || (!keys && n) || (!vals && n) || bits > 32) return RTNN_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(ctx->tmp.ensure(4 * radix_tmp_words(n)));
  CK(radix_sort_pairs(keys, vals, n, (int)bits, ctx->tmp.as<uint32_t>(), ctx->st, nullptr));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

int rtnn_exclusive_scan(rtnn_ctx ctx, const uint32_t* in, uint32_t* out, uint64_t n) {
  if (!ctx || !out || (!in && n)) return RTNN_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(ctx->tmp.ensure(4 * (uint64_t)scan_tmp_words(n)));
  CK(exclusive_scan(in, out, n, ctx->tmp.as<uint32_t>(), ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

int rtnn_morton3(rtnn_ctx ctx, const uint32_t* xyz, uint32_t* codes, uint64_t n) {
  if (!ctx || ((!xyz || !codes) && n)) return RTNN_ERR_INVALID;
  CK(hipSetDevice(ctx->device));
  CK(morton3_codes(xyz, n, codes, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  return RTNN_OK;
}

}  // extern "C"

// ================================================ graph-captured batches ==

struct rtnn_graph_s {
  rtnn_ctx ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Buffers owned by the graph, so calls on the context between replays
  // cannot move memory the graph's nodes point at.
  DevBuf offs, bboxes, status, grids, keys, vals, cell_start, spts, tmp, tiles, ccells, bbox_part;
  uint64_t n = 0;
  uint32_t nclouds = 0, ntiles = 0;
  uint64_t cap_cells = 0;
  // Options at capture: the whole batch's, or one per size class with the
  // class-major tile ranges class_tiles[j] .. [j+1].
  rtnn_options opt;
  std::vector<rtnn_options> class_opt;
  std::vector<uint32_t> class_tiles;
  ~rtnn_graph_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    DevBuf* all[] = {&offs, &bboxes, &status, &grids, &keys, &vals, &cell_start, &spts, &tmp,
                     &tiles, &ccells, &bbox_part};
    for (DevBuf* b : all) b->release();
  }
};

// Live graph handles of the process.  rtnn_destroy frees a context's graphs,
// so a handle may outlive its graph; every graph call checks the handle here
// first and rejects a freed one (RTNN_ERR_INVALID) instead of touching it.
namespace {
std::mutex g_graphs_mu;
std::set<rtnn_graph_s*> g_live_graphs;
bool graph_live(rtnn_graph_s* g) {
  std::lock_guard<std::mutex> lk(g_graphs_mu);
  return g_live_graphs.count(g) != 0;
}
}  // namespace

static void destroy_graphs(rtnn_ctx ctx) {
  {
    std::lock_guard<std::mutex> lk(g_graphs_mu);
    for (rtnn_graph_s* g : ctx->graphs) g_live_graphs.erase(g);
  }
  for (rtnn_graph_s* g : ctx->graphs) delete g;
  ctx->graphs.clear();
}

namespace {

// The whole batched pipeline on ctx->st with the graph's buffers: status
// clear -> per-cloud bbox -> device-side grid plan -> cell keys -> radix sort
// -> cell table -> gather -> one persistent search.  Kernels only: no host
// synchronisation, allocation, memset or host copy, so it can be captured.
hipError_t enqueue_batched(rtnn_graph g, const float* xyz, float cell, int mode, float r,
                           uint32_t k, int32_t* out_idx, float* out_d2) {
  hipStream_t st = g->ctx->st;
  uint32_t* status = g->status.as<uint32_t>();
  const uint64_t* off = g->offs.as<uint64_t>();
  uint32_t* keys = g->keys.as<uint32_t>();
  uint32_t* vals = g->vals.as<uint32_t>();
  hipError_t e;
  if ((e = fill_u32(status, 4, 0, st)) != hipSuccess) return e;
  // One cloud: the multi-block reduction (the per-cloud kernel gives each
  // cloud one block, which for a single large cloud is the slowest step).
  if (g->nclouds == 1)
    e = bbox(xyz, g->n, g->bbox_part.as<float>(), g->bboxes.as<float>(), status, st);
  else
    e = bbox_segmented(xyz, off, g->nclouds, g->bboxes.as<float>(), status, st);
  if (e != hipSuccess) return e;
  const float* cells = g->class_opt.empty() ? nullptr : g->ccells.as<float>();
  if ((e = plan_batched(g->bboxes.as<float>(), off, g->nclouds, cell, cells, g->cap_cells,
                        g->grids.as<GridDesc>(), status, st)) != hipSuccess)
    return e;
  if ((e = cell_keys(xyz, g->n, g->grids.as<GridDesc>(), off, g->nclouds, keys, vals, nullptr,
                     st)) != hipSuccess)
    return e;
  if ((e = radix_sort_pairs(keys, vals, g->n, bits_for(g->cap_cells - 1), g->tmp.as<uint32_t>(),
                            st, nullptr)) != hipSuccess)
    return e;
  if ((e = cell_bounds(keys, g->n, (uint32_t)g->cap_cells, g->cell_start.as<uint32_t>(),
                       g->tmp.as<uint32_t>(), st)) != hipSuccess)
    return e;
  if ((e = gather_points(xyz, vals, g->n, g->spts.as<float4>(), st)) != hipSuccess) return e;
  SearchArgs a;
  std::memset(&a, 0, sizeof a);
  a.spts = g->spts.as<float4>();
  a.cell_start = g->cell_start.as<uint32_t>();
  a.grids = g->grids.as<GridDesc>();
  a.queries = g->spts.as<float4>();
  a.tiles = g->tiles.as<Tile>();
  a.ntiles = g->ntiles;
  a.nq = (uint32_t)g->n;
  a.r = r;
  a.qradius_scale = 1.0f;
  a.k = k;
  a.out_idx = out_idx;
  a.out_d2 = out_d2;
  a.local_index = 1;
  a.persistent = 1;
  return launch_batched(a, mode, k, g->ctx->num_cus, g->opt, g->class_tiles, g->class_opt, st,
                        g->nclouds == 1);
}

}  // namespace

extern "C" {

int rtnn_graph_batched(rtnn_ctx ctx, const float* xyz, const uint64_t* offsets,
                       uint32_t nclouds, int mode, float r, uint32_t k, int32_t* out_idx,
                       float* out_d2, rtnn_graph* out) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (!out) return fail(ctx, RTNN_ERR_INVALID, "out is NULL");
  *out = nullptr;
  int rc = check_search_args(ctx, mode, r, k, out_idx);
  if (rc) return rc;
  rc = check_offsets(ctx, xyz, offsets, nclouds);
  if (rc) return rc;
  const uint64_t n = offsets[nclouds];
  if (n == 0) return fail(ctx, RTNN_ERR_INVALID, "the batch holds no points");
  rc = check_batched_lanes(ctx, mode, k);
  if (rc) return rc;
  CK(hipSetDevice(ctx->device));
  const float f = ctx->opt.cell_factor > 0 ? ctx->opt.cell_factor : default_cell_factor(mode);
  const float cell = r / f;
  const std::vector<float> cells = batched_cells(ctx, offsets, nclouds, mode, r, cell);
  // One host-planned build sizes the cell table and the tiles; replays plan
  // on the device within that table (growing the edges if new data needs it).
  rc = build_batched(ctx, xyz, offsets, nclouds, cells);
  if (rc) return rc;
  std::unique_ptr<rtnn_graph_s> g(new rtnn_graph_s);
  g->ctx = ctx;
  g->n = n;
  g->nclouds = nclouds;
  g->ntiles = ctx->batch_tiles;
  g->opt = ctx->opt;
  g->class_opt = ctx->class_opt;
  g->class_tiles = ctx->class_tiles;
  g->cap_cells = std::min<uint64_t>(max_cells_for(ctx, n), ctx->ncells + ctx->ncells / 4 + 4096);
  g->cap_cells = std::max<uint64_t>(g->cap_cells, ctx->ncells);
  CK(g->offs.ensure(8 * (nclouds + 1)));
  CK(g->bboxes.ensure(24 * (uint64_t)nclouds));
  CK(g->status.ensure(16));
  CK(g->grids.ensure(sizeof(GridDesc) * nclouds));
  CK(g->keys.ensure(4 * n));
  CK(g->vals.ensure(4 * n));
  CK(g->cell_start.ensure(4 * (g->cap_cells + 1)));
  CK(g->spts.ensure(16 * n + kSptsPadBytes));
  CK(g->tmp.ensure(4 * std::max(radix_tmp_words(n), cell_bounds_tmp_words(g->cap_cells))));
  CK(g->tiles.ensure(sizeof(Tile) * g->ntiles));
  CK(g->ccells.ensure(4 * (uint64_t)nclouds));
  if (nclouds == 1) CK(g->bbox_part.ensure(4 * 6 * kBboxBlocks));
  CK(hipMemcpyAsync(g->ccells.p, cells.data(), 4 * (uint64_t)nclouds, hipMemcpyHostToDevice,
                    ctx->st));
  CK(hipMemcpyAsync(g->offs.p, offsets, 8 * (nclouds + 1), hipMemcpyHostToDevice, ctx->st));
  CK(hipMemcpyAsync(g->tiles.p, ctx->tiles.p, sizeof(Tile) * g->ntiles, hipMemcpyDeviceToDevice,
                    ctx->st));
  // Eager run first: loads every kernel and checks the inputs.
  CK(enqueue_batched(g.get(), xyz, cell, mode, r, k, out_idx, out_d2));
  uint32_t hs[4];
  CK(hipMemcpyAsync(hs, g->status.p, 16, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (hs[0]) return fail(ctx, RTNN_ERR_INVALID, "%u non-finite point coordinates", hs[0]);
  CK(hipStreamBeginCapture(ctx->st, hipStreamCaptureModeThreadLocal));
  const hipError_t e = enqueue_batched(g.get(), xyz, cell, mode, r, k, out_idx, out_d2);
  hipGraph_t graph = nullptr;
  const hipError_t e2 = hipStreamEndCapture(ctx->st, &graph);
  g->graph = graph;
  if (e != hipSuccess) return hip_fail(ctx, e, "capturing the batched pipeline");
  CK(e2);
  CK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  ctx->graphs.push_back(*out);
  {
    std::lock_guard<std::mutex> lk(g_graphs_mu);
    g_live_graphs.insert(*out);
  }
  return RTNN_OK;
}

int rtnn_graph_launch(rtnn_graph g, int sync, rtnn_stats* stats) {
  if (!g || !graph_live(g)) return RTNN_ERR_INVALID;
  rtnn_ctx ctx = g->ctx;
  CK(hipSetDevice(ctx->device));
  if (stats) CK(hipEventRecord(ctx->ev[6], ctx->st));
  CK(hipGraphLaunch(g->exec, ctx->st));
  if (stats) CK(hipEventRecord(ctx->ev[7], ctx->st));
  if (!sync && !stats) return RTNN_OK;
  uint32_t hs[4];
  CK(hipMemcpyAsync(hs, g->status.p, 16, hipMemcpyDeviceToHost, ctx->st));
  CK(hipStreamSynchronize(ctx->st));
  if (stats) {
    std::memset(stats, 0, sizeof *stats);
    stats->ms_total = ev_ms(ctx->ev[6], ctx->ev[7]);
    stats->search_launches = (uint32_t)std::max<size_t>(1, g->class_opt.size());
  }
  if (hs[0])
    return fail(ctx, RTNN_ERR_INVALID, "%u non-finite point coordinates (rows undefined)", hs[0]);
  return RTNN_OK;
}

int rtnn_graph_destroy(rtnn_graph g) {
  if (!g) return RTNN_OK;
  {
    // Already freed (by an earlier destroy, or with its context): rejected.
    std::lock_guard<std::mutex> lk(g_graphs_mu);
    if (!g_live_graphs.erase(g)) return RTNN_ERR_INVALID;
  }
  rtnn_ctx ctx = g->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->st);
  auto it = std::find(ctx->graphs.begin(), ctx->graphs.end(), g);
  if (it != ctx->graphs.end()) ctx->graphs.erase(it);
  delete g;
  return RTNN_OK;
}

}  // extern "C"
