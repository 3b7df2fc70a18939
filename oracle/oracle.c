/*
 * oracle.c -- CPU restatement of RTNN's neighbor-search semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of the
 * benchmark use it, and only as the checker.
 *
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -shared -fPIC (see
 * oracle/build.py).  -ffp-contract=off keeps every multiply and add a separately
 * rounded FP32 operation, matching the parity rule for d2.
 *
 * Functions and the reference lines they restate (paths relative to the
 * reference's src/):
 *   orc_d2              optixNSearch/geometry.cu:61-62,173-174 with
 *                       sutil/vec_math.h:433-436 (float3 operator-) and
 *                       :527-530 (dot = a.x*b.x + a.y*b.y + a.z*b.z)
 *   orc_bruteforce      optixNSearch/check.cpp:20-48 (KNN: 0 < d2 < r2, K-bounded
 *                       max-queue) and geometry.cu:61-66,72-83 (range: d2 < r2,
 *                       self included, at most `limit` results), canonicalised
 *                       to the K smallest by (d2, index) in ascending order
 *   orc_grid_search     same result as orc_bruteforce via a uniform cell grid
 *                       (the exact CPU grid used for checks at 10M+ points)
 *   orc_morton3         optixNSearch/helper_mortonCode.h:19-34 (Part1By2,
 *                       MortonCode3)
 *   orc_search_size     optixNSearch/grid.cu:33-44,84-211 (getWidthFromIter,
 *                       calcSearchSize) on a raster-ordered count grid
 *   orc_boundary_pairs  pairs whose FP64 distance lies within 1e-6*r of r
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MODE_RADIUS 0
#define ORC_MODE_KNN 1

/* d2 between query q and point p: O = q - p (ray_orig - center), then
 * dot(O, O) = (O.x*O.x + O.y*O.y) + O.z*O.z, each operation rounded to FP32. */
float orc_d2(const float* q, const float* p) {
  float dx = q[0] - p[0];
  float dy = q[1] - p[1];
  float dz = q[2] - p[2];
  float sx = dx * dx;
  float sy = dy * dy;
  float sz = dz * dz;
  float s = sx + sy;
  return s + sz;
}

/* (d2, index) packed so that unsigned comparison is the lexicographic order:
 * d2 >= 0, so its IEEE bits order like the value. */
static inline uint64_t pack_key(float d2, uint32_t idx) {
  uint32_t b;
  memcpy(&b, &d2, 4);
  return ((uint64_t)b << 32) | idx;
}
static inline float key_d2(uint64_t key) {
  uint32_t b = (uint32_t)(key >> 32);
  float f;
  memcpy(&f, &b, 4);
  return f;
}

/* Bounded max-heap of the k smallest keys (the reference's knn_queue,
 * check.cpp:9-18, with the index as the tie-break). */
typedef struct {
  uint64_t* a;
  uint32_t n, cap;
} heap_t;

static void heap_push_bounded(heap_t* h, uint64_t key) {
  if (h->n < h->cap) {
    uint32_t i = h->n++;
    h->a[i] = key;
    while (i > 0) {
      uint32_t p = (i - 1) / 2;
      if (h->a[p] >= h->a[i]) break;
      uint64_t t = h->a[p]; h->a[p] = h->a[i]; h->a[i] = t;
      i = p;
    }
  } else if (key < h->a[0]) {
    h->a[0] = key;
    uint32_t i = 0;
    for (;;) {
      uint32_t l = 2 * i + 1, r = l + 1, m = i;
      if (l < h->n && h->a[l] > h->a[m]) m = l;
      if (r < h->n && h->a[r] > h->a[m]) m = r;
      if (m == i) break;
      uint64_t t = h->a[m]; h->a[m] = h->a[i]; h->a[i] = t;
      i = m;
    }
  }
}

static int cmp_u64(const void* x, const void* y) {
  uint64_t a = *(const uint64_t*)x, b = *(const uint64_t*)y;
  return a < b ? -1 : (a > b ? 1 : 0);
}

static void heap_emit(heap_t* h, uint32_t k, int32_t* out_idx, float* out_d2) {
  qsort(h->a, h->n, sizeof(uint64_t), cmp_u64);
  for (uint32_t i = 0; i < k; ++i) {
    if (i < h->n) {
      out_idx[i] = (int32_t)(uint32_t)(h->a[i] & 0xffffffffu);
      if (out_d2) out_d2[i] = key_d2(h->a[i]);
    } else {
      out_idx[i] = -1;
      if (out_d2) out_d2[i] = INFINITY;
    }
  }
}

static inline int accept(float d2, float r2, int mode) {
  if (mode == ORC_MODE_KNN) return (d2 > 0) && (d2 < r2);
  return d2 < r2;
}

/* ------------------------------------------------------------------ */
/* brute force                                                          */

typedef struct {
  const float* pts; uint64_t n;
  const float* qs; const uint64_t* qsel; uint64_t nq;
  int mode; float r; uint32_t k;
  int32_t* out_idx; float* out_d2;
  uint64_t q0, q1;
} bf_job;

static void* bf_worker(void* arg) {
  bf_job* j = (bf_job*)arg;
  float r2 = j->r * j->r;
  heap_t h;
  h.a = (uint64_t*)malloc(sizeof(uint64_t) * (j->k ? j->k : 1));
  h.cap = j->k;
  for (uint64_t t = j->q0; t < j->q1; ++t) {
    uint64_t qi = j->qsel ? j->qsel[t] : t;
    const float* q = j->qs + 3 * qi;
    h.n = 0;
    for (uint64_t p = 0; p < j->n; ++p) {
      float d2 = orc_d2(q, j->pts + 3 * p);
      if (accept(d2, r2, j->mode)) heap_push_bounded(&h, pack_key(d2, (uint32_t)p));
    }
    heap_emit(&h, j->k, j->out_idx + t * j->k, j->out_d2 ? j->out_d2 + t * j->k : 0);
  }
  free(h.a);
  return 0;
}

static int run_jobs(void* (*fn)(void*), void* jobs, size_t job_size, int threads) {
  if (threads <= 1) {
    fn(jobs);
    return 0;
  }
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  for (int t = 0; t < threads; ++t)
    pthread_create(&th[t], 0, fn, (char*)jobs + job_size * t);
  for (int t = 0; t < threads; ++t) pthread_join(th[t], 0);
  free(th);
  return 0;
}

/* Brute-force search.  qsel (nullable) selects which queries to answer; row t
 * of the output belongs to query qsel[t] (or t). */
int orc_bruteforce(const float* pts, uint64_t n, const float* qs, const uint64_t* qsel,
                   uint64_t nq, int mode, float r, uint32_t k, int32_t* out_idx,
                   float* out_d2, int threads) {
  if (k == 0) return -1;
  if (threads < 1) threads = 1;
  bf_job* jobs = (bf_job*)calloc(threads, sizeof(bf_job));
  for (int t = 0; t < threads; ++t) {
    bf_job* j = &jobs[t];
    j->pts = pts; j->n = n; j->qs = qs; j->qsel = qsel; j->nq = nq;
    j->mode = mode; j->r = r; j->k = k; j->out_idx = out_idx; j->out_d2 = out_d2;
    j->q0 = nq * t / threads;
    j->q1 = nq * (t + 1) / threads;
  }
  run_jobs(bf_worker, jobs, sizeof(bf_job), threads);
  free(jobs);
  return 0;
}

/* ------------------------------------------------------------------ */
/* exact CPU grid                                                       */

typedef struct {
  float mn[3];
  float cell, inv;
  int64_t dim[3];
  uint64_t* start;   /* ncells + 1 */
  uint32_t* order;   /* point indices sorted by (cell, index) */
  float* spts;       /* points in that order */
} cgrid;

/* Cell of a point inside the grid's box; clamped for safety. */
static int64_t cell_coord(float x, float mn, float inv, int64_t dim) {
  float c = floorf((x - mn) * inv);
  if (!(c >= 0.0f)) return 0;
  if (c >= (float)dim) return dim - 1;
  return (int64_t)c;
}

static int cgrid_build(cgrid* g, const float* pts, uint64_t n, float cell) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      float v = pts[3 * i + a];
      if (v < mn[a]) mn[a] = v;
      if (v > mx[a]) mx[a] = v;
    }
  if (n == 0) { mn[0] = mn[1] = mn[2] = 0; mx[0] = mx[1] = mx[2] = 0; }
  g->cell = cell;
  g->inv = 1.0f / cell;
  uint64_t ncells = 1;
  for (int a = 0; a < 3; ++a) {
    g->mn[a] = mn[a];
    float span = floorf((mx[a] - mn[a]) * g->inv);
    if (!(span < 2147483648.0f)) return -1;  /* too many cells (or no finite span) */
    g->dim[a] = (int64_t)span + 1;
    ncells *= (uint64_t)g->dim[a];
    if (ncells > (1ull << 31)) return -1;
  }
  uint32_t* key = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  g->start = (uint64_t*)calloc(ncells + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i < n; ++i) {
    int64_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = cell_coord(pts[3 * i + a], g->mn[a], g->inv, g->dim[a]);
    key[i] = (uint32_t)((c[2] * g->dim[1] + c[1]) * g->dim[0] + c[0]);
    g->start[key[i] + 1]++;
  }
  for (uint64_t c = 0; c < ncells; ++c) g->start[c + 1] += g->start[c];
  uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (ncells + 1));
  memcpy(fill, g->start, sizeof(uint64_t) * (ncells + 1));
  g->order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  g->spts = (float*)malloc(sizeof(float) * 3 * (n ? n : 1));
  for (uint64_t i = 0; i < n; ++i) {  /* index order -> stable within a cell */
    uint64_t pos = fill[key[i]]++;
    g->order[pos] = (uint32_t)i;
    memcpy(g->spts + 3 * pos, pts + 3 * i, 12);
  }
  free(fill);
  free(key);
  return 0;
}

static void cgrid_free(cgrid* g) {
  free(g->start); free(g->order); free(g->spts);
}

typedef struct {
  const cgrid* g;
  const float* qs; const uint64_t* qsel;
  int mode; float r; uint32_t k;
  int32_t* out_idx; float* out_d2;
  uint64_t q0, q1;
} gs_job;

static void* gs_worker(void* arg) {
  gs_job* j = (gs_job*)arg;
  const cgrid* g = j->g;
  float r2 = j->r * j->r;
  heap_t h;
  h.a = (uint64_t*)malloc(sizeof(uint64_t) * j->k);
  h.cap = j->k;
  for (uint64_t t = j->q0; t < j->q1; ++t) {
    uint64_t qi = j->qsel ? j->qsel[t] : t;
    const float* q = j->qs + 3 * qi;
    h.n = 0;
    int64_t lo[3], hi[3];
    int empty = 0;
    for (int a = 0; a < 3; ++a) {
      /* Generous pad: a point with fl(d2) < fl(r2) is within r*(1+4u) of q on
       * every axis, and its FP32 cell coordinate is off by at most a few ulps
       * of (|p| + |mn|)/cell; both are far below this pad. */
      double pad = (double)j->r * (1.0 + 1e-4) +
                   (fabs((double)q[a]) + fabs((double)g->mn[a]) + j->r) * 1e-5;
      double l = floor(((double)q[a] - pad - g->mn[a]) * g->inv);
      double u = floor(((double)q[a] + pad - g->mn[a]) * g->inv);
      if (l < 0) l = 0;
      if (u > (double)(g->dim[a] - 1)) u = (double)(g->dim[a] - 1);
      if (!(l <= u)) { empty = 1; break; }  /* beyond the box, or NaN */
      lo[a] = (int64_t)l; hi[a] = (int64_t)u;
    }
    if (!empty) {
      for (int64_t cz = lo[2]; cz <= hi[2]; ++cz)
        for (int64_t cy = lo[1]; cy <= hi[1]; ++cy) {
          uint64_t row = (uint64_t)((cz * g->dim[1] + cy) * g->dim[0]);
          uint64_t s0 = g->start[row + lo[0]], s1 = g->start[row + hi[0] + 1];
          for (uint64_t s = s0; s < s1; ++s) {
            float d2 = orc_d2(q, g->spts + 3 * s);
            if (accept(d2, r2, j->mode)) heap_push_bounded(&h, pack_key(d2, g->order[s]));
          }
        }
    }
    heap_emit(&h, j->k, j->out_idx + t * j->k, j->out_d2 ? j->out_d2 + t * j->k : 0);
  }
  free(h.a);
  return 0;
}

/* Same result as orc_bruteforce, computed over a uniform grid of edge `cell`
 * (cell <= 0 means cell = r). */
int orc_grid_search(const float* pts, uint64_t n, const float* qs, const uint64_t* qsel,
                    uint64_t nq, int mode, float r, uint32_t k, float cell,
                    int32_t* out_idx, float* out_d2, int threads) {
  if (k == 0) return -1;
  if (threads < 1) threads = 1;
  if (!(cell > 0)) cell = r;
  cgrid g;
  for (;;) {
    int rc = cgrid_build(&g, pts, n, cell);
    if (rc == 0) break;
    cell *= 2.0f;  /* too many cells: coarsen (any edge gives the same result) */
    if (!(cell < INFINITY)) return -1;  /* non-finite points: no grid */
  }
  gs_job* jobs = (gs_job*)calloc(threads, sizeof(gs_job));
  for (int t = 0; t < threads; ++t) {
    gs_job* j = &jobs[t];
    j->g = &g; j->qs = qs; j->qsel = qsel; j->mode = mode; j->r = r; j->k = k;
    j->out_idx = out_idx; j->out_d2 = out_d2;
    j->q0 = nq * t / threads; j->q1 = nq * (t + 1) / threads;
  }
  run_jobs(gs_worker, jobs, sizeof(gs_job), threads);
  free(jobs);
  cgrid_free(&g);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Morton codes (helper_mortonCode.h:19-34)                             */

static uint32_t part1by2(uint32_t x) {
  x &= 0x000003ff;
  x = (x ^ (x << 16)) & 0xff0000ff;
  x = (x ^ (x << 8)) & 0x0300f00f;
  x = (x ^ (x << 4)) & 0x030c30c3;
  x = (x ^ (x << 2)) & 0x09249249;
  return x;
}

uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z) {
  return (part1by2(z) << 2) + (part1by2(y) << 1) + part1by2(x);
}

/* ------------------------------------------------------------------ */
/* Density partition mask (grid.cu:33-44,84-211)                        */

/* counts: raster-ordered cell counts, index (z*dy + y)*dx + x.  Returns the
 * smallest iter such that the (2*iter+1)^3 block of cells centred on (x,y,z)
 * holds >= knn+1 points, or the first iter whose width (2*iter+2)*cell
 * exceeds maxWidth (the reference's loop order: the width test comes first). */
int orc_search_size(const uint32_t* counts, int dx, int dy, int dz, int x, int y, int z,
                    float cell, float max_width, uint32_t knn) {
  int iter = 0;
  for (;;) {
    uint64_t count = 0;
    for (int cz = z - iter; cz <= z + iter; ++cz)
      for (int cy = y - iter; cy <= y + iter; ++cy)
        for (int cx = x - iter; cx <= x + iter; ++cx) {
          if (cx < 0 || cy < 0 || cz < 0 || cx >= dx || cy >= dy || cz >= dz) continue;
          count += counts[((uint64_t)cz * dy + cy) * dx + cx];
        }
    float width = (float)(iter * 2 + 2) * cell;
    if (width > max_width) return iter;
    if (count >= (uint64_t)knn + 1) return iter;
    iter++;
  }
}

/* ------------------------------------------------------------------ */
/* boundary pairs                                                       */

/* For the queries in qsel (or all), count pairs whose FP64 distance lies
 * within tol*r of r; up to max_out (query, point) pairs are written. */
uint64_t orc_boundary_pairs(const float* pts, uint64_t n, const float* qs,
                            const uint64_t* qsel, uint64_t nq, float r, double tol,
                            uint64_t* out_pairs, uint64_t max_out) {
  uint64_t cnt = 0;
  double rr = (double)r;
  for (uint64_t t = 0; t < nq; ++t) {
    uint64_t qi = qsel ? qsel[t] : t;
    const float* q = qs + 3 * qi;
    for (uint64_t p = 0; p < n; ++p) {
      double dx = (double)q[0] - pts[3 * p], dy = (double)q[1] - pts[3 * p + 1],
             dz = (double)q[2] - pts[3 * p + 2];
      double d = sqrt(dx * dx + dy * dy + dz * dz);
      if (fabs(d - rr) <= tol * rr) {
        if (cnt < max_out) { out_pairs[2 * cnt] = qi; out_pairs[2 * cnt + 1] = p; }
        cnt++;
      }
    }
  }
  return cnt;
}
