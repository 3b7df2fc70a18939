// Test double of the search library's C ABI (include/rtnn_hip.h), for the
// sanitizer build of the rtnn CLI only: it lets the CLI's host code (flag
// parsing, file reading, the -fq report, MLP weight files, the -c check, the
// result and boundary files) run under AddressSanitizer and
// UndefinedBehaviorSanitizer in a process with no GPU.  Searches are a plain
// O(n * nq) scan with the library's result contract; nothing here is shipped
// or linked into librtnn_hip.so.  Every output buffer is written in full, so
// a buffer the CLI sized too small is reported by AddressSanitizer.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rtnn_hip.h"

struct rtnn_ctx_s {
  std::vector<float> pts;
  std::vector<float> mlp = std::vector<float>(RTNN_MLP_NBLOB, 0.0f);
  rtnn_options opt{};
};

namespace {

float d2_of(const float* q, const float* p) {
  volatile float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
  volatile float sx = dx * dx, sy = dy * dy, sz = dz * dz;
  volatile float s = sx + sy;
  return s + sz;
}

int scan(rtnn_ctx c, const float* q, uint64_t nq, int mode, float r, uint32_t k,
         int32_t* idx, float* d2, rtnn_stats* st) {
  if (!c || !idx || k == 0 || !(r > 0) || c->pts.empty()) return RTNN_ERR_INVALID;
  const float* Q = q ? q : c->pts.data();
  const uint64_t np = c->pts.size() / 3;
  const float r2 = r * r;
  std::vector<std::pair<float, int32_t>> row;
  for (uint64_t i = 0; i < nq; ++i) {
    row.clear();
    for (uint64_t p = 0; p < np; ++p) {
      float d = d2_of(&Q[3 * i], &c->pts[3 * p]);
      if (d < r2 && (mode == RTNN_MODE_RADIUS || d > 0)) row.emplace_back(d, (int32_t)p);
    }
    std::sort(row.begin(), row.end());
    for (uint32_t j = 0; j < k; ++j) {
      idx[i * k + j] = j < row.size() ? row[j].second : RTNN_EMPTY_SLOT;
      if (d2) d2[i * k + j] = j < row.size() ? row[j].first : INFINITY;
    }
  }
  if (st) *st = rtnn_stats{};
  return RTNN_OK;
}

}  // namespace

extern "C" {

int rtnn_create(int device, rtnn_ctx* out) {
  if (!out || device < 0) return RTNN_ERR_INVALID;
  *out = new rtnn_ctx_s;
  return RTNN_OK;
}
int rtnn_destroy(rtnn_ctx ctx) {
  delete ctx;
  return RTNN_OK;
}
const char* rtnn_last_error(rtnn_ctx) { return "test double"; }
const char* rtnn_create_error(void) { return "test double"; }
const char* rtnn_version(void) { return "test double of the C ABI (no GPU)"; }
int rtnn_set_options(rtnn_ctx ctx, const rtnn_options* opt) {
  if (!ctx) return RTNN_ERR_INVALID;
  ctx->opt = opt ? *opt : rtnn_options{};
  return RTNN_OK;
}
int rtnn_set_points(rtnn_ctx ctx, const float* xyz, uint64_t n, int on_device) {
  if (!ctx || !xyz || on_device) return RTNN_ERR_INVALID;
  ctx->pts.assign(xyz, xyz + 3 * n);
  return RTNN_OK;
}
int rtnn_search(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int, int mode, float r,
                uint32_t k, int32_t* out_idx, float* out_d2, int, rtnn_stats* stats) {
  return scan(ctx, qxyz, nq, mode, r, k, out_idx, out_d2, stats);
}
int rtnn_search_partitioned(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int, float r,
                            uint32_t k, int32_t* out_idx, float* out_d2, int,
                            rtnn_stats* stats) {
  return scan(ctx, qxyz, nq, RTNN_MODE_KNN, r, k, out_idx, out_d2, stats);
}
int rtnn_search_learned(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int, float r,
                        uint32_t k, int32_t* out_idx, float* out_d2, int, rtnn_stats* stats) {
  return scan(ctx, qxyz, nq, RTNN_MODE_KNN, r, k, out_idx, out_d2, stats);
}
int rtnn_mlp_train(rtnn_ctx ctx, float, uint32_t, uint32_t, uint32_t, uint32_t, float, float,
                   uint64_t seed, float* final_loss) {
  if (!ctx) return RTNN_ERR_INVALID;
  for (size_t i = 0; i < ctx->mlp.size(); ++i) ctx->mlp[i] = (float)((seed + i) % 7) * 0.01f;
  if (final_loss) *final_loss = 0.5f;
  return RTNN_OK;
}
int rtnn_mlp_get_weights(rtnn_ctx ctx, float* blob) {
  if (!ctx || !blob) return RTNN_ERR_INVALID;
  std::memcpy(blob, ctx->mlp.data(), sizeof(float) * RTNN_MLP_NBLOB);
  return RTNN_OK;
}
int rtnn_mlp_set_weights(rtnn_ctx ctx, const float* blob) {
  if (!ctx || !blob) return RTNN_ERR_INVALID;
  std::memcpy(ctx->mlp.data(), blob, sizeof(float) * RTNN_MLP_NBLOB);
  return RTNN_OK;
}
int rtnn_autotune(rtnn_ctx ctx, const float*, uint64_t, int, int, int, float, uint32_t,
                  int32_t*, float*, rtnn_options* best, double* best_ms) {
  if (!ctx) return RTNN_ERR_INVALID;
  if (best) *best = ctx->opt;
  if (best_ms) *best_ms = 1.0;
  return RTNN_OK;
}
int rtnn_boundary_pairs(rtnn_ctx ctx, const float* qxyz, uint64_t nq, int, float r, double tol,
                        uint32_t* out_pairs, uint64_t capacity, int, uint64_t* count) {
  if (!ctx || !count || (capacity && !out_pairs)) return RTNN_ERR_INVALID;
  const float* Q = qxyz ? qxyz : ctx->pts.data();
  const uint64_t np = ctx->pts.size() / 3;
  uint64_t c = 0;
  for (uint64_t i = 0; i < nq; ++i)
    for (uint64_t p = 0; p < np; ++p) {
      double dx = (double)Q[3 * i] - ctx->pts[3 * p];
      double dy = (double)Q[3 * i + 1] - ctx->pts[3 * p + 1];
      double dz = (double)Q[3 * i + 2] - ctx->pts[3 * p + 2];
      if (std::fabs(std::sqrt(dx * dx + dy * dy + dz * dz) - r) > tol * r) continue;
      if (c < capacity) {
        out_pairs[2 * c] = (uint32_t)i;
        out_pairs[2 * c + 1] = (uint32_t)p;
      }
      ++c;
    }
  *count = c;
  return RTNN_OK;
}
int rtnn_mem_stats(uint64_t* a, uint64_t* b, uint64_t* c, uint64_t* d) {
  for (uint64_t* p : {a, b, c, d})
    if (p) *p = 0;
  return RTNN_OK;
}

}  // extern "C"
