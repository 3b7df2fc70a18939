// rtnn -- command-line driver with the optixNSearch flag surface.
//
// Flags, defaults and their meaning follow the reference driver
// (src/optixNSearch/util.cpp:137-383, state.h:70-96, main.cpp:47-161):
//   -f/--pfile, -q/--qfile (queries default to the points), -sm/--searchmode
//   radius|knn, -r/--radius (default 2), -k/--knn (default 50; ignored in KNN
//   mode, where K is fixed at build time: -DRTNN_KNN=<K>, the reference's
//   cmake -DKNN), -d/--device, -c/--check, -p/--partition, and the advanced
//   flags, which are accepted (see below).  The reference's flag bugs are not
//   reproduced: -a takes 0/1/2, and --crStep has no trailing space.
// Added: -o <file> writes the result (format below), --learned 0|1 selects the
// learned-radius KNN, --mlp-in/--mlp-out load/save its weights, --cell-factor
// and --qpw set the grid cell factor and queries per wave, --boundary <file>
// lists the pairs whose distance lies within 1e-6 r of r.  Of the reference's
// scheduling flags, -cr sets the cell factor (cell edge = r / crRatio, as
// sort.cpp:80) and -ac 1 runs the autotuner (measured instead of modelled,
// util.cpp:551-676), -fq 1 reports the queries beyond the points' box + r
// (their rows stay, empty), -ab 0 -nb n gives the partitioned KNN n radius
// classes (grid levels); the others are accepted and do not change results:
// -a 1/2 (approximate partition radii) run exact, -gmu/-egs feed the
// reference's GPU-memory model of crRatio, which the measured autotuner
// replaces, and -mc scales its meta-grid order, which full Morton order
// replaces.
//
// Output file (little endian): "RTNNRES1", u64 nq, u32 k, u32 flags (bit 0:
// d2 present), then int32 idx[nq][k] (empty slot = -1), then float d2[nq][k].
//
// All GPU work goes through the C ABI of librtnn_hip.so (include/rtnn_hip.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <queue>
#include <string>
#include <vector>

#include "rtnn_data.h"
#include "rtnn_hip.h"

#ifndef RTNN_KNN
#define RTNN_KNN 50
#endif

namespace {

struct Args {
  std::string pfile, qfile, out, mlp_in, mlp_out, mode = "radius";
  float radius = 2.0f;
  unsigned knn = 50;
  int device = 0;
  bool check = false, partition = true, learned = false;
  int approx = 2;
  float cell_factor = 0.0f;
  float cr_ratio = 0.0f;  // -cr: cell edge = r / crRatio (sort.cpp:80)
  bool autotune = false;  // -ac 1: measure the grid configurations first
  int qpw = 0;
  std::string boundary;  // --boundary: boundary-pair report file
  bool filter_queries = false;  // -fq 1: report queries beyond the points' box + r
  bool autobatch = true;  // -ab: with -ab 0, -nb n fixes the partition's radius classes
  int numbatch = -1;      // -nb: radius classes (hierarchical grid levels) of -p 1 KNN
};

struct Timer {
  std::chrono::high_resolution_clock::time_point t0 = std::chrono::high_resolution_clock::now();
  // Same line format as the reference's Timing::stopTiming (sutil/Timing.h:69).
  double stop(const char* name) {
    double ms = std::chrono::duration<double, std::milli>(
                    std::chrono::high_resolution_clock::now() - t0).count();
    std::printf("time %s: %g ms\n\n", name, ms);
    std::fflush(stdout);
    return ms;
  }
};

[[noreturn]] void usage(const char* argv0, int code) {
  std::fprintf(stderr,
               "Usage: %s [options]\n\n"
               "Basic Options:\n"
               "  --pfile      | -f   File for search points (also the queries unless -q).\n"
               "  --qfile      | -q   File for queries.\n"
               "  --searchmode | -sm  \"knn\" or \"radius\" (default radius).\n"
               "  --radius     | -r   Search radius (default 2).\n"
               "  --knn        | -k   Max K returned in radius mode (default 50); KNN mode uses "
               "the build-time K (%d).\n"
               "  --device     | -d   GPU id (default 0).\n"
               "  --check      | -c   Sanity check on the CPU (default 0).\n"
               "  --help       | -h   This message.\n\n"
               "Advanced Options (accepted for compatibility):\n"
               "  --partition | -p   Density partition for KNN (exact; default 1).\n"
               "  --approx | -a, --autobatch | -ab, --numbatch | -nb, --gassort | -s,\n"
               "  --gsrRatio | -sg, --gather | -g, --pointsort | -ps, --querysort | -qs,\n"
               "  --autocrratio | -ac, --crratio | -cr, --gpumemused | -gmu, --crStep | -crs,\n"
               "  --metacellScale | -mc, --estgassize | -egs, --interleave | -i, --msr | -m,\n"
               "  --deferFree | -df, --filterQueries | -fq\n\n"
               "Added Options:\n"
               "  -o <file>          Write results (RTNNRES1 binary).\n"
               "  --learned 0|1      Learned-radius KNN (trains the MLP unless --mlp-in).\n"
               "  --mlp-in <file>    Load MLP weights;  --mlp-out <file>  save them.\n"
               "  --cell-factor <f>  Grid cell edge = r / f;  --qpw <n>  queries per wave.\n"
               "  --boundary <file>  Write pairs with |d - r| <= 1e-6 r (query point d).\n"
               "  (-cr sets the cell factor, -ac 1 runs the autotuner first.)\n",
               argv0, RTNN_KNN);
  std::exit(code);
}

Args parse(int argc, char** argv) {
  Args a;
  auto need = [&](int& i) -> const char* {
    if (i >= argc - 1) usage(argv[0], 1);
    return argv[++i];
  };
  // Numeric values: a value that is not a number in full is a usage error
  // (the reference's stof/atoi would throw or read 0).
  auto num_f = [&](int& i) -> float {
    const char* s = need(i);
    char* end = nullptr;
    float v = std::strtof(s, &end);
    if (end == s || *end != '\0') usage(argv[0], 1);
    return v;
  };
  auto num_i = [&](int& i) -> int {
    const char* s = need(i);
    char* end = nullptr;
    long v = std::strtol(s, &end, 10);
    if (end == s || *end != '\0' || v < -2147483647L || v > 2147483647L) usage(argv[0], 1);
    return (int)v;
  };
  for (int i = 1; i < argc; ++i) {
    std::string s = argv[i];
    if (s == "--help" || s == "-h") usage(argv[0], 0);
    else if (s == "--pfile" || s == "-f") a.pfile = need(i);
    else if (s == "--qfile" || s == "-q") a.qfile = need(i);
    else if (s == "--knn" || s == "-k") {
      const int v = num_i(i);
      if (v < 1) usage(argv[0], 1);
      a.knn = (unsigned)v;
    }
    else if (s == "--searchmode" || s == "-sm") {
      a.mode = need(i);
      if (a.mode != "knn" && a.mode != "radius") usage(argv[0], 1);
    } else if (s == "--radius" || s == "-r") a.radius = num_f(i);
    else if (s == "--device" || s == "-d") a.device = num_i(i);
    else if (s == "--check" || s == "-c") a.check = num_i(i) != 0;
    else if (s == "--partition" || s == "-p") a.partition = num_i(i) != 0;
    else if (s == "--approx" || s == "-a") a.approx = num_i(i);
    else if (s == "--gassort" || s == "-s") {
      int v = num_i(i);
      if (v < 0 || v > 2) usage(argv[0], 1);
    } else if (s == "--gsrRatio" || s == "-sg") {
      if (!(num_f(i) > 0)) usage(argv[0], 1);
    } else if (s == "--estgassize" || s == "-egs") {
      if (num_f(i) < 0) usage(argv[0], 1);
    } else if (s == "--crStep" || s == "--crStep " || s == "-crs") {
      if (!(num_f(i) > 1)) {
        std::fprintf(stderr, "crStep has to be greater than 1.\n");
        usage(argv[0], 1);
      }
    } else if (s == "--filterQueries" || s == "-fq") {
      a.filter_queries = num_i(i) != 0;
    } else if (s == "--autobatch" || s == "-ab") {
      a.autobatch = num_i(i) != 0;
    } else if (s == "--numbatch" || s == "-nb") {
      a.numbatch = num_i(i);
    } else if (s == "--msr" || s == "-m" || s == "--deferFree" || s == "-df" ||
               s == "--interleave" || s == "-i" || s == "--pointsort" || s == "-ps" ||
               s == "--querysort" || s == "-qs" ||
               s == "--gpumemused" || s == "-gmu" || s == "--gather" || s == "-g" ||
               s == "--metacellScale" || s == "-mc") {
      need(i);  // scheduling knobs of the BVH pipeline: no effect on results here
    } else if (s == "--crratio" || s == "-cr") {
      a.cr_ratio = num_f(i);
      if (!(a.cr_ratio > 0)) usage(argv[0], 1);
    } else if (s == "--autocrratio" || s == "-ac") a.autotune = num_i(i) != 0;
    else if (s == "--boundary") a.boundary = need(i);
    else if (s == "-o" || s == "--out") a.out = need(i);
    else if (s == "--learned") a.learned = num_i(i) != 0;
    else if (s == "--mlp-in") a.mlp_in = need(i);
    else if (s == "--mlp-out") a.mlp_out = need(i);
    else if (s == "--cell-factor") a.cell_factor = num_f(i);
    else if (s == "--qpw") a.qpw = num_i(i);
    else {
      std::fprintf(stderr, "Unknown option '%s'\n", argv[i]);
      usage(argv[0], 1);
    }
  }
  if (a.mode == "knn") a.knn = RTNN_KNN;  // util.cpp:368-369
  if (a.pfile.empty()) usage(argv[0], 1);
  return a;
}

std::vector<float> read_file(const std::string& path) {
  uint64_t n = 0;
  int rc = rtnn_read_points(path.c_str(), nullptr, &n);
  if (rc == -2) {
    std::fprintf(stderr, "Could not read the frame data %s\n", path.c_str());
    std::exit(1);
  }
  std::vector<float> v(3 * n);
  rc = rtnn_read_points(path.c_str(), v.data(), &n);
  if (rc == -1) {
    std::fprintf(stderr, "%s: line %llu does not hold three numbers\n", path.c_str(),
                 (unsigned long long)n);
    std::exit(1);
  }
  if (rc != 0) {
    std::fprintf(stderr, "%s changed while it was read (code %d)\n", path.c_str(), rc);
    std::exit(1);
  }
  v.resize(3 * n);  // a file that shrank between the two calls
  return v;
}

float d2_of(const float* q, const float* p) {
  volatile float dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
  volatile float sx = dx * dx, sy = dy * dy, sz = dz * dz;
  volatile float s = sx + sy;
  return s + sz;
}

// Sanity check with the reference's rules (check.cpp:20-139): KNN compares 5
// sampled queries against a brute force; radius checks every returned index is
// within r and reports the average neighbour count.  This check also compares
// the KNN indices and their (d2, index) order, and radius completeness for the
// sampled queries.
bool sanity_check(const Args& a, const std::vector<float>& P, const std::vector<float>& Q,
                  const std::vector<int32_t>& idx, unsigned k) {
  const uint64_t np = P.size() / 3, nq = Q.size() / 3;
  const float r2 = a.radius * a.radius;
  const bool knn = a.mode == "knn";
  bool ok = true;
  if (!knn) {
    uint64_t total = 0;
    for (uint64_t q = 0; q < nq && ok; ++q)
      for (unsigned n = 0; n < k; ++n) {
        int32_t p = idx[q * k + n];
        if (p < 0) break;
        ++total;
        if (!(d2_of(&Q[3 * q], &P[3 * (uint64_t)p]) < r2)) {
          std::printf("Point %d is not a neighbor of query %llu.\n", p, (unsigned long long)q);
          ok = false;
          break;
        }
      }
    std::fprintf(stderr, "Avg neighbor/query: %g\n", (double)total / (double)nq);
  }
  uint64_t seed = 0x1234567ull;
  for (int t = 0; t < 5 && ok; ++t) {
    seed = seed * 6364136223846793005ull + 1442695040888963407ull;
    const uint64_t q = (seed >> 17) % nq;
    std::priority_queue<std::pair<float, uint32_t>> heap;
    for (uint64_t p = 0; p < np; ++p) {
      float d2 = d2_of(&Q[3 * q], &P[3 * p]);
      bool keep = knn ? (d2 > 0 && d2 < r2) : (d2 < r2);
      if (!keep) continue;
      std::pair<float, uint32_t> e(d2, (uint32_t)p);
      if (heap.size() < k) heap.push(e);
      else if (e < heap.top()) { heap.pop(); heap.push(e); }
    }
    std::vector<int32_t> want(k, -1);
    for (int i = (int)heap.size() - 1; i >= 0; --i) { want[i] = (int32_t)heap.top().second; heap.pop(); }
    for (unsigned n = 0; n < k; ++n)
      if (want[n] != idx[q * k + n]) {
        std::printf("Incorrect query [%llu] %f, %f, %f\n", (unsigned long long)q, Q[3 * q],
                    Q[3 * q + 1], Q[3 * q + 2]);
        ok = false;
        break;
      }
  }
  if (ok) std::fprintf(stderr, "Sanity check done.\n");
  return ok;
}

int check_rc(rtnn_ctx ctx, int rc, const char* what) {
  if (rc != RTNN_OK) {
    std::fprintf(stderr, "Caught exception: %s failed (%d): %s\n", what, rc,
                 ctx ? rtnn_last_error(ctx) : rtnn_create_error());
    std::exit(1);
  }
  return rc;
}

}  // namespace

int main(int argc, char** argv) {
  Args a = parse(argc, argv);
  std::vector<float> P = read_file(a.pfile);
  const bool samepq = a.qfile.empty() || a.qfile == a.pfile;
  std::vector<float> Q = samepq ? P : read_file(a.qfile);
  const uint64_t np = P.size() / 3, nq = Q.size() / 3;
  if (np == 0 || nq == 0) {  // util.cpp:401-404
    std::printf("empty query and/or points\n");
    return 0;
  }
  const bool knn = a.mode == "knn";
  const unsigned k = a.knn;
  std::printf("========================================\n");
  std::printf("numPoints: %llu\nnumQueries: %llu\nsearchMode: %s\nradius: %g\nK: %u\n",
              (unsigned long long)np, (unsigned long long)nq, a.mode.c_str(), a.radius, k);
  std::printf("Same P and Q? %s\nQuery partition? %s\nLearned radius? %s\n",
              samepq ? "true" : "false", a.partition ? "true" : "false",
              a.learned ? "true" : "false");
  std::printf("Device: %s\n========================================\n\n", rtnn_version());

  if (a.filter_queries && !samepq) {
    // optix.cpp:67-115 drops the queries outside the points' box grown by r
    // (they have no neighbour).  Here their rows stay, holding only empty
    // slots, so row i is still query i; the count is reported the same way.
    // Boxes as computeMinMax (sort.cpp:34-71): floor of the minimum, floor of
    // the maximum + 1, per axis.
    auto box = [](const std::vector<float>& X, float mn[3], float mx[3]) {
      for (int d = 0; d < 3; ++d) mn[d] = mx[d] = std::floor(X[d]);
      for (size_t i = 0; i < X.size(); i += 3)
        for (int d = 0; d < 3; ++d) {
          mn[d] = std::min(mn[d], std::floor(X[i + d]));
          mx[d] = std::max(mx[d], std::floor(X[i + d]));
        }
      for (int d = 0; d < 3; ++d) mx[d] += 1.0f;
    };
    float pmn[3], pmx[3], qmn[3], qmx[3];
    box(P, pmn, pmx);
    box(Q, qmn, qmx);
    bool inside = true;
    for (int d = 0; d < 3; ++d) inside = inside && qmn[d] >= pmn[d] && qmx[d] <= pmx[d];
    if (!inside) {
      uint64_t far = 0;
      for (uint64_t i = 0; i < nq; ++i) {
        bool in = true;
        for (int d = 0; d < 3; ++d)
          in = in && Q[3 * i + d] >= pmn[d] - a.radius && Q[3 * i + d] <= pmx[d] + a.radius;
        far += in ? 0 : 1;
      }
      std::printf("Filter queries: %llu (%.3f)\n", (unsigned long long)far,
                  100.0 * (double)far / (double)nq);
    }
  }
  rtnn_ctx ctx = nullptr;
  check_rc(nullptr, rtnn_create(a.device, &ctx), "rtnn_create");
  // every return from here on releases the context
  std::unique_ptr<rtnn_ctx_s, int (*)(rtnn_ctx)> ctx_guard(ctx, &rtnn_destroy);
  rtnn_options opt{};
  opt.cell_factor = a.cell_factor > 0 ? a.cell_factor : a.cr_ratio;
  opt.queries_per_wave = a.qpw;
  // -ab 0 -nb n: the reference launches the partition in n batches of radius
  // classes (sort.cpp:326-345); here one launch walks n grid levels, one per
  // radius class (2..8).
  if (!a.autobatch && a.numbatch > 0) {
    opt.grid_levels = std::min(std::max(a.numbatch, 2), 8);
    std::printf("grid levels: %d (radius classes of the partition, -nb)\n", opt.grid_levels);
  }
  check_rc(ctx, rtnn_set_options(ctx, &opt), "rtnn_set_options");
  {
    Timer t;
    check_rc(ctx, rtnn_set_points(ctx, P.data(), np, 0), "rtnn_set_points");
    t.stop("upload points and/or queries");
  }
  if (knn && a.learned) {
    if (!a.mlp_in.empty()) {
      std::vector<float> blob(RTNN_MLP_NBLOB);
      FILE* f = std::fopen(a.mlp_in.c_str(), "rb");
      const bool ok = f && std::fread(blob.data(), 4, blob.size(), f) == blob.size();
      if (f) std::fclose(f);
      if (!ok) {
        std::fprintf(stderr, "cannot read MLP weights %s\n", a.mlp_in.c_str());
        return 1;
      }
      check_rc(ctx, rtnn_mlp_set_weights(ctx, blob.data()), "rtnn_mlp_set_weights");
    } else {
      Timer t;
      float loss = 0;
      check_rc(ctx, rtnn_mlp_train(ctx, a.radius, k, 1u << 21, 2000, 4096, 1e-3f, 0.9f, 1, &loss),
               "rtnn_mlp_train");
      t.stop("train radius MLP");
      std::printf("MLP pinball loss: %g\n", loss);
    }
    if (!a.mlp_out.empty()) {
      std::vector<float> blob(RTNN_MLP_NBLOB);
      check_rc(ctx, rtnn_mlp_get_weights(ctx, blob.data()), "rtnn_mlp_get_weights");
      FILE* f = std::fopen(a.mlp_out.c_str(), "wb");
      if (!f) {
        std::fprintf(stderr, "cannot write %s\n", a.mlp_out.c_str());
        return 1;
      }
      std::fwrite(blob.data(), 4, blob.size(), f);
      std::fclose(f);
    }
  }
  const int method = knn && a.learned ? RTNN_METHOD_LEARNED
                     : knn && a.partition ? RTNN_METHOD_PARTITIONED : RTNN_METHOD_PLAIN;
  if (a.autotune) {
    Timer t;
    rtnn_options best{};
    double ms = 0;
    check_rc(ctx, rtnn_autotune(ctx, samepq ? nullptr : Q.data(), nq, 0, method,
                                knn ? RTNN_MODE_KNN : RTNN_MODE_RADIUS, a.radius, k, nullptr,
                                nullptr, &best, &ms),
             "rtnn_autotune");
    t.stop("autotune");
    std::printf("autotune: cell_factor %g, queries_per_wave %d, grid_levels %d, level_alpha %g"
                " (%g ms)\n", best.cell_factor, best.queries_per_wave, best.grid_levels,
                best.level_alpha, ms);
  }
  std::vector<int32_t> idx(nq * (uint64_t)k);
  std::vector<float> d2(a.out.empty() ? 0 : nq * (uint64_t)k);
  float* d2p = d2.empty() ? nullptr : d2.data();
  const float* qp = samepq ? nullptr : Q.data();
  rtnn_stats st{};
  {
    Timer t;
    int rc;
    if (knn && a.learned)
      rc = rtnn_search_learned(ctx, qp, nq, 0, a.radius, k, idx.data(), d2p, 0, &st);
    else if (knn && a.partition)
      rc = rtnn_search_partitioned(ctx, qp, nq, 0, a.radius, k, idx.data(), d2p, 0, &st);
    else
      rc = rtnn_search(ctx, qp, nq, 0, knn ? RTNN_MODE_KNN : RTNN_MODE_RADIUS, a.radius, k,
                       idx.data(), d2p, 0, &st);
    check_rc(ctx, rc, "search");
    t.stop("total search time");
  }
  // The reference's inner scopes (sort.cpp:671,676; search.cpp:10-57), from
  // the library's stage events: the point sort into the grid, the query
  // ordering and radius proposal (its partition), the search kernels with
  // the fallback stages, and the copy of the rows to host memory.
  std::printf("time sort points: %g ms\n\n", st.ms_grid);
  std::printf("time sort and/or partition queries: %g ms\n\n", st.ms_qsort + st.ms_mlp);
  std::printf("time search compute: %g ms\n\n", st.ms_search + st.ms_fallback);
  std::printf("time result copy D2H: %g ms\n\n", st.ms_copy);
  if (knn && (a.learned || a.partition))
    std::printf("fallback queries: %llu (%.3f%%)\n", (unsigned long long)st.fallback_queries,
                100.0 * (double)st.fallback_queries / (double)nq);
  if (!a.out.empty()) {
    FILE* f = std::fopen(a.out.c_str(), "wb");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", a.out.c_str());
      return 1;
    }
    uint64_t hq = nq;
    uint32_t hk = k, flags = 1;
    std::fwrite("RTNNRES1", 1, 8, f);
    std::fwrite(&hq, 8, 1, f);
    std::fwrite(&hk, 4, 1, f);
    std::fwrite(&flags, 4, 1, f);
    std::fwrite(idx.data(), 4, idx.size(), f);
    std::fwrite(d2.data(), 4, d2.size(), f);
    std::fclose(f);
  }
  if (!a.boundary.empty()) {
    // One "query point distance" line per pair with |d - r| <= 1e-6 r.
    uint64_t cnt = 0;
    check_rc(ctx, rtnn_boundary_pairs(ctx, samepq ? nullptr : Q.data(), nq, 0, a.radius, 1e-6,
                                      nullptr, 0, 0, &cnt), "rtnn_boundary_pairs");
    std::vector<uint32_t> pr(2 * cnt);
    if (cnt)
      check_rc(ctx, rtnn_boundary_pairs(ctx, samepq ? nullptr : Q.data(), nq, 0, a.radius, 1e-6,
                                        pr.data(), cnt, 0, &cnt), "rtnn_boundary_pairs");
    FILE* f = std::fopen(a.boundary.c_str(), "w");
    if (!f) {
      std::fprintf(stderr, "cannot write %s\n", a.boundary.c_str());
      return 1;
    }
    for (uint64_t i = 0; i < cnt; ++i) {
      const float* q = &Q[3 * (uint64_t)pr[2 * i]];
      const float* p = &P[3 * (uint64_t)pr[2 * i + 1]];
      const double dx = (double)q[0] - p[0], dy = (double)q[1] - p[1], dz = (double)q[2] - p[2];
      std::fprintf(f, "%u %u %.9g\n", pr[2 * i], pr[2 * i + 1],
                   std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    std::fclose(f);
    std::printf("boundary pairs (|d - r| <= 1e-6 r): %llu\n", (unsigned long long)cnt);
  }
  int code = 0;
  if (a.check && !sanity_check(a, P, Q, idx, k)) code = 1;
  if (std::getenv("RTNN_MEMSTAT")) {
    uint64_t dev = 0, peak = 0, pin = 0, pin_peak = 0;
    rtnn_mem_stats(&dev, &peak, &pin, &pin_peak);
    std::printf("[MEM_STATS] peak device %.3f MB, peak pinned host %.3f MB\n",
                peak / 1048576.0, pin_peak / 1048576.0);
  }
  return code;
}
