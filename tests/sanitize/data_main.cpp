// Drives the host data utilities (rtnn_amd/csrc/data.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitizers.py): the point-file readers on well-formed, blank,
// malformed, over-long, empty, missing and changed-between-calls files; the
// writer round trip; every generator at sizes 0, 1 and a few thousand.
// argv[1] is a scratch directory.  Exit status 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rtnn_data.h"

namespace {

int failures = 0;
#define CHECK(c)                                                           \
  do {                                                                     \
    if (!(c)) {                                                            \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

std::string dir;

std::string put(const char* name, const std::string& text) {
  std::string path = dir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  std::fwrite(text.data(), 1, text.size(), f);
  std::fclose(f);
  return path;
}

// Two-call read; returns the code of the call that failed (or 0).
int read_all(const std::string& path, std::vector<float>& out) {
  uint64_t n = 0;
  int rc = rtnn_read_points(path.c_str(), nullptr, &n);
  if (rc) return rc;
  out.assign(3 * n, 0.0f);
  rc = rtnn_read_points(path.c_str(), out.data(), &n);
  if (rc == 0) out.resize(3 * n);
  return rc;
}

void readers() {
  std::vector<float> v;
  CHECK(read_all(put("ok.txt", "1,2,3\n\n  4.5 , -6,7e1\r\n\t\n8,9,10"), v) == 0);
  CHECK(v.size() == 9 && v[3] == 4.5f && v[4] == -6.0f && v[5] == 70.0f && v[8] == 10.0f);
  CHECK(read_all(put("empty.txt", ""), v) == 0 && v.empty());
  CHECK(read_all(put("blank.txt", "\n\n   \n"), v) == 0 && v.empty());
  uint64_t n = 0;
  CHECK(rtnn_read_points((dir + "/missing.txt").c_str(), nullptr, &n) == -2);
  CHECK(rtnn_read_points(nullptr, nullptr, &n) == -3);
  // malformed: two numbers, a word, a trailing comma
  for (const char* bad : {"1,2,3\n4,5\n", "1,2,3\nx,y,z\n", "1,2,\n"}) {
    std::string p = put("bad.txt", bad);
    n = 0;
    CHECK(rtnn_read_points(p.c_str(), nullptr, &n) == 0);
    std::vector<float> buf(3 * n + 3);
    CHECK(rtnn_read_points(p.c_str(), buf.data(), &n) == -1);
  }
  // a line longer than the reader's buffer
  std::string longline(70000, ' ');
  longline += "1,2,3\n";
  CHECK(read_all(put("long.txt", longline + "4,5,6\n"), v) == 0);
  // the file grew between the two calls: refused, the buffer not overrun
  std::string p = put("grow.txt", "1,2,3\n4,5,6\n");
  n = 0;
  CHECK(rtnn_read_points(p.c_str(), nullptr, &n) == 0 && n == 2);
  put("grow.txt", "1,2,3\n4,5,6\n7,8,9\n");
  std::vector<float> two(6);
  CHECK(rtnn_read_points(p.c_str(), two.data(), &n) == -4);

  // point sets: 2 sets of 2 lines, a short first line padded, malformed lines
  uint32_t ns = 0;
  p = put("sets.txt", "1,2,3,4,5,6\n7,8,9,10,11,12\n");
  n = 0;
  CHECK(rtnn_read_point_sets(p.c_str(), nullptr, &n, &ns) == 0 && n == 2 && ns == 2);
  std::vector<float> sets(3 * ns * n);
  CHECK(rtnn_read_point_sets(p.c_str(), sets.data(), &n, &ns) == 0);
  CHECK(sets[0] == 1 && sets[3] == 7 && sets[6] == 4 && sets[11] == 12);
  p = put("sets4.txt", "1,2,3,4\n5,6,7,8\n");
  n = 0;
  CHECK(rtnn_read_point_sets(p.c_str(), nullptr, &n, &ns) == 0 && ns == 2);
  sets.assign(3 * ns * n, -1.0f);
  CHECK(rtnn_read_point_sets(p.c_str(), sets.data(), &n, &ns) == 0);
  CHECK(sets[6] == 4.0f && sets[7] == 0.0f && sets[8] == 0.0f && sets[9] == 8.0f);
  for (const char* bad : {"1,2,3\n1,2\n", "1,2,3\n1,2,z\n", "1,2,3 x\n"}) {
    std::string q = put("sbad.txt", bad);
    n = 0;
    CHECK(rtnn_read_point_sets(q.c_str(), nullptr, &n, &ns) == -1);
  }
  // lines widened between the two calls: refused, the buffer not overrun
  p = put("widen.txt", "1,2,3\n4,5,6\n");
  n = 0;
  CHECK(rtnn_read_point_sets(p.c_str(), nullptr, &n, &ns) == 0 && ns == 1);
  put("widen.txt", "1,2,3,4,5,6,7,8,9\n1,2,3,4,5,6,7,8,9\n");
  std::vector<float> narrow(3 * ns * n);
  CHECK(rtnn_read_point_sets(p.c_str(), narrow.data(), &n, &ns) == -4);
  // grown between the two calls
  p = put("sgrow.txt", "1,2,3\n");
  n = 0;
  CHECK(rtnn_read_point_sets(p.c_str(), nullptr, &n, &ns) == 0 && n == 1);
  put("sgrow.txt", "1,2,3\n4,5,6\n");
  std::vector<float> one(3);
  CHECK(rtnn_read_point_sets(p.c_str(), one.data(), &n, &ns) == -4);
}

void generators_and_writer() {
  for (uint64_t n : {0ull, 1ull, 3000ull}) {
    std::vector<float> a(3 * n + 3), b;
    CHECK(rtnn_gen_uniform(7, n, -1.0f, 1.0f, a.data()) == 0);
    for (uint64_t i = 0; i < 3 * n; ++i) CHECK(a[i] >= -1.0f && a[i] < 1.0f);
    CHECK(rtnn_gen_gaussian_mixture(1, 2, n, 4, 50.0f, 2.0f, 0.2f, a.data()) == 0);
    CHECK(rtnn_gen_spheres(3, 4, n, 5, 100.0f, 2.0f, 9.0f, 0.01f, a.data()) == 0);
    CHECK(rtnn_gen_levy(5, n, 1.5f, 0.01f, 1000.0f, a.data()) == 0);
    for (uint64_t i = 0; i < 3 * n; ++i) CHECK(std::isfinite(a[i]));
    a.resize(3 * n);
    std::string p = dir + "/w.txt";
    CHECK(rtnn_write_points(p.c_str(), a.data(), n) == 0);
    CHECK(read_all(p, b) == 0 && b == a);  // 9 significant digits round-trip FP32
  }
  for (uint32_t nc : {1u, 17u}) {
    std::vector<uint64_t> off(nc + 1);
    CHECK(rtnn_gen_cloud_sizes(9, nc, 10, 500, off.data()) == 0);
    CHECK(off[0] == 0);
    for (uint32_t c = 0; c < nc; ++c) CHECK(off[c + 1] - off[c] >= 10 && off[c + 1] - off[c] <= 500);
    std::vector<float> xyz(3 * off[nc]);
    CHECK(rtnn_gen_lattice_clouds(9, nc, off.data(), 0.1f, 0.02f, xyz.data()) == 0);
    for (float x : xyz) CHECK(std::isfinite(x));
  }
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  dir = argv[1];
  readers();
  generators_and_writer();
  if (failures) {
    std::fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  std::printf("data checks passed\n");
  return 0;
}
