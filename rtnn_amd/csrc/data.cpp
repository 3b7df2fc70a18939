// Host data utilities: point-file reader/writer and seeded workloads.
// Reader semantics follow the reference driver's read_pc_data
// (src/optixNSearch/util.cpp:101-135).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/rtnn_data.h"

namespace {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// xoshiro256** (Blackman & Vigna), seeded through splitmix64.
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& v : s) v = splitmix64(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }         // [0, 1)
  double u01_open() { return ((double)(next() >> 11) + 0.5) * 0x1.0p-53; }  // (0, 1)
  double normal() {  // Box-Muller
    double u1 = u01_open(), u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  void unit_vec(double v[3]) {  // uniform direction (Marsaglia 1972)
    for (;;) {
      double a = 2 * u01() - 1, b = 2 * u01() - 1, s2 = a * a + b * b;
      if (s2 >= 1.0 || s2 == 0.0) continue;
      double f = 2 * std::sqrt(1 - s2);
      v[0] = a * f; v[1] = b * f; v[2] = 1 - 2 * s2;
      return;
    }
  }
};

}  // namespace

extern "C" {

int rtnn_read_points(const char* path, float* xyz, uint64_t* n) {
  if (!path || !n) return -3;
  FILE* f = std::fopen(path, "rb");
  if (!f) return -2;
  // The reference reads with getline(line, 1024): lines are counted first,
  // then each is parsed as three comma-separated doubles.
  std::vector<char> line(1 << 16);
  uint64_t count = 0;
  const uint64_t cap = *n;  // second call: the buffer holds cap points
  int rc = 0;
  while (std::fgets(line.data(), (int)line.size(), f)) {
    const char* c = line.data();
    while (*c == ' ' || *c == '\t' || *c == '\r' || *c == '\n') ++c;
    if (*c == '\0') continue;  // blank line: skipped (the reference would parse garbage)
    if (xyz) {
      if (count >= cap) { rc = -4; break; }  // the file grew since the counting call
      double v[3];
      char* p = line.data();
      int got = 0;
      for (; got < 3; ++got) {
        char* end = nullptr;
        v[got] = std::strtod(p, &end);
        if (end == p) break;
        p = end;
        if (got < 2) {
          while (*p == ' ' || *p == '\t') ++p;
          if (*p != ',') { ++got; break; }
          ++p;
        }
      }
      if (got < 3) { rc = -1; *n = count + 1; break; }
      xyz[3 * count + 0] = (float)v[0];
      xyz[3 * count + 1] = (float)v[1];
      xyz[3 * count + 2] = (float)v[2];
    }
    ++count;
  }
  std::fclose(f);
  if (rc == 0) *n = count;
  return rc;
}

int rtnn_read_point_sets(const char* path, float* xyz, uint64_t* n, uint32_t* nsets) {
  if (!path || !n || !nsets) return -3;
  const uint64_t n_in = *n;       // second call: the line count of the first
  const uint32_t sets_in = *nsets;  // and its set count (the buffer's size)
  FILE* f = std::fopen(path, "rb");
  if (!f) return -2;
  std::vector<char> line(1 << 16);
  std::vector<float> vals;
  uint64_t count = 0;
  size_t first = 0;  // numbers on the first line
  uint32_t sets = 0;
  int rc = 0;
  while (std::fgets(line.data(), (int)line.size(), f)) {
    const char* c = line.data();
    while (*c == ' ' || *c == '\t' || *c == '\r' || *c == '\n') ++c;
    if (*c == '\0') continue;  // blank line
    // comma-separated numbers, each parsed straight to float (stof)
    vals.clear();
    char* p = line.data();
    for (;;) {
      char* end = nullptr;
      errno = 0;
      const float v = std::strtof(p, &end);
      if (end == p) { rc = -1; break; }
      vals.push_back(v);
      p = end;
      while (*p == ' ' || *p == '\t') ++p;
      if (*p == ',') { ++p; continue; }
      while (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n') ++p;
      if (*p != '\0') rc = -1;
      break;
    }
    if (rc == 0 && count == 0) {
      first = vals.size();
      sets = (uint32_t)((first + 2) / 3);
    } else if (rc == 0 && vals.size() != first) {
      rc = -1;
    }
    if (rc) { *n = count + 1; break; }
    if (xyz) {
      // the file grew, or its lines widened, since the first call
      if (count >= n_in || sets > sets_in) { rc = -4; break; }
      vals.resize(3 * (size_t)sets, 0.0f);  // a short first line is padded with zeros
      for (uint32_t b = 0; b < sets; ++b)
        for (int a = 0; a < 3; ++a) xyz[(b * n_in + count) * 3 + a] = vals[3 * b + a];
    }
    ++count;
  }
  std::fclose(f);
  if (rc == 0) {
    if (xyz && (count != n_in || sets != sets_in)) return -4;  // the file changed
    *n = count;
    *nsets = sets;
  }
  return rc;
}

int rtnn_write_points(const char* path, const float* xyz, uint64_t n) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return -2;
  for (uint64_t i = 0; i < n; ++i)
    std::fprintf(f, "%.9g,%.9g,%.9g\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  std::fclose(f);
  return 0;
}

int rtnn_gen_uniform(uint64_t sample_seed, uint64_t n, float lo, float hi, float* xyz) {
  Rng g(sample_seed);
  double w = (double)hi - (double)lo;
  for (uint64_t i = 0; i < 3 * n; ++i) {
    float v = (float)(lo + w * g.u01());
    if (v >= hi) v = std::nextafter(hi, lo);
    xyz[i] = v;
  }
  return 0;
}

int rtnn_gen_gaussian_mixture(uint64_t dist_seed, uint64_t sample_seed, uint64_t n,
                              uint32_t ncomp, float box, float sigma, float bg_frac,
                              float* xyz) {
  if (ncomp == 0) return -3;
  Rng d(dist_seed);
  std::vector<double> c(3 * ncomp);
  for (auto& v : c) v = box * d.u01();
  Rng g(sample_seed);
  for (uint64_t i = 0; i < n; ++i) {
    if (g.u01() < bg_frac) {
      for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (float)(box * g.u01());
    } else {
      uint32_t k = (uint32_t)(g.next() % ncomp);
      for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (float)(c[3 * k + a] + sigma * g.normal());
    }
  }
  return 0;
}

int rtnn_gen_spheres(uint64_t dist_seed, uint64_t sample_seed, uint64_t n,
                     uint32_t nspheres, float box, float rmin, float rmax, float noise,
                     float* xyz) {
  if (nspheres == 0) return -3;
  Rng d(dist_seed);
  std::vector<double> c(3 * nspheres), rad(nspheres), cum(nspheres);
  double tot = 0;
  for (uint32_t s = 0; s < nspheres; ++s) {
    for (int a = 0; a < 3; ++a) c[3 * s + a] = box * d.u01();
    rad[s] = rmin + (rmax - rmin) * d.u01();
    tot += rad[s] * rad[s];  // samples proportional to area: uniform on all surfaces
    cum[s] = tot;
  }
  Rng g(sample_seed);
  for (uint64_t i = 0; i < n; ++i) {
    double u = g.u01() * tot;
    uint32_t lo = 0, hi = nspheres - 1;
    while (lo < hi) {
      uint32_t m = (lo + hi) / 2;
      if (cum[m] > u) hi = m; else lo = m + 1;
    }
    double v[3];
    g.unit_vec(v);
    double rr = rad[lo] + noise * g.normal();
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (float)(c[3 * lo + a] + rr * v[a]);
  }
  return 0;
}

int rtnn_gen_cloud_sizes(uint64_t seed, uint32_t nclouds, uint32_t nmin, uint32_t nmax,
                         uint64_t* offsets) {
  if (nmin == 0 || nmax < nmin) return -3;
  Rng g(seed);
  double a = std::log((double)nmin), b = std::log((double)nmax);
  offsets[0] = 0;
  for (uint32_t c = 0; c < nclouds; ++c) {
    uint64_t s = (uint64_t)std::llround(std::exp(a + (b - a) * g.u01()));
    if (s < nmin) s = nmin;
    if (s > nmax) s = nmax;
    offsets[c + 1] = offsets[c] + s;
  }
  return 0;
}

int rtnn_gen_lattice_clouds(uint64_t seed, uint32_t nclouds, const uint64_t* offsets,
                            float h, float jit, float* xyz) {
  for (uint32_t c = 0; c < nclouds; ++c) {
    Rng g(seed * 0x100000001b3ull + c);
    uint64_t n = offsets[c + 1] - offsets[c];
    uint64_t side = (uint64_t)std::ceil(std::cbrt((double)n) - 1e-9);
    if (side * side * side < n) ++side;
    float* out = xyz + 3 * offsets[c];
    for (uint64_t i = 0; i < n; ++i) {
      uint64_t ix = i % side, iy = (i / side) % side, iz = i / (side * side);
      double base[3] = {(double)ix, (double)iy, (double)iz};
      for (int a = 0; a < 3; ++a)
        out[3 * i + a] = (float)(h * (base[a] + jit * (2 * g.u01() - 1)));
    }
  }
  return 0;
}

int rtnn_gen_levy(uint64_t seed, uint64_t n, float alpha, float min_step, float box,
                  float* xyz) {
  Rng g(seed);
  double p[3] = {box * g.u01(), box * g.u01(), box * g.u01()};
  double inv_alpha = 1.0 / alpha;
  for (uint64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) xyz[3 * i + a] = (float)p[a];
    double len = min_step * std::pow(g.u01_open(), -inv_alpha);
    double v[3];
    g.unit_vec(v);
    for (int a = 0; a < 3; ++a) {
      double x = std::fmod(p[a] + len * v[a], (double)box);
      if (x < 0) x += box;
      if (x >= box) x = 0;
      p[a] = x;
    }
  }
  return 0;
}

}  // extern "C"
