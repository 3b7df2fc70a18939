/*
 * rtnn_data.h -- host-side data utilities of the neighbor-search library:
 * the point-file reader of the reference driver and seeded synthetic
 * workloads.  Plain C ABI; all buffers are host memory owned by the caller
 * unless stated otherwise.
 */
#ifndef RTNN_DATA_H
#define RTNN_DATA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Reads a point file: one point per line, "x,y,z", parsed as doubles and
 * rounded to float (util.cpp:101-135: getline into a 1024-byte buffer, then
 * sscanf("%lf,%lf,%lf")).  Two calls: with xyz == NULL it only counts lines
 * into *n; then with a buffer of 3 * *n floats (and *n as returned) it fills
 * it, returning -4 if the file now holds more points than *n.  Differences from
 * the reference, on purpose: a line with fewer than three numbers is an error
 * (-1, *n = 1-based line number) instead of leaving uninitialised values,
 * blank lines are skipped, and an unreadable file is an error (-2) instead of
 * assert(0). */
int rtnn_read_points(const char* path, float* xyz, uint64_t* n);
/* Reads a file of point sets: every line holds the same count of
 * comma-separated numbers, d (set by the first line, rounded up to a multiple
 * of 3); triple b of line i is point i of set b (util.cpp:18-99, tokenize and
 * read_pc_data(..., int* d): each number parsed with stof, decimal -> float
 * directly).  Two calls: with xyz == NULL it counts lines into *n and sets
 * into *nsets; then, with *n and *nsets as returned and xyz of 3 * nsets * n
 * floats, it fills xyz set-major ([set][line][3]).  Differences, on purpose: a line with
 * another count of numbers than the first, or a token that is not a number,
 * is an error (-1, *n = its 1-based line number) where the reference reads
 * past its vector or throws; a first line whose count is not a multiple of 3
 * is padded with zeros; blank lines are skipped; -2 unreadable file, -4 the
 * file changed between the calls. */
int rtnn_read_point_sets(const char* path, float* xyz, uint64_t* n, uint32_t* nsets);
/* Writes "x,y,z" lines with 9 significant digits (round-trips FP32). */
int rtnn_write_points(const char* path, const float* xyz, uint64_t n);

/* Seeded synthetic workloads (splitmix64-seeded xoshiro256**).  dist_seed
 * fixes the distribution (centres, radii); sample_seed the draws from it, so
 * points and queries of one workload share dist_seed. */
int rtnn_gen_uniform(uint64_t sample_seed, uint64_t n, float lo, float hi, float* xyz);
int rtnn_gen_gaussian_mixture(uint64_t dist_seed, uint64_t sample_seed, uint64_t n,
                              uint32_t ncomp, float box, float sigma, float bg_frac,
                              float* xyz);
int rtnn_gen_spheres(uint64_t dist_seed, uint64_t sample_seed, uint64_t n,
                     uint32_t nspheres, float box, float rmin, float rmax,
                     float noise, float* xyz);
/* Cloud sizes log-uniform in [nmin, nmax]; offsets has nclouds + 1 entries. */
int rtnn_gen_cloud_sizes(uint64_t seed, uint32_t nclouds, uint32_t nmin, uint32_t nmax,
                         uint64_t* offsets);
/* Each cloud: the first n sites (raster order) of a cubic lattice of spacing
 * h with a side of ceil(cbrt(n)) sites, each jittered by U(-jit*h, +jit*h). */
int rtnn_gen_lattice_clouds(uint64_t seed, uint32_t nclouds, const uint64_t* offsets,
                            float h, float jit, float* xyz);
/* Levy flight: step length min_step * U^(-1/alpha), isotropic direction,
 * periodic wrap into [0, box)^3. */
int rtnn_gen_levy(uint64_t seed, uint64_t n, float alpha, float min_step, float box,
                  float* xyz);

#ifdef __cplusplus
}
#endif

#endif /* RTNN_DATA_H */
