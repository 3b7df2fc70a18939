/*
 * Drives every function of oracle/oracle.c under AddressSanitizer and
 * UndefinedBehaviorSanitizer (built by tests/test_sanitizers.py): random and
 * degenerate inputs (no points, k = 0, NaN and far queries, one-cell and
 * fine grids, thread counts above the query count), with the cheap
 * cross-checks that make a wrong answer visible too (grid search equal to the
 * brute force).  Exit status 0 when every check holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

float orc_d2(const float* q, const float* p);
int orc_bruteforce(const float* pts, uint64_t n, const float* qs, const uint64_t* qsel,
                   uint64_t nq, int mode, float r, uint32_t k, int32_t* out_idx,
                   float* out_d2, int threads);
int orc_grid_search(const float* pts, uint64_t n, const float* qs, const uint64_t* qsel,
                    uint64_t nq, int mode, float r, uint32_t k, float cell,
                    int32_t* out_idx, float* out_d2, int threads);
uint32_t orc_morton3(uint32_t x, uint32_t y, uint32_t z);
int orc_search_size(const uint32_t* counts, int dx, int dy, int dz, int x, int y, int z,
                    float cell, float max_width, uint32_t knn);
uint64_t orc_boundary_pairs(const float* pts, uint64_t n, const float* qs,
                            const uint64_t* qsel, uint64_t nq, float r, double tol,
                            uint64_t* out_pairs, uint64_t max_out);

static int failures = 0;
#define CHECK(c)                                                 \
  do {                                                           \
    if (!(c)) {                                                  \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      failures++;                                                \
    }                                                            \
  } while (0)

static uint64_t rng = 0x2545F4914F6CDD1Dull;
static float urand(float lo, float hi) {
  rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
  return lo + (hi - lo) * (float)((rng >> 40) * 0x1.0p-24);
}

/* Brute force and grid search on the same input agree bit for bit. */
static void compare(const float* pts, uint64_t n, const float* qs, uint64_t nq,
                    const uint64_t* qsel, uint64_t nsel, int mode, float r, uint32_t k,
                    float cell, int threads) {
  uint64_t rows = qsel ? nsel : nq;
  int32_t* bi = malloc(sizeof(int32_t) * (rows * k + 1));
  int32_t* gi = malloc(sizeof(int32_t) * (rows * k + 1));
  float* bd = malloc(sizeof(float) * (rows * k + 1));
  float* gd = malloc(sizeof(float) * (rows * k + 1));
  CHECK(orc_bruteforce(pts, n, qs, qsel, rows, mode, r, k, bi, bd, threads) == 0);
  CHECK(orc_grid_search(pts, n, qs, qsel, rows, mode, r, k, cell, gi, gd, threads) == 0);
  CHECK(memcmp(bi, gi, sizeof(int32_t) * rows * k) == 0);
  CHECK(memcmp(bd, gd, sizeof(float) * rows * k) == 0);
  for (uint64_t t = 0; t < rows; ++t)
    for (uint32_t j = 0; j < k; ++j) {
      int32_t p = bi[t * k + j];
      if (p < 0) { CHECK(isinf(bd[t * k + j])); continue; }
      CHECK((uint64_t)p < n);
      uint64_t q = qsel ? qsel[t] : t;
      CHECK(orc_d2(qs + 3 * q, pts + 3 * (uint64_t)p) == bd[t * k + j]);
      if (j) CHECK(bd[t * k + j - 1] <= bd[t * k + j]);
    }
  /* out_d2 may be NULL */
  CHECK(orc_bruteforce(pts, n, qs, qsel, rows, mode, r, k, bi, NULL, threads) == 0);
  CHECK(orc_grid_search(pts, n, qs, qsel, rows, mode, r, k, cell, gi, NULL, threads) == 0);
  CHECK(memcmp(bi, gi, sizeof(int32_t) * rows * k) == 0);
  free(bi); free(gi); free(bd); free(gd);
}

int main(void) {
  enum { N = 600, NQ = 64 };
  float* pts = malloc(sizeof(float) * 3 * N);
  float* qs = malloc(sizeof(float) * 3 * NQ);
  for (int i = 0; i < N; ++i)
    for (int a = 0; a < 3; ++a) pts[3 * i + a] = urand(0.0f, 10.0f);
  memcpy(pts + 3 * 10, pts + 3 * 11, 12);  /* a duplicate point */
  for (int i = 0; i < NQ; ++i)
    for (int a = 0; a < 3; ++a) qs[3 * i + a] = urand(-1.0f, 11.0f);
  qs[0] = NAN;                                      /* NaN query: empty row */
  qs[3] = 1e30f; qs[4] = -1e30f; qs[5] = 1e30f;     /* far beyond the box */
  qs[6] = -3e9f; qs[7] = 5.0f; qs[8] = 5.0f;        /* beyond on one axis */
  memcpy(qs + 9, pts + 3 * 10, 12);                 /* on the duplicate point */

  const float cells[] = {0.0f, 0.3f, 1.0f, 100.0f};
  const uint32_t ks[] = {1, 7, 40};
  for (int mode = 0; mode < 2; ++mode)
    for (int c = 0; c < 4; ++c)
      for (int kk = 0; kk < 3; ++kk)
        compare(pts, N, qs, NQ, NULL, 0, mode, 1.5f, ks[kk], cells[c], 1 + kk * 2);
  /* the points as queries, a selection of rows, more threads than rows */
  uint64_t sel[5] = {599, 0, 10, 11, 300};
  compare(pts, N, pts, N, sel, 5, 0, 2.0f, 9, 0.5f, 16);
  compare(pts, N, pts, N, sel, 5, 1, 2.0f, 9, 0.5f, 16);
  /* no points: every slot empty */
  compare(pts, 0, qs, NQ, NULL, 0, 0, 1.0f, 4, 0.0f, 2);
  /* one point, one cell */
  compare(pts, 1, qs, NQ, NULL, 0, 1, 50.0f, 3, 0.0f, 1);
  /* a fine grid: most cells empty, a query's block spans many cells */
  compare(pts, N, qs, 8, NULL, 0, 0, 0.5f, 5, 0.05f, 1);
  /* no queries */
  compare(pts, N, qs, 0, NULL, 0, 1, 1.0f, 4, 0.0f, 4);
  /* k = 0 is refused */
  int32_t dummy_i[1];
  float dummy_d[1];
  CHECK(orc_bruteforce(pts, N, qs, NULL, 1, 0, 1.0f, 0, dummy_i, dummy_d, 1) == -1);
  CHECK(orc_grid_search(pts, N, qs, NULL, 1, 0, 1.0f, 0, 0.0f, dummy_i, dummy_d, 1) == -1);

  /* Morton codes: the top bit pattern of 10-bit coordinates, masked inputs */
  CHECK(orc_morton3(1023, 1023, 1023) == 0x3FFFFFFFu);
  CHECK(orc_morton3(1, 0, 0) == 1 && orc_morton3(0, 1, 0) == 2 && orc_morton3(0, 0, 1) == 4);
  CHECK(orc_morton3(0xFFFFFFFFu, 0, 0) == orc_morton3(1023, 0, 0));

  /* search size on a 4 x 3 x 2 count grid: inside, on the corner, outside */
  uint32_t counts[24];
  for (int i = 0; i < 24; ++i) counts[i] = (uint32_t)(i % 3);
  CHECK(orc_search_size(counts, 4, 3, 2, 0, 0, 0, 1.0f, 100.0f, 0) == 1);
  CHECK(orc_search_size(counts, 4, 3, 2, 3, 2, 1, 1.0f, 100.0f, 2) == 1);
  CHECK(orc_search_size(counts, 4, 3, 2, -1, 3, 2, 1.0f, 3.5f, 1000) == 1);
  CHECK(orc_search_size(counts, 4, 3, 2, 1, 1, 0, 1.0f, 1.0f, 5) == 0);

  /* boundary pairs: a capacity below the count writes only what fits */
  uint64_t pairs[2 * 4];
  uint64_t all = orc_boundary_pairs(pts, N, pts, NULL, N, 1.0f, 1e-2, NULL, 0);
  uint64_t again = orc_boundary_pairs(pts, N, pts, NULL, N, 1.0f, 1e-2, pairs, 4);
  CHECK(all == again && all > 4);
  for (int i = 0; i < 4 && (uint64_t)i < all; ++i) CHECK(pairs[2 * i] < N && pairs[2 * i + 1] < N);
  CHECK(orc_boundary_pairs(pts, 0, qs, NULL, NQ, 1.0f, 1e-6, pairs, 4) == 0);

  free(pts);
  free(qs);
  if (failures) {
    fprintf(stderr, "%d checks failed\n", failures);
    return 1;
  }
  printf("oracle checks passed\n");
  return 0;
}
