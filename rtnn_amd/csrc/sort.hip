// Device-wide primitives written for gfx950 (wave64): exclusive scan of u32
// and a stable LSD radix sort of (u32 key, u32 value) pairs.
//
// They replace the Thrust calls of the reference's grid build
// (src/optixNSearch/thrust_helper.cu:13-99: sort_by_key, exclusive_scan) that
// sort.cpp:535-612 uses for the counting sort of particles into cells.
//
// Radix sort, one 8-bit digit per pass, three kernels per pass:
//   rs_hist    : per 4096-key tile, a digit histogram (LDS atomics), stored
//                digit-major (row d = every tile's count of digit d);
//   rs_rowscan : one block per digit scans its row (the tile's start within
//                the digit) and writes the digit's total;
//   rs_scatter : per tile, the global digit starts from the totals; each
//                wave ranks a contiguous quarter of the tile by wave64 ballot
//                matching (one ballot per digit bit) against its own running
//                digit counts, a scan over the four waves orders the quarters,
//                keys are staged in LDS in digit order and written out so
//                consecutive lanes store consecutive addresses of a digit run.
// Stability (equal keys keep input order) makes the sort deterministic.
#include "common.h"
#include "internal.h"

namespace rtnn {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096

constexpr int kRsThreads = 256;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;  // 4096
constexpr int kRsWaves = kRsThreads / kWave;

// ---------------------------------------------------------------- scan ----

__device__ inline uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (kScanThreads threads).
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* sh /*waves+1*/,
                                           uint32_t& total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  constexpr int nw = kScanThreads / kWave;
  uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) sh[w] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (int i = 0; i < nw; ++i) {
      uint32_t t = sh[i];
      sh[i] = run;
      run += t;
    }
    sh[nw] = run;
  }
  __syncthreads();
  uint32_t r = sh[w] + incl - v;
  total = sh[nw];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              uint64_t n,
                                                              uint32_t* __restrict__ part) {
  __shared__ uint32_t sh[kScanThreads / kWave + 1];
  uint64_t base = (uint64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t e = base + (uint64_t)i * kScanThreads + threadIdx.x;
    if (e < n) s += in[e];
  }
  uint32_t tot;
  block_excl_scan(s, sh, tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// One block scans all partials in place (exclusive), chunk by chunk.
__global__ __launch_bounds__(kScanThreads) void k_scan_partials(uint32_t* part, uint32_t np,
                                                                uint32_t* total_out) {
  __shared__ uint32_t sh[kScanThreads / kWave + 1];
  uint32_t run = 0;
  for (uint32_t b = 0; b < np; b += kScanThreads) {
    uint32_t i = b + threadIdx.x;
    uint32_t v = i < np ? part[i] : 0;
    uint32_t tot;
    uint32_t ex = block_excl_scan(v, sh, tot);
    if (i < np) part[i] = run + ex;
    run += tot;
  }
  if (threadIdx.x == 0 && total_out) *total_out = run;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(const uint32_t* __restrict__ in,
                                                            uint64_t n,
                                                            const uint32_t* __restrict__ part,
                                                            uint32_t* __restrict__ out) {
  __shared__ uint32_t sh[kScanThreads / kWave + 1];
  // Blocked arrangement: thread t owns kScanItems consecutive elements.
  uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t e = base + i;
    v[i] = e < n ? in[e] : 0;
    s += v[i];
  }
  uint32_t tot;
  uint32_t run = block_excl_scan(s, sh, tot) + part[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    uint64_t e = base + i;
    if (e < n) out[e] = run;
    run += v[i];
  }
}

__global__ void k_scan_total(const uint32_t* total, uint32_t* dst) { *dst = *total; }

int scan_tmp_words(uint64_t n) { return (int)((n + kScanTile - 1) / kScanTile) + 1; }

// out[0..n) = exclusive scan of in; out[n] = total.  tmp: scan_tmp_words(n) u32.
hipError_t exclusive_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp,
                          hipStream_t st) {
  uint32_t nb = (uint32_t)((n + kScanTile - 1) / kScanTile);
  if (nb == 0) {
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, st, tmp, 0u, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanThreads), 0, st, in, n, tmp);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanThreads), 0, st, tmp, nb, tmp + nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(kScanThreads), 0, st, in, n, tmp, out);
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, st, tmp + nb, out + n);
  return hipGetLastError();
}

// ----------------------------------------------------------- radix sort --

// 8-bit digits: ceil(bits / 8) passes.  (9- and 10-bit digits, three passes
// for 27- or 30-bit keys, measured slower on gfx950: 0.35 vs 0.31 ms for 10M
// 27-bit keys, 7.9 vs 4.6 ms for 150M 28-bit keys; the scatter's digit runs
// get shorter and its LDS use cuts the blocks per CU.)
constexpr int kRsDigit = 8;

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* __restrict__ keys,
                                                        uint64_t n, int shift, uint32_t lowmask,
                                                        uint32_t ntiles,
                                                        uint32_t* __restrict__ hist) {
  constexpr int NB = 1 << kRsDigit;
  __shared__ uint32_t h[NB];
  for (int i = threadIdx.x; i < NB; i += kRsThreads) h[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t rk[kRsItems];
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const uint64_t e = base + (uint64_t)i * kRsThreads + threadIdx.x;
    rk[i] = e < n ? keys[e] : 0u;
  }
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const uint64_t e = base + (uint64_t)i * kRsThreads + threadIdx.x;
    const bool valid = e < n;
    const uint32_t d = valid ? ((rk[i] & lowmask) >> shift) & (NB - 1) : 0u;
    // A wave whose keys share one digit (common for nearly sorted input:
    // clouds laid out cell by cell) adds its count with one atomic instead
    // of 64 to the same LDS word.
    const uint32_t d0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)d);
    const uint64_t act = __ballot(valid);
    const uint64_t same = __ballot(valid && d == d0);
    if (same == act) {
      if (act != 0 && lane == __ffsll((unsigned long long)act) - 1)
        atomicAdd(&h[d0], (uint32_t)__popcll(act));
    } else if (valid) {
      atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < NB; i += kRsThreads)
    hist[(uint64_t)i * ntiles + blockIdx.x] = h[i];
}

// Block d: exclusive scan of row d of the digit-major tile histogram (tile t's
// count of digit d at d * ntiles + t) into hscan, and the row's total into
// totals[d].  One launch per pass instead of a device-wide scan of the table.
__global__ __launch_bounds__(kScanThreads) void k_rs_rowscan(const uint32_t* __restrict__ hraw,
                                                             uint32_t ntiles,
                                                             uint32_t* __restrict__ hscan,
                                                             uint32_t* __restrict__ totals) {
  __shared__ uint32_t sh[kScanThreads / kWave + 1];
  const uint64_t row = (uint64_t)blockIdx.x * ntiles;
  uint32_t run = 0;
  for (uint32_t b = 0; b < ntiles; b += kScanTile) {
    const uint32_t first = b + threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      v[i] = first + i < ntiles ? hraw[row + first + i] : 0u;
      s += v[i];
    }
    uint32_t tot;
    uint32_t ex = block_excl_scan(s, sh, tot) + run;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
      if (first + i < ntiles) hscan[row + first + i] = ex;
      ex += v[i];
    }
    run += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = run;
}

__device__ inline uint64_t lanemask_lt() {
  const int lane = threadIdx.x & (kWave - 1);
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n, int shift,
    uint32_t lowmask, uint32_t ntiles, const uint32_t* __restrict__ hist_raw,
    const uint32_t* __restrict__ hist_scan, const uint32_t* __restrict__ totals) {
  constexpr int DB = kRsDigit;
  constexpr int NB = 1 << DB;
  constexpr int PT = NB / kRsThreads;               // digits per thread in the digit scans
  constexpr int kWaveItems = kRsTile / kRsWaves;    // consecutive tile items of one wave
  __shared__ uint32_t s_keys[kRsTile];
  __shared__ uint32_t s_vals[kRsTile];
  __shared__ uint32_t s_lbase[NB];  // tile-local start of each digit
  __shared__ uint32_t s_gbase[NB];  // global start of this tile's digit run
  __shared__ uint32_t s_cnt[kRsWaves][NB];
  __shared__ uint32_t s_scan[kRsThreads / kWave + 1];

  const int t = threadIdx.x;
  const int w = t / kWave;
  const int lane = t & (kWave - 1);
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  const uint32_t tile_n = (uint32_t)((n - base) < (uint64_t)kRsTile ? (n - base) : kRsTile);

  {
    // Thread t owns digits t*PT .. t*PT+PT-1: the tile-local digit starts
    // scan this tile's counts; the global starts add the exclusive scan of
    // the digit totals to the tile's prefix within its digit row.
    uint32_t c[PT], tt[PT];
    uint32_t cs = 0, ts = 0;
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      const uint64_t d = (uint64_t)(t * PT + j);
      c[j] = hist_raw[d * ntiles + blockIdx.x];
      tt[j] = totals[d];
      cs += c[j];
      ts += tt[j];
    }
    uint32_t tot;
    uint32_t lex = block_excl_scan(cs, s_scan, tot);
    uint32_t gex = block_excl_scan(ts, s_scan, tot);
#pragma unroll
    for (int j = 0; j < PT; ++j) {
      const int d = t * PT + j;
      s_lbase[d] = lex;
      lex += c[j];
      s_gbase[d] = gex + hist_scan[(uint64_t)d * ntiles + blockIdx.x];
      gex += tt[j];
#pragma unroll
      for (int i = 0; i < kRsWaves; ++i) s_cnt[i][d] = 0;
    }
  }
  __syncthreads();

  // Wave w ranks the tile's items [w*kWaveItems, (w+1)*kWaveItems) on its
  // own, with running per-digit counts in its LDS row: no block barrier
  // inside the loop.  All loads are issued before the first is used.
  uint32_t rk[kRsItems], rv[kRsItems], rr[kRsItems];
#pragma unroll
  for (int it = 0; it < kRsItems; ++it) {
    const uint32_t j = (uint32_t)(w * kWaveItems + it * kWave + lane);
    rk[it] = j < tile_n ? kin[base + j] : 0u;
    rv[it] = j < tile_n ? vin[base + j] : 0u;
  }
#pragma unroll
  for (int it = 0; it < kRsItems; ++it) {
    const uint32_t j = (uint32_t)(w * kWaveItems + it * kWave + lane);
    const bool valid = j < tile_n;
    const uint32_t d = valid ? ((rk[it] & lowmask) >> shift) & (NB - 1) : 0u;
    // lanes of this wave with the same digit, from one ballot per digit bit
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < DB; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      match &= bit ? bb : ~bb;
    }
    const uint32_t below = (uint32_t)__popcll(match & lanemask_lt());
    const uint32_t cnt = (uint32_t)__popcll(match);
    // Every lane reads its digit's running count before the group's last
    // lane writes the new one (one wave: LDS operations in program order).
    const uint32_t cur = valid ? s_cnt[w][d] : 0u;
    rr[it] = cur + below;
    if (valid && below == cnt - 1) s_cnt[w][d] = cur + cnt;
  }
  __syncthreads();
  // Per digit: waves in order -> each wave's start inside the digit's run.
#pragma unroll
  for (int j = 0; j < PT; ++j) {
    const int dd = t * PT + j;
    uint32_t run = 0;
#pragma unroll
    for (int i = 0; i < kRsWaves; ++i) {
      const uint32_t c = s_cnt[i][dd];
      s_cnt[i][dd] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kRsItems; ++it) {
    const uint32_t j = (uint32_t)(w * kWaveItems + it * kWave + lane);
    if (j < tile_n) {
      const uint32_t d = ((rk[it] & lowmask) >> shift) & (NB - 1);
      const uint32_t lp = s_lbase[d] + s_cnt[w][d] + rr[it];
      s_keys[lp] = rk[it];
      s_vals[lp] = rv[it];
    }
  }
  __syncthreads();
  for (int it = 0; it < kRsItems; ++it) {
    const uint32_t j = (uint32_t)it * kRsThreads + t;
    if (j < tile_n) {
      const uint32_t key = s_keys[j];
      const uint32_t d = ((key & lowmask) >> shift) & (NB - 1);
      const uint64_t g = (uint64_t)s_gbase[d] + (j - s_lbase[d]);
      kout[g] = key;
      vout[g] = s_vals[j];
    }
  }
}

uint64_t radix_tmp_words(uint64_t n) {
  const uint64_t ntiles = (n + kRsTile - 1) / kRsTile;
  const uint64_t hist = (1ull << kRsDigit) * ntiles;
  // two key/value ping-pong buffers, raw and scanned tile histograms, totals
  return 4 * n + 2 * hist + (1ull << kRsDigit) + 64;
}


// Stable sort of (keys, vals) on key bits [0, bits).  Result in keys/vals.
// Passes after the first alternate between two scratch buffers and the last
// writes keys/vals, so no copy follows unless there is a single pass.
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint64_t n, int bits,
                            uint32_t* tmp, hipStream_t st, int* passes_out) {
  const int passes = bits <= 0 ? 0 : (bits + kRsDigit - 1) / kRsDigit;
  if (passes_out) *passes_out = passes;
  if (n <= 1 || passes == 0) return hipSuccess;
  const uint32_t lowmask = bits >= 32 ? 0xffffffffu : ((1u << bits) - 1u);
  const uint32_t ntiles = (uint32_t)((n + kRsTile - 1) / kRsTile);
  const uint64_t hsz = (1ull << kRsDigit) * ntiles;
  uint32_t* bk[2] = {tmp, tmp + 2 * n};
  uint32_t* bv[2] = {tmp + n, tmp + 3 * n};
  uint32_t* hraw = tmp + 4 * n;
  uint32_t* hscan = hraw + hsz;
  uint32_t* totals = hscan + hsz;
  const uint32_t* ka = keys;
  const uint32_t* va = vals;
  for (int p = 0; p < passes; ++p) {
    const bool last = p == passes - 1 && passes > 1;
    uint32_t* kb = last ? keys : bk[p & 1];
    uint32_t* vb = last ? vals : bv[p & 1];
    const int shift = kRsDigit * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, st, ka, n, shift, lowmask,
                       ntiles, hraw);
    hipLaunchKernelGGL(k_rs_rowscan, dim3(1u << kRsDigit), dim3(kScanThreads), 0, st, hraw,
                       ntiles, hscan, totals);
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, st, ka, va, kb, vb, n,
                       shift, lowmask, ntiles, hraw, hscan, totals);
    ka = kb;
    va = vb;
  }
  if (passes == 1) {
    hipError_t e = hipMemcpyAsync(keys, ka, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(vals, va, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace rtnn
