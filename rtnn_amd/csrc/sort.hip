// Device-wide primitives written for gfx950 (wave64): exclusive scan of u32
// and a stable LSD radix sort of (u32 key, u32 value) pairs.
//
// They replace the Thrust calls of the reference's grid build
// (src/optixNSearch/thrust_helper.cu:13-99: sort_by_key, exclusive_scan) that
// sort.cpp:535-612 uses for the counting sort of particles into cells.
//
// Radix sort, one 8-bit digit per pass, three kernels per pass:
//   rs_hist    : per 4096-key tile, a 256-bin digit histogram (LDS atomics),
//                stored digit-major so one exclusive scan gives every tile's
//                global base for every digit;
//   scan       : exclusive scan of that table;
//   rs_scatter : per tile, stable local ranks from wave64 ballot matching
//                (8 ballots per key) plus per-wave digit counts, keys staged
//                in LDS in digit order, then written out so consecutive
//                lanes store consecutive addresses of one digit run.
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

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* __restrict__ keys,
                                                        uint64_t n, int shift, uint32_t ntiles,
                                                        uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll 4
  for (int i = 0; i < kRsItems; ++i) {
    uint64_t e = base + (uint64_t)i * kRsThreads + threadIdx.x;
    const bool valid = e < n;
    const uint32_t d = valid ? (keys[e] >> shift) & 255u : 0u;
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
  hist[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__device__ inline uint64_t lanemask_lt() {
  const int lane = threadIdx.x & (kWave - 1);
  return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout, uint64_t n, int shift,
    uint32_t ntiles, const uint32_t* __restrict__ hist_raw,
    const uint32_t* __restrict__ hist_scan) {
  __shared__ uint32_t s_keys[kRsTile];
  __shared__ uint32_t s_vals[kRsTile];
  __shared__ uint32_t s_lbase[256];    // tile-local start of each digit
  __shared__ uint32_t s_gbase[256];    // global start of this tile's digit run
  __shared__ uint32_t s_run[256];      // keys of each digit placed so far
  __shared__ uint32_t s_wcnt[kRsWaves][256];
  __shared__ uint32_t s_wofs[kRsWaves][256];
  __shared__ uint32_t s_scan[kRsThreads / kWave + 1];

  const int t = threadIdx.x;
  const int w = t / kWave;
  const uint64_t base = (uint64_t)blockIdx.x * kRsTile;
  const uint32_t tile_n = (uint32_t)((n - base) < (uint64_t)kRsTile ? (n - base) : kRsTile);

  {
    uint32_t c = hist_raw[(uint64_t)t * ntiles + blockIdx.x];
    s_gbase[t] = hist_scan[(uint64_t)t * ntiles + blockIdx.x];
    uint32_t tot;
    // block_excl_scan expects kScanThreads == kRsThreads threads.
    uint32_t lane_ex = block_excl_scan(c, s_scan, tot);
    s_lbase[t] = lane_ex;
    s_run[t] = 0;
#pragma unroll
    for (int i = 0; i < kRsWaves; ++i) s_wcnt[i][t] = 0;
  }
  __syncthreads();

  for (int it = 0; it < kRsItems; ++it) {
    const uint32_t j = (uint32_t)it * kRsThreads + t;
    const bool valid = j < tile_n;
    uint32_t key = 0, val = 0, d = 0;
    if (valid) {
      key = kin[base + j];
      val = vin[base + j];
      d = (key >> shift) & 255u;
    }
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      match &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(match & lanemask_lt());
    if (valid && rank == 0) s_wcnt[w][d] = (uint32_t)__popcll(match);
    __syncthreads();
    {
      uint32_t run = s_run[t];
#pragma unroll
      for (int i = 0; i < kRsWaves; ++i) {
        uint32_t c = s_wcnt[i][t];
        s_wofs[i][t] = run;
        run += c;
        s_wcnt[i][t] = 0;
      }
      s_run[t] = run;
    }
    __syncthreads();
    if (valid) {
      const uint32_t lp = s_lbase[d] + s_wofs[w][d] + rank;
      s_keys[lp] = key;
      s_vals[lp] = val;
    }
  }
  __syncthreads();
  for (int it = 0; it < kRsItems; ++it) {
    const uint32_t j = (uint32_t)it * kRsThreads + t;
    if (j < tile_n) {
      const uint32_t key = s_keys[j];
      const uint32_t d = (key >> shift) & 255u;
      const uint64_t g = (uint64_t)s_gbase[d] + (j - s_lbase[d]);
      kout[g] = key;
      vout[g] = s_vals[j];
    }
  }
}

uint64_t radix_tmp_words(uint64_t n) {
  uint64_t ntiles = (n + kRsTile - 1) / kRsTile;
  uint64_t hist = 256 * ntiles;
  // keys + vals ping-pong, raw hist, scanned hist (+1), scan partials
  return 2 * n + hist + (hist + 1) + (uint64_t)scan_tmp_words(hist) + 64;
}

// Stable sort of (keys, vals) on key bits [0, bits).  Result in keys/vals.
hipError_t radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint64_t n, int bits,
                            uint32_t* tmp, hipStream_t st, int* passes_out) {
  int passes = (bits + 7) / 8;
  if (passes_out) *passes_out = passes;
  if (n <= 1 || passes == 0) return hipSuccess;
  uint32_t ntiles = (uint32_t)((n + kRsTile - 1) / kRsTile);
  uint64_t hsz = 256ull * ntiles;
  uint32_t* k2 = tmp;
  uint32_t* v2 = k2 + n;
  uint32_t* hraw = v2 + n;
  uint32_t* hscan = hraw + hsz;
  uint32_t* stmp = hscan + hsz + 1;
  uint32_t *ka = keys, *va = vals, *kb = k2, *vb = v2;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(kRsThreads), 0, st, ka, n, shift, ntiles,
                       hraw);
    hipError_t e = exclusive_scan(hraw, hscan, hsz, stmp, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(kRsThreads), 0, st, ka, va, kb, vb, n,
                       shift, ntiles, hraw, hscan);
    uint32_t* t;
    t = ka; ka = kb; kb = t;
    t = va; va = vb; vb = t;
  }
  if (ka != keys) {
    hipError_t e = hipMemcpyAsync(keys, ka, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(vals, va, n * 4, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace rtnn
