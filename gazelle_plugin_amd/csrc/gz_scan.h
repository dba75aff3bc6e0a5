// Exclusive scan over tiles of a [T][R] count matrix (tile-major: a tile's R
// counters are contiguous), in three launches over chunks of CH tiles: chunk
// sums, a scan of the chunk sums (which also gives each column's total), then
// the per-tile prefixes.  Every access walks tile rows, which are contiguous.
// Shared by the splitter and the partitioned aggregate (templates, so both
// translation units may instantiate them).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gz_device.h"

namespace gz {

template <int CH>
__global__ __launch_bounds__(256) void tile_scan_sum_kernel(const uint32_t* counts, int R, int64_t T,
                                                            uint64_t* csum) {
  int64_t c = blockIdx.x;
  int64_t t0 = c * CH, t1 = min<int64_t>(T, t0 + CH);
  for (int e = threadIdx.x; e < R; e += 256) {
    uint64_t s = 0;
    for (int64_t tb = t0; tb < t1; tb += 16) {
      uint32_t v[16];  // sixteen independent loads in flight
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = tb + k < t1 ? gld(counts + (tb + k) * R + e) : 0u;
#pragma unroll
      for (int k = 0; k < 16; k++) s += v[k];
    }
    csum[c * R + e] = s;
  }
}

// One block per column: exclusive scan of the column's chunk sums in place
// (thread i owns a contiguous run of chunks, block scan of the run sums); the
// column total goes to tot[e].  Grid = R.
__global__ __launch_bounds__(256) inline void tile_scan_chunks_kernel(uint64_t* csum, int R, int64_t C,
                                                                      uint64_t* tot) {
  __shared__ uint64_t s_scan[256 / kWave + 1];
  const int e = blockIdx.x;
  const int64_t per = (C + 255) / 256;
  const int64_t c0 = min<int64_t>(C, (int64_t)threadIdx.x * per), c1 = min<int64_t>(C, c0 + per);
  uint64_t s = 0;
  for (int64_t cb = c0; cb < c1; cb += 16) {
    uint64_t v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = cb + k < c1 ? csum[(cb + k) * R + e] : 0ull;
#pragma unroll
    for (int k = 0; k < 16; k++) s += v[k];
  }
  uint64_t total;
  uint64_t run = block_exclusive_scan<uint64_t>(s, s_scan, &total);
  for (int64_t cb = c0; cb < c1; cb += 16) {
    uint64_t v[16];
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = cb + k < c1 ? csum[(cb + k) * R + e] : 0ull;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      if (cb + k < c1) csum[(cb + k) * R + e] = run;
      run += v[k];
    }
  }
  if (threadIdx.x == 0) tot[e] = total;
}

template <int CH>
__global__ __launch_bounds__(256) void tile_scan_write_kernel(const uint32_t* counts, const uint64_t* cpre, int R,
                                                              int64_t T, uint32_t* prefix) {
  int64_t c = blockIdx.x;
  int64_t t0 = c * CH, t1 = min<int64_t>(T, t0 + CH);
  for (int e = threadIdx.x; e < R; e += 256) {
    uint64_t run = cpre[c * R + e];
    for (int64_t tb = t0; tb < t1; tb += 16) {
      uint32_t v[16];
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = tb + k < t1 ? gld(counts + (tb + k) * R + e) : 0u;
#pragma unroll
      for (int k = 0; k < 16; k++) {
        if (tb + k < t1) gst(prefix + (tb + k) * R + e, (uint32_t)run);
        run += v[k];
      }
    }
  }
}

}  // namespace gz
