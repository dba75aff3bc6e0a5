// Device-side helpers shared by the split, aggregate and join kernels (gfx950,
// wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gz {

constexpr int kWave = 64;

// ---- Spark Murmur3_x86_32 (hashInt / hashLong / hashUnsafeBytes) ------------
// The published MurmurHash3_x86_32 block and finalization steps; Spark's
// hashLong feeds the low then the high word of the key
// (third_party/murmurhash/murmurhash32.h:74-108 has the same arithmetic).
__device__ __forceinline__ uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
__device__ __forceinline__ uint32_t mix_k1(uint32_t k1) {
  k1 *= 0xcc9e2d51u;
  k1 = rotl32(k1, 15);
  return k1 * 0x1b873593u;
}
__device__ __forceinline__ uint32_t mix_h1(uint32_t h1, uint32_t k1) {
  h1 ^= k1;
  h1 = rotl32(h1, 13);
  return h1 * 5u + 0xe6546b64u;
}
__device__ __forceinline__ uint32_t fmix32(uint32_t h1, uint32_t len) {
  h1 ^= len;
  h1 ^= h1 >> 16;
  h1 *= 0x85ebca6bu;
  h1 ^= h1 >> 13;
  h1 *= 0xc2b2ae35u;
  h1 ^= h1 >> 16;
  return h1;
}
__device__ __forceinline__ uint32_t spark_hash_int(uint32_t v, uint32_t seed) {
  return fmix32(mix_h1(seed, mix_k1(v)), 4);
}
__device__ __forceinline__ uint32_t spark_hash_long(uint64_t v, uint32_t seed) {
  uint32_t h = mix_h1(seed, mix_k1((uint32_t)v));
  h = mix_h1(h, mix_k1((uint32_t)(v >> 32)));
  return fmix32(h, 8);
}

// pmod(h, P) for the int32 hash h: floored modulo, equal to the reference's
// `pid = h % P; if (pid < 0) pid += P` (splitter.cc:1456-1465).  Computed with
// Lemire's fastmod on x = h + 2^31 (unsigned), then shifted back by 2^31 mod P.
struct FastMod {
  uint64_t M;    // floor((2^64 - 1) / P) + 1
  uint32_t P;
  uint32_t c;    // 2^31 mod P
};
__device__ __forceinline__ uint32_t fastmod_u32(uint32_t x, const FastMod& f) {
  uint64_t low = f.M * (uint64_t)x;
  return (uint32_t)__umul64hi(low, (uint64_t)f.P);
}
__device__ __forceinline__ int pmod_hash(uint32_t h, const FastMod& f) {
  uint32_t x = h ^ 0x80000000u;  // (int32)h + 2^31 as unsigned
  int r = (int)fastmod_u32(x, f) - (int)f.c;
  return r < 0 ? r + (int)f.P : r;
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ---- global memory ------------------------------------------------------------
// Pointers read from LDS or from pointer tables are generic to the compiler, so
// it emits flat instructions, which count against both the vector-memory and the
// LDS counters and make LDS traffic wait on HBM loads.  These go through the
// global address space instead.
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// ---- bits ---------------------------------------------------------------------
__device__ __forceinline__ int bit_get(const uint8_t* bm, int64_t i) {
  return (gld(bm + (i >> 3)) >> (i & 7)) & 1;
}

// ---- join key normalisation ---------------------------------------------------
// Spark's "normalize" of a floating join key, normalize_nan_zero
// (precompile/gandiva.h:214-222): every NaN becomes one NaN and |x| < 1e-7
// becomes 0.0.  On raw key bits of width 4 (float) or 8 (double).
__device__ __forceinline__ double norm_nan_zero(double d) {
  if (d != d) return __longlong_as_double(0x7ff8000000000000ll);
  return fabs(d) < 0.0000001 ? 0.0 : d;
}
__device__ __forceinline__ uint64_t norm_key_bits(uint64_t raw, int width) {
  if (width == 8) return (uint64_t)__double_as_longlong(norm_nan_zero(__longlong_as_double((long long)raw)));
  const float f = (float)norm_nan_zero((double)__uint_as_float((uint32_t)raw));
  return (uint64_t)__float_as_uint(f);
}

// ---- wave helpers -------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  int l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Lanes of the wave holding the same 8-bit value v (all 64 lanes participate,
// inactive lanes pass a value outside [0, 256) together with active=false).
__device__ __forceinline__ uint64_t match_u8(uint32_t v, bool active) {
  uint64_t m = __ballot(active);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    uint64_t bal = __ballot(active && ((v >> b) & 1));
    m &= ((v >> b) & 1) ? bal : ~bal;
  }
  return active ? m : 0ull;
}
// Same for a value of `bits` bits (up to 16).
__device__ __forceinline__ uint64_t match_bits(uint32_t v, bool active, int bits) {
  uint64_t m = __ballot(active);
  for (int b = 0; b < bits; b++) {
    uint64_t bal = __ballot(active && ((v >> b) & 1));
    m &= ((v >> b) & 1) ? bal : ~bal;
  }
  return active ? m : 0ull;
}

// Inclusive wave prefix sum.
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  int l = __lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    T n = __shfl_up(v, o, kWave);
    if (l >= o) v += n;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_reduce_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Exclusive block-wide scan of one value per thread; the block total goes to
// *total when non-null.  `smem` must hold blockDim.x / 64 + 1 entries; it may be
// reused once the call returns.
template <typename T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* smem, T* total) {
  int w = threadIdx.x / kWave, nw = blockDim.x / kWave, l = __lane_id();
  T inc = wave_inclusive_scan(v);
  if (l == kWave - 1) smem[w] = inc;
  __syncthreads();
  if (w == 0) {
    T s = l < nw ? smem[l] : T(0);
    T si = wave_inclusive_scan(s);
    if (l < nw) smem[l] = si - s;
    if (l == nw - 1) smem[nw] = si;
  }
  __syncthreads();
  T ex = smem[w] + inc - v;
  if (total) *total = smem[nw];
  __syncthreads();
  return ex;
}

// XCD-aware tile index: consecutive blocks are dispatched round-robin over the
// 8 XCDs of MI355X, so block b runs on XCD b % 8.  Giving each XCD a contiguous
// range of tiles keeps neighbouring tiles (which share output cache lines at
// partition-run boundaries) in the same L2.  grid = 8 * ceil(T / 8).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t grid, int64_t T) {
  int64_t per = grid / 8;
  int64_t t = (b % 8) * per + b / 8;
  return t < T ? t : -1;
}

}  // namespace gz
