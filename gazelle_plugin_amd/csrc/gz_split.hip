// Shuffle hash Splitter for MI355X (gfx950).
//
// Replaces sparkcolumnarplugin::shuffle::Splitter (cpp/src/shuffle/splitter.cc).
// One call (one batch, or many batches at once) produces one *segment*: the rows
// of every partition contiguous and in input order, which is the order the
// reference's DoSplit appends them in (splitter.cc:855-938).
//
// Three launches per segment:
//   split_count   one workgroup per tile: partition id of every row (Spark
//                 Murmur3 hashLong seed 42 + pmod, or rr / range / single), an
//                 LDS histogram of rows and of utf8 bytes per partition, null
//                 counts.  Reads the key (and utf8 offsets) only.
//   split_scan    exclusive scan of the per-tile counts, partition-major, so each
//                 tile knows where its run of partition p starts.
//   split_scatter one workgroup per tile: a stable in-tile rank per partition from
//                 wave ballots, an LDS permutation into partition order, then
//                 coalesced writes of each column's runs; validity and boolean bits
//                 are cleared with atomic AND in bitmaps pre-filled with ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

constexpr int kMaxCols = 32;
constexpr int kMaxKeys = 4;
constexpr int kMaxPartitions = 2048;

enum ColKind : int { K_NA = 0, K_FIXED = 1, K_BOOL = 2, K_BINARY = 3 };

struct TileDesc {
  int32_t batch;
  int32_t row_begin;
  int32_t nrows;
  int32_t rr;  // round-robin partition of the tile's first row
};

struct SplitParams {
  const TileDesc* tiles;
  int64_t T;
  const void* const* bufs;  // bufs[batch * bstride + col * 3 + k], k: 0 validity, 1 offsets, 2 values
  int bstride;
  int mode;  // 0 hash, 1 rr, 2 range, 3 single
  int P;
  int pbits;
  FastMod fm;
  uint32_t seed;
  int nkeys;
  int key_col[kMaxKeys];
  int key_type[kMaxKeys];
  int ncols;                 // input columns
  int kind[kMaxCols];
  int width[kMaxCols];
  int out_col[kMaxCols];     // output column of an input column, -1 if dropped
  int bin[kMaxCols];         // binary column ordinal, -1 otherwise
  int nbin;
  int bin_col0;              // input column of the first binary column
  // per-segment buffers
  uint32_t* counts;          // [P][T]
  uint32_t* prefix;          // [P][T] exclusive within partition
  uint32_t* bytes;           // [nbin][P][T]
  uint32_t* bprefix;         // [nbin][P][T]
  const int64_t* row_start;  // [P+1]
  const int64_t* byte_start; // [nbin][P+1]
  void* out_values[kMaxCols];
  uint32_t* out_valid[kMaxCols];
  int32_t* out_off[kMaxCols];
  uint8_t* out_data[kMaxCols];
  unsigned long long* null_counts;  // [ncols]
  int* err;
};

// Buffer pointers of the tile's batch, staged in LDS once per workgroup.
struct TilePtrs {
  const void* v[kMaxCols * 3];
};

__device__ __forceinline__ void load_tile_ptrs(const SplitParams& p, int batch, const void** s_bp) {
  for (int i = threadIdx.x; i < p.ncols * 3; i += blockDim.x)
    s_bp[i] = p.bufs[(int64_t)batch * p.bstride + i];
}

// Spark hash of key column c at row r, chained through h (Murmur3Hash semantics:
// a null value leaves the running hash unchanged).
__device__ __forceinline__ uint32_t hash_key(const void* const* bp, int c, int t, int64_t r, uint32_t h) {
  const uint8_t* vb = (const uint8_t*)bp[c * 3 + 0];
  if (vb && !bit_get(vb, r)) return h;
  const void* v = bp[c * 3 + 2];
  switch (t) {
    case GZ_TYPE_INT64: return spark_hash_long((uint64_t)((const int64_t*)v)[r], h);
    case GZ_TYPE_INT32:
    case GZ_TYPE_DATE32: return spark_hash_int((uint32_t)((const int32_t*)v)[r], h);
    case GZ_TYPE_BOOL: return spark_hash_int(bit_get((const uint8_t*)v, r), h);
    case GZ_TYPE_INT8: return spark_hash_int((uint32_t)(int32_t)((const int8_t*)v)[r], h);
    case GZ_TYPE_INT16: return spark_hash_int((uint32_t)(int32_t)((const int16_t*)v)[r], h);
    case GZ_TYPE_FLOAT: {
      float f = ((const float*)v)[r];
      uint32_t bits;
      if (f == 0.0f) f = 0.0f;
      if (f != f) bits = 0x7fc00000u; else bits = __float_as_uint(f);
      return spark_hash_int(bits, h);
    }
    case GZ_TYPE_DOUBLE: {
      double d = ((const double*)v)[r];
      uint64_t bits;
      if (d == 0.0) d = 0.0;
      if (d != d) bits = 0x7ff8000000000000ull; else bits = (uint64_t)__double_as_longlong(d);
      return spark_hash_long(bits, h);
    }
    case GZ_TYPE_STRING:
    case GZ_TYPE_BINARY: {
      const int32_t* off = (const int32_t*)bp[c * 3 + 1];
      const uint8_t* d = (const uint8_t*)v + off[r];
      int len = off[r + 1] - off[r];
      int aligned = len & ~3;
      for (int i = 0; i < aligned; i += 4) {
        uint32_t w = (uint32_t)d[i] | ((uint32_t)d[i + 1] << 8) | ((uint32_t)d[i + 2] << 16) |
                     ((uint32_t)d[i + 3] << 24);
        h = mix_h1(h, mix_k1(w));
      }
      for (int i = aligned; i < len; i++) h = mix_h1(h, mix_k1((uint32_t)(int32_t)(int8_t)d[i]));
      return fmix32(h, (uint32_t)len);
    }
    default: return h;
  }
}

// Partition id of row lr of a tile.  KEYK = 1: one int64 key without
// validity (the hot case), else the generic path.
template <int KEYK>
__device__ __forceinline__ int compute_pid(const SplitParams& p, const void* const* bp, const TileDesc& td,
                                           int lr) {
  int64_t r = (int64_t)td.row_begin + lr;
  if (KEYK == 1) {
    uint64_t k = (uint64_t)((const int64_t*)bp[p.key_col[0] * 3 + 2])[r];
    return pmod_hash(spark_hash_long(k, p.seed), p.fm);
  }
  if (p.mode == 0) {
    uint32_t h = p.seed;
    for (int k = 0; k < p.nkeys; k++) h = hash_key(bp, p.key_col[k], p.key_type[k], r, h);
    return pmod_hash(h, p.fm);
  }
  if (p.mode == 1) return (int)fastmod_u32((uint32_t)(td.rr + lr), p.fm);
  if (p.mode == 2) {
    int v = ((const int32_t*)bp[2])[r];
    if (v < 0 || v >= p.P) {
      atomicOr(p.err, 1);
      return 0;
    }
    return v;
  }
  return 0;
}

// ---------------------------------------------------------------- pass 1
template <int BLOCK, int KEYK>
__global__ __launch_bounds__(BLOCK) void split_count_kernel(SplitParams p) {
  extern __shared__ uint32_t s_cnt[];  // [P] rows, then [nbin][P] bytes
  __shared__ const void* s_bp[kMaxCols * 3];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const TileDesc td = p.tiles[t];
  const int P = p.P;
  load_tile_ptrs(p, td.batch, s_bp);
  for (int i = threadIdx.x; i < P * (1 + p.nbin); i += BLOCK) s_cnt[i] = 0;
  __syncthreads();
  if (KEYK == 1 && p.nbin == 1) {
    // hot path: int64 key, one utf8 column; four independent rows per step
    const int64_t* key = (const int64_t*)s_bp[p.key_col[0] * 3 + 2] + td.row_begin;
    int bc = p.bin_col0;
    const int32_t* off = (const int32_t*)s_bp[bc * 3 + 1] + td.row_begin;
    for (int lr0 = threadIdx.x; lr0 < td.nrows; lr0 += 4 * BLOCK) {
      uint64_t k[4];
      int32_t o0[4], o1[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        int lr = lr0 + u * BLOCK;
        bool a = lr < td.nrows;
        k[u] = a ? (uint64_t)key[lr] : 0;
        o0[u] = a ? off[lr] : 0;
        o1[u] = a ? off[lr + 1] : 0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        if (lr0 + u * BLOCK < td.nrows) {
          int pid = pmod_hash(spark_hash_long(k[u], p.seed), p.fm);
          atomicAdd(&s_cnt[pid], 1u);
          atomicAdd(&s_cnt[P + pid], (uint32_t)(o1[u] - o0[u]));
        }
      }
    }
  } else {
    for (int lr = threadIdx.x; lr < td.nrows; lr += BLOCK) {
      int pid = compute_pid<KEYK>(p, s_bp, td, lr);
      atomicAdd(&s_cnt[pid], 1u);
      if (p.nbin) {
        int64_t r = (int64_t)td.row_begin + lr;
        for (int c = 0; c < p.ncols; c++) {
          int j = p.bin[c];
          if (j < 0) continue;
          const int32_t* off = (const int32_t*)s_bp[c * 3 + 1];
          atomicAdd(&s_cnt[P * (1 + j) + pid], (uint32_t)(off[r + 1] - off[r]));
        }
      }
    }
  }
  // null counts: zero bits of each validity bitmap over the tile's rows
  // (tiles start on multiples of 8 rows inside a batch).
  for (int c = 0; c < p.ncols; c++) {
    if (p.out_col[c] < 0 || p.kind[c] == K_NA) continue;
    const uint8_t* vb = (const uint8_t*)s_bp[c * 3 + 0];
    if (!vb) continue;
    uint32_t nulls = 0;
    int nbytes = (td.nrows + 7) >> 3;
    const uint8_t* base = vb + (td.row_begin >> 3);
    for (int i = threadIdx.x; i < nbytes; i += BLOCK) {
      uint32_t b = base[i];
      int valid_bits = min(8, td.nrows - i * 8);
      uint32_t mask = (1u << valid_bits) - 1u;
      nulls += __popc((~b) & mask);
    }
    nulls = wave_reduce_sum(nulls);
    if (__lane_id() == 0 && nulls) atomicAdd(&p.null_counts[p.out_col[c]], (unsigned long long)nulls);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P; i += BLOCK) p.counts[(int64_t)i * p.T + t] = s_cnt[i];
  for (int j = 0; j < p.nbin; j++)
    for (int i = threadIdx.x; i < P; i += BLOCK)
      p.bytes[((int64_t)j * P + i) * p.T + t] = s_cnt[P * (1 + j) + i];
}

// ---------------------------------------------------------------- scan
// One block per row (partition, or (binary column, partition)): exclusive scan
// over T tiles; the row total goes to tot[row].
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void split_scan_kernel(const uint32_t* in, uint32_t* out,
                                                            uint64_t* tot, int64_t T) {
  __shared__ uint64_t s_scan[BLOCK / kWave + 1];
  int64_t row = blockIdx.x;
  const uint32_t* a = in + row * T;
  uint32_t* o = out + row * T;
  uint64_t carry = 0;
  for (int64_t base = 0; base < T; base += (int64_t)BLOCK * 8) {
    // each thread scans 8 consecutive entries
    int64_t i0 = base + (int64_t)threadIdx.x * 8;
    uint32_t v[8];
    uint64_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      v[k] = (i0 + k < T) ? a[i0 + k] : 0u;
      s += v[k];
    }
    uint64_t total;
    uint64_t ex = block_exclusive_scan<uint64_t>(s, s_scan, &total);
    uint64_t run = carry + ex;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (i0 + k < T) o[i0 + k] = (uint32_t)run;
      run += v[k];
    }
    carry += total;
  }
  if (threadIdx.x == 0) tot[row] = carry;
}

// One block: exclusive scans of the partition totals into row_start / byte_start,
// first utf8 offset of every partition run set to 0, error on a partition whose
// bytes exceed int32 offsets.
__global__ __launch_bounds__(256) void split_layout_kernel(const uint64_t* tot, int P, int nbin,
                                                            int64_t* row_start,
                                                            int64_t* byte_start,
                                                            int32_t** out_off, int* err) {
  __shared__ uint64_t s_scan[256 / kWave + 1];
  for (int j = -1; j < nbin; j++) {
    const uint64_t* tt = tot + (int64_t)(j + 1) * P;
    int64_t* dst = j < 0 ? row_start : byte_start + (int64_t)j * (P + 1);
    uint64_t carry = 0;
    for (int base = 0; base < P; base += 256) {
      int i = base + threadIdx.x;
      uint64_t v = i < P ? tt[i] : 0;
      if (j >= 0 && v >= 0x80000000ull) atomicOr(err, 2);
      uint64_t total;
      uint64_t ex = block_exclusive_scan<uint64_t>(v, s_scan, &total);
      if (i < P) dst[i] = (int64_t)(carry + ex);
      carry += total;
    }
    if (threadIdx.x == 0) dst[P] = (int64_t)carry;
    __syncthreads();
  }
  // offsets[row_start[p] + p] = 0 for every utf8 column
  for (int j = 0; j < nbin; j++)
    for (int i = threadIdx.x; i < P; i += 256) out_off[j][row_start[i] + i] = 0;
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// ---------------------------------------------------------------- pass 2
// Copy len bytes: byte head to a 4-byte aligned destination, 4-byte stores in
// the middle (unaligned 4-byte loads), byte tail.  Exactly [dst, dst+len) is
// written, so neighbouring strings owned by other threads are untouched.
__device__ __forceinline__ void copy_bytes(uint8_t* dst, const uint8_t* src, int len) {
  int i = 0;
  int head = (int)((4 - ((uintptr_t)dst & 3)) & 3);
  if (head > len) head = len;
  for (; i < head; i++) dst[i] = src[i];
  for (; i + 4 <= len; i += 4) {
    uint32_t w;
    __builtin_memcpy(&w, src + i, 4);
    *(uint32_t*)(dst + i) = w;
  }
  for (; i < len; i++) dst[i] = src[i];
}

// LDS layout of the scatter kernel (bytes from the start of dynamic LDS).
struct ScatterLds {
  size_t scan, pbase, bbase, brel, toff, wcnt, bp, pid, pos, spid, stage, total;
};
__host__ __device__ inline ScatterLds scatter_layout(int TILE, int BLOCK, int P) {
  ScatterLds L;
  size_t o = 0;
  auto take = [&](size_t bytes, size_t align) {
    o = (o + align - 1) / align * align;
    size_t r = o;
    o += bytes;
    return r;
  };
  int NW = BLOCK / 64;
  L.scan = take(8 * (NW + 1), 8);
  L.pbase = take(8 * (size_t)P, 8);
  L.bbase = take(8 * (size_t)P, 8);
  L.brel = take(4 * (size_t)P, 4);
  L.toff = take(4 * (size_t)(P + 1), 4);
  L.wcnt = take(2 * (size_t)NW * P, 4);
  L.bp = take(8 * (size_t)kMaxCols * 3, 8);
  L.pid = take(2 * (size_t)TILE, 4);
  L.pos = take(2 * (size_t)TILE, 4);
  L.spid = take(2 * (size_t)TILE, 4);
  L.stage = take((size_t)TILE * 8 + 4, 16);
  L.total = o;
  return L;
}

template <int TILE, int BLOCK, int KEYK>
__global__ __launch_bounds__(BLOCK) void split_scatter_kernel(SplitParams p) {
  constexpr int NW = BLOCK / kWave;
  constexpr int RW = TILE / NW;  // rows per wave in the ranking pass
  static_assert(RW % kWave == 0, "tile must be a multiple of the block");
  extern __shared__ __align__(16) uint8_t smem[];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const TileDesc td = p.tiles[t];
  const int n = td.nrows;
  const int P = p.P;
  const int tid = threadIdx.x, w = tid / kWave, lane = __lane_id();
  const ScatterLds Lo = scatter_layout(TILE, BLOCK, P);
  uint64_t* s_scan = (uint64_t*)(smem + Lo.scan);
  int64_t* s_pbase = (int64_t*)(smem + Lo.pbase);
  int64_t* s_bbase = (int64_t*)(smem + Lo.bbase);
  uint32_t* s_brel = (uint32_t*)(smem + Lo.brel);
  uint32_t* s_toff = (uint32_t*)(smem + Lo.toff);
  uint16_t* s_wcnt = (uint16_t*)(smem + Lo.wcnt);
  const void** s_bp = (const void**)(smem + Lo.bp);
  uint16_t* s_pid = (uint16_t*)(smem + Lo.pid);
  uint16_t* s_pos = (uint16_t*)(smem + Lo.pos);
  uint16_t* s_spid = (uint16_t*)(smem + Lo.spid);
  uint8_t* s_stage = smem + Lo.stage;

  load_tile_ptrs(p, td.batch, s_bp);
  for (int i = tid; i < NW * P; i += BLOCK) s_wcnt[i] = 0;
  __syncthreads();

  // ---- 1: partition id and stable rank of every row.  Wave w ranks rows
  // [w*RW, (w+1)*RW) in chunks of 64 consecutive rows.
  for (int j = 0; j < RW / kWave; j++) {
    int lr = w * RW + j * kWave + lane;
    bool act = lr < n;
    int pid = act ? compute_pid<KEYK>(p, s_bp, td, lr) : 0;
    uint64_t m = match_bits((uint32_t)pid, act, p.pbits);
    uint64_t lt = m & lanemask_lt();
    int before = act ? s_wcnt[w * P + pid] : 0;
    if (act) {
      s_pid[lr] = (uint16_t)pid;
      s_pos[lr] = (uint16_t)(before + __popcll(lt));
      if (lt == 0) s_wcnt[w * P + pid] = (uint16_t)(before + __popcll(m));
    }
  }
  __syncthreads();

  // ---- 2: per partition: wave offsets (in place), tile count, tile offsets,
  // destination bases
  {
    constexpr int PPT = (kMaxPartitions + BLOCK - 1) / BLOCK;
    uint32_t sum = 0;
    for (int k = 0; k < PPT; k++) {
      int pp = tid * PPT + k;
      if (pp >= P) break;
      uint32_t c = 0;
      for (int ww = 0; ww < NW; ww++) {
        uint32_t x = s_wcnt[ww * P + pp];
        s_wcnt[ww * P + pp] = (uint16_t)c;
        c += x;
      }
      s_toff[pp] = c;  // count for now
      sum += c;
    }
    uint64_t total;
    uint64_t ex = block_exclusive_scan<uint64_t>(sum, s_scan, &total);
    uint32_t run = (uint32_t)ex;
    for (int k = 0; k < PPT; k++) {
      int pp = tid * PPT + k;
      if (pp >= P) break;
      uint32_t c = s_toff[pp];
      s_toff[pp] = run;
      s_pbase[pp] = p.row_start[pp] + p.prefix[(int64_t)pp * p.T + t];
      run += c;
    }
    if (tid == 0) s_toff[P] = (uint32_t)total;
  }
  __syncthreads();

  // ---- 3: sorted position of every row; pid of every sorted position
  for (int lr = tid; lr < n; lr += BLOCK) {
    int pid = s_pid[lr];
    int spos = (int)s_toff[pid] + s_wcnt[(lr / RW) * P + pid] + s_pos[lr];
    s_pos[lr] = (uint16_t)spos;
    s_spid[spos] = (uint16_t)pid;
  }
  __syncthreads();

  // ---- 4: columns
  const int64_t rb = td.row_begin;
  for (int c = 0; c < p.ncols; c++) {
    int oc = p.out_col[c];
    if (oc < 0) continue;
    int kind = p.kind[c];
    if (kind == K_NA) continue;
    const uint8_t* vb = (const uint8_t*)s_bp[c * 3 + 0];
    const uint8_t* src = (const uint8_t*)s_bp[c * 3 + 2];

    // validity: clear the bit of every null row at its destination
    if (vb && p.out_valid[oc]) {
      uint32_t* ov = p.out_valid[oc];
      for (int lr = tid; lr < n; lr += BLOCK) {
        if (!bit_get(vb, rb + lr)) {
          int pid = s_pid[lr];
          int64_t d = s_pbase[pid] + ((int)s_pos[lr] - (int)s_toff[pid]);
          atomicAnd(&ov[d >> 5], ~(1u << (d & 31)));
        }
      }
    }

    if (kind == K_BOOL) {
      uint32_t* ov = (uint32_t*)p.out_values[oc];
      for (int lr = tid; lr < n; lr += BLOCK) {
        if (!bit_get(src, rb + lr)) {
          int pid = s_pid[lr];
          int64_t d = s_pbase[pid] + ((int)s_pos[lr] - (int)s_toff[pid]);
          atomicAnd(&ov[d >> 5], ~(1u << (d & 31)));
        }
      }
    } else if (kind == K_FIXED) {
      const int W = p.width[c];
      const int parts = W > 8 ? W / 8 : 1;  // 16-byte values move as two 8-byte halves
      const int EW = W > 8 ? 8 : W;
      uint8_t* dst = (uint8_t*)p.out_values[oc];
      for (int part = 0; part < parts; part++) {
        if (EW == 8) {
          const uint64_t* s8 = (const uint64_t*)src;
          for (int lr = tid; lr < n; lr += BLOCK)
            ((uint64_t*)s_stage)[s_pos[lr]] = s8[(rb + lr) * parts + part];
        } else if (EW == 4) {
          const uint32_t* s4 = (const uint32_t*)src;
          for (int lr = tid; lr < n; lr += BLOCK) ((uint32_t*)s_stage)[s_pos[lr]] = s4[rb + lr];
        } else if (EW == 2) {
          const uint16_t* s2 = (const uint16_t*)src;
          for (int lr = tid; lr < n; lr += BLOCK) ((uint16_t*)s_stage)[s_pos[lr]] = s2[rb + lr];
        } else {
          for (int lr = tid; lr < n; lr += BLOCK) s_stage[s_pos[lr]] = src[rb + lr];
        }
        __syncthreads();
        for (int i = tid; i < n; i += BLOCK) {
          int pid = s_spid[i];
          int64_t d = s_pbase[pid] + (i - (int)s_toff[pid]);
          if (EW == 8) ((uint64_t*)dst)[d * parts + part] = ((uint64_t*)s_stage)[i];
          else if (EW == 4) ((uint32_t*)dst)[d] = ((uint32_t*)s_stage)[i];
          else if (EW == 2) ((uint16_t*)dst)[d] = ((uint16_t*)s_stage)[i];
          else dst[d] = s_stage[i];
        }
        __syncthreads();
      }
    } else if (kind == K_BINARY) {
      const int j_bin = p.bin[c];
      const int32_t* soff = (const int32_t*)s_bp[c * 3 + 1];
      uint32_t* s_src = (uint32_t*)s_stage;  // [TILE] source byte offset
      uint32_t* s_boff = s_src + TILE;       // [TILE+1] sorted byte prefix
      for (int lr = tid; lr < n; lr += BLOCK) {
        int o0 = soff[rb + lr], o1 = soff[rb + lr + 1];
        int sp = s_pos[lr];
        s_src[sp] = (uint32_t)o0;
        s_boff[sp] = (uint32_t)(o1 - o0);
      }
      for (int pp = tid; pp < P; pp += BLOCK) {
        uint32_t rel = p.bprefix[((int64_t)j_bin * P + pp) * p.T + t];
        s_brel[pp] = rel;
        s_bbase[pp] = p.byte_start[(int64_t)j_bin * (P + 1) + pp] + rel;
      }
      __syncthreads();
      {
        constexpr int EPT = TILE / BLOCK;
        uint32_t s = 0;
        int i0 = tid * EPT;
        for (int k = 0; k < EPT; k++)
          if (i0 + k < n) s += s_boff[i0 + k];
        uint64_t total;
        uint64_t ex = block_exclusive_scan<uint64_t>(s, s_scan, &total);
        uint32_t run = (uint32_t)ex;
        for (int k = 0; k < EPT; k++) {
          if (i0 + k < n) {
            uint32_t v = s_boff[i0 + k];
            s_boff[i0 + k] = run;
            run += v;
          }
        }
        if (tid == 0) s_boff[n] = (uint32_t)total;
      }
      __syncthreads();
      uint8_t* odata = p.out_data[oc];
      int32_t* ooff = p.out_off[oc];
      for (int i = tid; i < n; i += BLOCK) {
        int pid = s_spid[i];
        uint32_t toff = s_toff[pid];
        uint32_t run0 = s_boff[toff];
        uint32_t b0 = s_boff[i], b1 = s_boff[i + 1];
        int len = (int)(b1 - b0);
        // offsets of partition pid start at index row_start + pid; this row is
        // (s_pbase - row_start) + (i - toff) rows into the partition
        ooff[s_pbase[pid] + pid + 1 + (i - (int)toff)] = (int32_t)(s_brel[pid] + (b0 - run0) + len);
        copy_bytes(odata + s_bbase[pid] + (b0 - run0), src + s_src[i], len);
      }
      __syncthreads();
    }
  }
}

// Host-side LDS bytes of the scatter kernel.
static size_t scatter_lds(int TILE, int BLOCK, int P) { return scatter_layout(TILE, BLOCK, P).total; }

}  // namespace gz

// =============================================================== host side
using namespace gz;

struct SplitSegment {
  int64_t N = 0;
  int32_t nbatches = 0;
  int64_t T = 0;
  DBuf counts, prefix, bytes, bprefix, tot, row_start, byte_start, null_counts, err, offtab;
  std::vector<DBuf> values, validity, offsets, data;
  std::vector<int32_t> tile_batch;  // batch of each tile
  // host copies, valid after gz_split_stop
  std::vector<int64_t> h_row_start, h_byte_start;
  std::vector<unsigned long long> h_null;
  bool resolved = false;
};

struct gz_splitter {
  gz_context* ctx = nullptr;
  int mode = 0;
  int P = 0;
  int ncols_in = 0;
  std::vector<int> types_in;
  std::vector<int> out_types;
  std::vector<int> keys;
  gz_split_options opts;
  int32_t rr_state = 0;
  int tile_rows = 4096;
  SplitParams base;  // constant part of the kernel parameters
  std::vector<SplitSegment*> segments;
  DBuf d_tiles, d_bufs;
  bool stopped = false;
  int64_t compute_pid_ns = 0;
};

static int split_release_segment(gz_context* ctx, SplitSegment* s) {
  for (DBuf* b : {&s->counts, &s->prefix, &s->bytes, &s->bprefix, &s->tot, &s->row_start,
                  &s->byte_start, &s->null_counts, &s->err, &s->offtab})
    ctx_free(ctx, b);
  for (auto* v : {&s->values, &s->validity, &s->offsets, &s->data})
    for (auto& b : *v) ctx_free(ctx, &b);
  delete s;
  return GZ_OK;
}

static FastMod make_fastmod(uint32_t P) {
  FastMod f;
  f.P = P;
  f.M = UINT64_MAX / P + 1;
  f.c = (uint32_t)((1ull << 31) % P);
  return f;
}

extern "C" {

void gz_split_options_defaults(gz_split_options* o) {
  if (!o) return;
  o->buffer_size = 4096;
  o->prefer_spill = 1;
  o->offheap_per_task = 0;
  o->tile_rows = 0;
  o->hash_seed = 42;
}

int gz_split_make(gz_context* ctx, const char* short_name, int32_t num_cols,
                  const int32_t* col_types, int32_t num_partitions, const int32_t* key_cols,
                  int32_t num_keys, const gz_split_options* opts, gz_splitter** out) {
  if (!ctx || !short_name || !out || (num_cols > 0 && !col_types))
    return fail(GZ_INVALID, "gz_split_make: null argument");
  std::string name(short_name);
  int mode;
  if (name == "hash") mode = 0;
  else if (name == "rr") mode = 1;
  else if (name == "range") mode = 2;
  else if (name == "single") mode = 3;
  else return fail(GZ_NOT_IMPLEMENTED, "Partitioning " + name + " not supported yet.");
  int P = mode == 3 ? 1 : num_partitions;
  // splitter.cc:218-221: partitions <= 64K, buffer_size <= 32K
  if (P <= 0 || P > 65536) return fail(GZ_INVALID, "num_partitions must be in [1, 65536]");
  if (P > kMaxPartitions)
    return fail(GZ_NOT_IMPLEMENTED, "more than " + std::to_string(kMaxPartitions) +
                                        " partitions is not supported by the HIP splitter");
  if (num_cols > kMaxCols) return fail(GZ_NOT_IMPLEMENTED, "more than 32 columns");
  gz_split_options o;
  gz_split_options_defaults(&o);
  if (opts) o = *opts;
  if (o.buffer_size <= 0 || o.buffer_size > 32 * 1024)
    return fail(GZ_INVALID, "buffer_size must be in [1, 32768]");
  for (int c = 0; c < num_cols; c++)
    if (!type_supported(col_types[c]))
      return fail(GZ_NOT_IMPLEMENTED, "column type " + std::to_string(col_types[c]) +
                                          " is not supported by the splitter");
  if (mode == 2 && (num_cols < 1 || col_types[0] != GZ_TYPE_INT32))
    return fail(GZ_INVALID, "RecordBatch field 0 should be int32");
  if (mode == 0) {
    if (num_keys < 0 || num_keys > kMaxKeys)
      return fail(GZ_NOT_IMPLEMENTED, "at most 4 hash keys");
    for (int k = 0; k < num_keys; k++) {
      int c = key_cols[k];
      if (c < 0 || c >= num_cols) return fail(GZ_INVALID, "key column out of range");
      int t = col_types[c];
      bool ok = t == GZ_TYPE_NA || t == GZ_TYPE_BOOL || t == GZ_TYPE_INT8 || t == GZ_TYPE_INT16 ||
                t == GZ_TYPE_INT32 || t == GZ_TYPE_DATE32 || t == GZ_TYPE_INT64 ||
                t == GZ_TYPE_FLOAT || t == GZ_TYPE_DOUBLE || type_is_binary(t);
      if (!ok)
        return fail(GZ_NOT_IMPLEMENTED, "hash partitioning on type " + std::to_string(t) +
                                            " has no Spark-compatible hash");
    }
  }
  auto* h = new gz_splitter();
  h->ctx = ctx;
  h->mode = mode;
  h->P = P;
  h->ncols_in = num_cols;
  h->types_in.assign(col_types, col_types + num_cols);
  if (mode == 0) h->keys.assign(key_cols, key_cols + num_keys);
  h->opts = o;
  h->tile_rows = o.tile_rows > 0 ? o.tile_rows : 2048;
  if (h->tile_rows != 4096 && h->tile_rows != 2048 && h->tile_rows != 1024) h->tile_rows = 4096;
  SplitParams& b = h->base;
  memset(&b, 0, sizeof(b));
  b.mode = mode;
  b.P = P;
  b.pbits = 1;
  while ((1 << b.pbits) < P) b.pbits++;
  b.fm = make_fastmod((uint32_t)P);
  b.seed = (uint32_t)o.hash_seed;
  b.nkeys = (int)h->keys.size();
  for (int k = 0; k < b.nkeys; k++) {
    b.key_col[k] = h->keys[k];
    b.key_type[k] = col_types[h->keys[k]];
  }
  b.ncols = num_cols;
  int oc = 0, nbin = 0;
  for (int c = 0; c < num_cols; c++) {
    int t = col_types[c];
    b.kind[c] = t == GZ_TYPE_NA ? K_NA : t == GZ_TYPE_BOOL ? K_BOOL : type_is_binary(t) ? K_BINARY : K_FIXED;
    b.width[c] = type_width(t);
    if (mode == 2 && c == 0) {
      b.out_col[c] = -1;
      b.bin[c] = -1;
      continue;
    }
    b.out_col[c] = oc++;
    h->out_types.push_back(t);
    if (type_is_binary(t) && nbin == 0) b.bin_col0 = c;
    b.bin[c] = type_is_binary(t) ? nbin++ : -1;
  }
  b.nbin = nbin;
  if (scatter_lds(h->tile_rows, 256, P) > 160 * 1024) {
    delete h;
    return fail(GZ_NOT_IMPLEMENTED, "partition count too large for the LDS-staged scatter");
  }
  *out = h;
  return GZ_OK;
}

}  // extern "C"

// Run one segment over a list of batches (device buffers in the JNI convention).
static int split_run(gz_splitter* h, int32_t nb, const gz_batch* batches) {
  gz_context* ctx = h->ctx;
  if (h->stopped) return fail(GZ_INVALID, "splitter already stopped");
  GZ_HIP(hipSetDevice(ctx->device));
  const int P = h->P;
  const int nc = h->ncols_in;
  const int TILE = h->tile_rows;
  // ---- validate and build tile / buffer tables
  std::vector<TileDesc> tiles;
  std::vector<const void*> bufs((size_t)nb * nc * 3, nullptr);
  int64_t N = 0;
  std::vector<size_t> data_bytes(h->base.nbin, 0);
  std::vector<bool> has_validity(nc, false);
  std::vector<int32_t> tile_batch;
  std::vector<const int32_t*> empty_data;
  std::vector<int64_t> empty_rows;
  int64_t rr = h->rr_state;
  for (int bi = 0; bi < nb; bi++) {
    const gz_batch& bt = batches[bi];
    if (bt.num_rows < 0 || bt.num_rows > INT32_MAX) return fail(GZ_INVALID, "bad num_rows");
    int need = 0;
    for (int c = 0; c < nc; c++) need += type_num_bufs(h->types_in[c]);
    if (bt.num_bufs < need)
      return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
    int k = 0;
    for (int c = 0; c < nc; c++) {
      int t = h->types_in[c];
      const void** slot = &bufs[((size_t)bi * nc + c) * 3];
      int64_t va = bt.buf_addrs[k], vs = bt.buf_sizes[k];
      slot[0] = (va != 0 && vs > 0) ? (const void*)va : nullptr;
      if (slot[0] && t != GZ_TYPE_NA) {
        if (vs < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
        has_validity[c] = true;
      }
      k++;
      if (type_is_binary(t)) {
        slot[1] = (const void*)bt.buf_addrs[k];
        if (bt.buf_sizes[k] < (bt.num_rows + 1) * 4) return fail(GZ_INVALID, "offsets buffer too small");
        k++;
      }
      slot[2] = (const void*)bt.buf_addrs[k];
      int64_t vsz = bt.buf_sizes[k];
      k++;
      if (t == GZ_TYPE_NA) continue;
      if (type_is_binary(t) && bt.num_rows > 0 && (!slot[2] || vsz == 0)) {
        // an empty data buffer is valid when every value is empty: checked below
        empty_data.push_back((const int32_t*)slot[1]);
        empty_rows.push_back(bt.num_rows);
        slot[2] = slot[1];  // any valid address; no byte is read from it
      }
      if (bt.num_rows > 0 && !slot[2]) return fail(GZ_INVALID, "missing values buffer");
      if (type_is_binary(t)) data_bytes[h->base.bin[c]] += (size_t)vsz;
      else if (t == GZ_TYPE_BOOL) {
        if (vsz < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "values buffer too small");
      } else if (vsz < bt.num_rows * type_width(t)) return fail(GZ_INVALID, "values buffer too small");
    }
    for (int64_t r0 = 0; r0 < bt.num_rows; r0 += TILE) {
      TileDesc td;
      td.batch = bi;
      td.row_begin = (int32_t)r0;
      td.nrows = (int32_t)std::min<int64_t>(TILE, bt.num_rows - r0);
      td.rr = (int32_t)((rr + r0) % P);
      tiles.push_back(td);
      tile_batch.push_back(bi);
    }
    rr = (rr + bt.num_rows) % P;
    N += bt.num_rows;
  }
  if (N > INT32_MAX) return fail(GZ_CAPACITY_ERROR, "more than 2^31 rows in one split call");
  if (!empty_data.empty()) {
    std::vector<int32_t> ends(2 * empty_data.size());
    for (size_t i = 0; i < empty_data.size(); i++) {
      GZ_HIP(hipMemcpyAsync(&ends[2 * i], empty_data[i], 4, hipMemcpyDeviceToHost, ctx->stream));
      GZ_HIP(hipMemcpyAsync(&ends[2 * i + 1], empty_data[i] + empty_rows[i], 4, hipMemcpyDeviceToHost,
                            ctx->stream));
    }
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < empty_data.size(); i++)
      if (ends[2 * i] != ends[2 * i + 1]) return fail(GZ_INVALID, "missing utf8 data buffer");
  }
  h->rr_state = (int32_t)rr;
  auto* seg = new SplitSegment();
  seg->N = N;
  seg->nbatches = nb;
  seg->T = (int64_t)tiles.size();
  seg->tile_batch = tile_batch;
  const int64_t T = seg->T;
  const int nout = (int)h->out_types.size();
  const int nbin = h->base.nbin;
  int rc = GZ_OK;
  auto bail = [&](int code) {
    split_release_segment(ctx, seg);
    return code;
  };
  // ---- allocate the segment
  if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * P * std::max<int64_t>(T, 1), &seg->counts))) return bail(rc);
  if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * P * std::max<int64_t>(T, 1), &seg->prefix))) return bail(rc);
  if (nbin) {
    if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * nbin * P * std::max<int64_t>(T, 1), &seg->bytes))) return bail(rc);
    if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * nbin * P * std::max<int64_t>(T, 1), &seg->bprefix))) return bail(rc);
    if ((rc = ctx_alloc(ctx, sizeof(int64_t) * nbin * (P + 1), &seg->byte_start))) return bail(rc);
  }
  if ((rc = ctx_alloc(ctx, sizeof(uint64_t) * (1 + nbin) * P, &seg->tot))) return bail(rc);
  if ((rc = ctx_alloc(ctx, sizeof(int64_t) * (P + 1), &seg->row_start))) return bail(rc);
  if ((rc = ctx_alloc(ctx, sizeof(unsigned long long) * std::max(nout, 1), &seg->null_counts))) return bail(rc);
  if ((rc = ctx_alloc(ctx, sizeof(int), &seg->err))) return bail(rc);
  seg->values.resize(nout);
  seg->validity.resize(nout);
  seg->offsets.resize(nout);
  seg->data.resize(nout);
  const int64_t nwords = (N + 63) / 64 * 2;  // 32-bit words of a 64-bit padded bitmap
  SplitParams prm = h->base;
  for (int c = 0; c < nc; c++) {
    int oc = prm.out_col[c];
    if (oc < 0) continue;
    int t = h->types_in[c];
    if (t == GZ_TYPE_NA) continue;
    if (has_validity[c]) {
      if ((rc = ctx_alloc(ctx, nwords * 4, &seg->validity[oc]))) return bail(rc);
    }
    if (t == GZ_TYPE_BOOL) {
      if ((rc = ctx_alloc(ctx, nwords * 4, &seg->values[oc]))) return bail(rc);
    } else if (type_is_binary(t)) {
      if ((rc = ctx_alloc(ctx, sizeof(int32_t) * (N + P), &seg->offsets[oc]))) return bail(rc);
    } else {
      if ((rc = ctx_alloc(ctx, (size_t)std::max<int64_t>(N, 1) * type_width(t), &seg->values[oc]))) return bail(rc);
    }
    prm.out_values[oc] = seg->values[oc].p;
    prm.out_valid[oc] = (uint32_t*)seg->validity[oc].p;
    prm.out_off[oc] = (int32_t*)seg->offsets[oc].p;
    prm.out_data[oc] = (uint8_t*)seg->data[oc].p;
  }
  // output columns of an input column are indexed by the output index
  if ((rc = upload(ctx, &h->d_tiles, tiles.data(), sizeof(TileDesc) * tiles.size()))) return bail(rc);
  if ((rc = upload(ctx, &h->d_bufs, bufs.data(), sizeof(void*) * bufs.size()))) return bail(rc);
  prm.tiles = (const TileDesc*)h->d_tiles.p;
  prm.T = T;
  prm.bufs = (const void* const*)h->d_bufs.p;
  prm.bstride = nc * 3;
  prm.counts = (uint32_t*)seg->counts.p;
  prm.prefix = (uint32_t*)seg->prefix.p;
  prm.bytes = (uint32_t*)seg->bytes.p;
  prm.bprefix = (uint32_t*)seg->bprefix.p;
  prm.row_start = (const int64_t*)seg->row_start.p;
  prm.byte_start = (const int64_t*)seg->byte_start.p;
  prm.null_counts = (unsigned long long*)seg->null_counts.p;
  prm.err = (int*)seg->err.p;
  hipStream_t st = ctx->stream;
  // ---- fills (kernels, not memsets, so the sequence can be graph-captured)
  auto fill = [&](void* p, int64_t n32, uint32_t v) {
    if (!p || n32 <= 0) return;
    int blocks = (int)std::min<int64_t>((n32 + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, st, (uint32_t*)p, n32, v);
  };
  fill(seg->null_counts.p, 2 * std::max(nout, 1), 0);
  fill(seg->err.p, 1, 0);
  for (int oc = 0; oc < nout; oc++) {
    if (seg->validity[oc].p) fill(seg->validity[oc].p, nwords, 0xFFFFFFFFu);
    if (h->out_types[oc] == GZ_TYPE_BOOL) fill(seg->values[oc].p, nwords, 0xFFFFFFFFu);
  }
  const int64_t grid = (T + 7) / 8 * 8;
  // hot path: hash of one int64 key that has no validity buffer in this call
  const int keyk = (h->mode == 0 && prm.nkeys == 1 && prm.key_type[0] == GZ_TYPE_INT64 &&
                    !has_validity[prm.key_col[0]]) ? 1 : 0;
  if (T > 0) {
    KernelTimer kt(ctx, "split_count");
    size_t lds = sizeof(uint32_t) * P * (1 + nbin);
    if (keyk)
      hipLaunchKernelGGL((split_count_kernel<256, 1>), dim3((unsigned)grid), dim3(256), lds, st, prm);
    else
      hipLaunchKernelGGL((split_count_kernel<256, 0>), dim3((unsigned)grid), dim3(256), lds, st, prm);
  }
  {
    KernelTimer kt(ctx, "split_scan");
    if (T > 0) {
      hipLaunchKernelGGL(split_scan_kernel<256>, dim3((unsigned)P), dim3(256), 0, st,
                         (const uint32_t*)seg->counts.p, (uint32_t*)seg->prefix.p,
                         (uint64_t*)seg->tot.p, T);
      if (nbin)
        hipLaunchKernelGGL(split_scan_kernel<256>, dim3((unsigned)(P * nbin)), dim3(256), 0, st,
                           (const uint32_t*)seg->bytes.p, (uint32_t*)seg->bprefix.p,
                           (uint64_t*)seg->tot.p + P, T);
    } else {
      fill(seg->tot.p, 2 * (int64_t)(1 + nbin) * P, 0);
    }
    std::vector<int32_t*> offp(std::max(nbin, 1), nullptr);
    for (int c = 0; c < nc; c++)
      if (prm.bin[c] >= 0) offp[prm.bin[c]] = prm.out_off[prm.out_col[c]];
    if ((rc = upload(ctx, &seg->offtab, offp.data(), sizeof(int32_t*) * offp.size()))) return bail(rc);
    hipLaunchKernelGGL(split_layout_kernel, dim3(1), dim3(256), 0, st, (const uint64_t*)seg->tot.p,
                       P, nbin, (int64_t*)seg->row_start.p, (int64_t*)seg->byte_start.p,
                       (int32_t**)seg->offtab.p, (int*)seg->err.p);
  }
  // utf8 data outputs: the sum of the input data buffers bounds them; when that
  // bound is loose (buffers shared by many batches), size them exactly from the
  // scanned byte totals instead.
  for (int c = 0; c < nc; c++) {
    int oc = prm.out_col[c];
    if (oc < 0 || !type_is_binary(h->types_in[c])) continue;
    size_t bytes = data_bytes[prm.bin[c]];
    if (bytes > ((size_t)1 << 30)) {
      int64_t total = 0;
      GZ_HIP(hipMemcpyAsync(&total, (int64_t*)seg->byte_start.p + (int64_t)prm.bin[c] * (P + 1) + P,
                            sizeof(int64_t), hipMemcpyDeviceToHost, st));
      GZ_HIP(hipStreamSynchronize(st));
      bytes = (size_t)total;
    }
    if ((rc = ctx_alloc(ctx, std::max<size_t>(bytes, 4), &seg->data[oc]))) return bail(rc);
    prm.out_data[oc] = (uint8_t*)seg->data[oc].p;
  }
  if (T > 0) {
    KernelTimer kt(ctx, "split_scatter");
    size_t lds = scatter_lds(TILE, 256, P);
#define GZ_SCATTER(TL, KK) \
  hipLaunchKernelGGL((split_scatter_kernel<TL, 256, KK>), dim3((unsigned)grid), dim3(256), lds, st, prm)
    if (TILE == 4096) { if (keyk) GZ_SCATTER(4096, 1); else GZ_SCATTER(4096, 0); }
    else if (TILE == 2048) { if (keyk) GZ_SCATTER(2048, 1); else GZ_SCATTER(2048, 0); }
    else { if (keyk) GZ_SCATTER(1024, 1); else GZ_SCATTER(1024, 0); }
#undef GZ_SCATTER
  }
  GZ_HIP(hipGetLastError());
  h->segments.push_back(seg);
  if (h->mode == 2) {
    // the range splitter validates partition ids like the reference (splitter.cc:1511-1516)
    int err = 0;
    GZ_HIP(hipMemcpyAsync(&err, seg->err.p, sizeof(int), hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (err & 1) return fail(GZ_INVALID, "Partition id is equal or greater than num_partitions");
  }
  return GZ_OK;
}

static int resolve_segment(gz_splitter* h, SplitSegment* s) {
  if (s->resolved) return GZ_OK;
  gz_context* ctx = h->ctx;
  const int P = h->P, nbin = h->base.nbin, nout = (int)h->out_types.size();
  s->h_row_start.resize(P + 1);
  s->h_byte_start.resize((size_t)nbin * (P + 1));
  s->h_null.resize(std::max(nout, 1));
  int err = 0;
  GZ_HIP(hipMemcpyAsync(s->h_row_start.data(), s->row_start.p, sizeof(int64_t) * (P + 1),
                        hipMemcpyDeviceToHost, ctx->stream));
  if (nbin)
    GZ_HIP(hipMemcpyAsync(s->h_byte_start.data(), s->byte_start.p,
                          sizeof(int64_t) * nbin * (P + 1), hipMemcpyDeviceToHost, ctx->stream));
  GZ_HIP(hipMemcpyAsync(s->h_null.data(), s->null_counts.p, sizeof(unsigned long long) * std::max(nout, 1),
                        hipMemcpyDeviceToHost, ctx->stream));
  GZ_HIP(hipMemcpyAsync(&err, s->err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  if (err & 1) return fail(GZ_INVALID, "Partition id is equal or greater than num_partitions");
  if (err & 2) return fail(GZ_CAPACITY_ERROR, "a partition's utf8 data exceeds 2 GiB");
  s->resolved = true;
  return GZ_OK;
}

extern "C" {

int gz_split_batch(gz_splitter* h, int64_t num_rows, const int64_t* buf_addrs,
                   const int64_t* buf_sizes, int32_t num_bufs) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (!buf_addrs || !buf_sizes) return fail(GZ_INVALID, "Native split: buf_addrs can't be null");
  gz_batch b{num_rows, buf_addrs, buf_sizes, num_bufs};
  return split_run(h, 1, &b);
}

int gz_split_batches(gz_splitter* h, int32_t num_batches, const gz_batch* batches) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (num_batches < 0 || (num_batches > 0 && !batches)) return fail(GZ_INVALID, "bad batch list");
  for (int i = 0; i < num_batches; i++)
    if (!batches[i].buf_addrs || !batches[i].buf_sizes)
      return fail(GZ_INVALID, "Native split: buf_addrs can't be null");
  return split_run(h, num_batches, batches);
}

int gz_split_stop(gz_splitter* h) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  for (auto* s : h->segments) GZ_TRY(resolve_segment(h, s));
  if (h->ctx->profiling) {
    double ms = 0;
    gz_context_kernel_time(h->ctx, "split_count", &ms, nullptr);
    h->compute_pid_ns = (int64_t)(ms * 1e6);
  }
  h->stopped = true;
  return GZ_OK;
}

int gz_split_num_partitions(gz_splitter* h) { return h ? h->P : 0; }
int gz_split_num_output_columns(gz_splitter* h) { return h ? (int)h->out_types.size() : 0; }
int gz_split_num_segments(gz_splitter* h) { return h ? (int)h->segments.size() : 0; }

int gz_split_partition_rows(gz_splitter* h, int64_t* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  for (int p = 0; p < h->P; p++) out[p] = 0;
  for (auto* s : h->segments) {
    GZ_TRY(resolve_segment(h, s));
    for (int p = 0; p < h->P; p++) out[p] += s->h_row_start[p + 1] - s->h_row_start[p];
  }
  return GZ_OK;
}

int gz_split_segment_rows(gz_splitter* h, int32_t seg, int64_t* row_start) {
  if (!h || !row_start) return fail(GZ_INVALID, "null argument");
  if (seg < 0 || seg >= (int)h->segments.size()) return fail(GZ_INDEX_ERROR, "segment out of range");
  SplitSegment* s = h->segments[seg];
  GZ_TRY(resolve_segment(h, s));
  memcpy(row_start, s->h_row_start.data(), sizeof(int64_t) * (h->P + 1));
  return GZ_OK;
}

int gz_split_segment_column(gz_splitter* h, int32_t seg, int32_t col, void** values,
                            void** validity, void** offsets, void** data, int64_t* byte_start,
                            int64_t* null_count) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (seg < 0 || seg >= (int)h->segments.size()) return fail(GZ_INDEX_ERROR, "segment out of range");
  if (col < 0 || col >= (int)h->out_types.size()) return fail(GZ_INDEX_ERROR, "column out of range");
  SplitSegment* s = h->segments[seg];
  GZ_TRY(resolve_segment(h, s));
  if (values) *values = s->values[col].p;
  if (validity) *validity = s->validity[col].p;
  if (offsets) *offsets = s->offsets[col].p;
  if (data) *data = s->data[col].p;
  if (null_count) *null_count = (int64_t)s->h_null[col];
  if (byte_start) {
    int in_col = -1;
    for (int c = 0; c < h->ncols_in; c++)
      if (h->base.out_col[c] == col) in_col = c;
    int j = in_col >= 0 ? h->base.bin[in_col] : -1;
    if (j >= 0) memcpy(byte_start, &s->h_byte_start[(size_t)j * (h->P + 1)], sizeof(int64_t) * (h->P + 1));
  }
  return GZ_OK;
}

int gz_split_segment_batch_counts(gz_splitter* h, int32_t seg, int32_t* num_batches,
                                  int32_t* counts) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (seg < 0 || seg >= (int)h->segments.size()) return fail(GZ_INDEX_ERROR, "segment out of range");
  SplitSegment* s = h->segments[seg];
  if (num_batches) *num_batches = s->nbatches;
  if (!counts) return GZ_OK;
  const int P = h->P;
  std::vector<uint32_t> tc((size_t)P * std::max<int64_t>(s->T, 1));
  if (s->T > 0) {
    GZ_HIP(hipMemcpyAsync(tc.data(), s->counts.p, sizeof(uint32_t) * P * s->T,
                          hipMemcpyDeviceToHost, h->ctx->stream));
    GZ_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  memset(counts, 0, sizeof(int32_t) * (size_t)s->nbatches * P);
  for (int64_t t = 0; t < s->T; t++) {
    int b = s->tile_batch[t];
    for (int p = 0; p < P; p++) counts[(size_t)b * P + p] += (int32_t)tc[(size_t)p * s->T + t];
  }
  return GZ_OK;
}

int64_t gz_split_compute_pid_time(gz_splitter* h) { return h ? h->compute_pid_ns : 0; }

int gz_split_release(gz_splitter* h) {
  if (!h) return GZ_OK;
  hipSetDevice(h->ctx->device);
  // No synchronization: the pool hands a freed block only to work enqueued
  // later on the context's stream, which runs after everything queued here.
  for (auto* s : h->segments) split_release_segment(h->ctx, s);
  ctx_free(h->ctx, &h->d_tiles);
  ctx_free(h->ctx, &h->d_bufs);
  delete h;
  return GZ_OK;
}

}  // extern "C"
