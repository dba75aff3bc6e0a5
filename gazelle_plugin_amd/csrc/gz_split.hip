// Shuffle hash Splitter for MI355X (gfx950).
//
// Replaces sparkcolumnarplugin::shuffle::Splitter (cpp/src/shuffle/splitter.cc).
// One call (one batch, or many batches at once) produces one *segment*: the rows
// of every partition contiguous and in input order, which is the order the
// reference's DoSplit appends them in (splitter.cc:855-938).
//
// Launches per segment:
//   split_count    one workgroup per tile: partition id of every row (Spark
//                  Murmur3 hashLong seed 42 + pmod, or rr / range / single), an
//                  LDS histogram of rows and of utf8 bytes per partition, and the
//                  null count of every nullable column.  Reads the key (and utf8
//                  offsets, validity) only.  Writes one row of R counters per
//                  tile (tile-major, so the row is one contiguous store).
//   split_scan_*   exclusive scan over tiles of every counter (three small
//                  launches: chunk sums, scan of chunk sums, per-tile prefixes).
//   split_layout   partition starts (rows and utf8 bytes) from the totals.
//   split_scatter  one 1024-thread workgroup per tile: a stable in-tile rank per
//                  partition from wave ballots, kept in registers; each column is
//                  loaded with independent coalesced loads, permuted through LDS
//                  into partition order and written as contiguous runs; utf8
//                  strings move with aligned dword loads; validity and boolean
//                  bits are cleared with atomic AND in bitmaps pre-filled with ones.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "gz_device.h"
#include "gz_internal.h"
#include "gz_scan.h"

namespace gz {

constexpr int kMaxCols = 128;  // input columns (and output columns)
constexpr int kArgCols = 32;   // columns whose tables ride in the kernel arguments
constexpr int kMaxKeys = 16;
constexpr int kMaxPartitions = 2048;
constexpr int kScanChunk = 16;  // tiles per chunk of the tile scan
constexpr size_t kCountLdsMax = 128 * 1024;  // LDS bytes of the count kernel's per-tile counters
constexpr int kPre = 4;         // fixed-width columns the scatter loads up front
// ... fewer when a thread owns more rows (registers)
__host__ __device__ constexpr int scatter_max_pre(int rows_per_thread) {
  return rows_per_thread <= 2 ? kPre : rows_per_thread <= 4 ? 2 : 1;
}

enum ColKind : int { K_NA = 0, K_FIXED = 1, K_BOOL = 2, K_BINARY = 3 };

struct TileDesc {
  int32_t batch;
  int32_t row_begin;
  int32_t nrows;
  int32_t rr;        // round-robin partition of the tile's first row
  int32_t call_row;  // index of the tile's first row among the rows of the call
  int32_t pad;
};

// A column of the split: constant per splitter except nullc and is_pre, which
// depend on the call's buffers.
struct SplitCol {
  int kind, width;
  int out_col;  // output column of an input column, -1 if dropped
  int bin;      // binary column ordinal, -1 otherwise
  int nullc;    // null-count slot of a nullable column, -1 otherwise
  int is_pre;   // loaded up front by the scatter
};

// Buffers of an output column in a segment.
struct SplitOut {
  void* values;
  uint32_t* valid;
  int32_t* off;
  uint8_t* data;
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
  SplitCol acol[kArgCols];   // the first kArgCols columns
  const SplitCol* col;       // [ncols] (device): the columns from kArgCols on
  int npre;                  // fixed-width columns loaded up front by the scatter
  int pre_col[kPre];         // (the key column first on the int64-key path)
  int nbin;
  int bin_col0;              // input column of the first binary column
  // per-segment buffers.  A tile's counter row has R = P * (1 + nbin) + nnull
  // entries: rows per partition, utf8 bytes per (binary column, partition),
  // nulls per nullable column.
  int R;
  uint32_t* counts;          // [T][R]
  uint32_t* prefix;          // [T][R] exclusive over tiles
  const int64_t* row_start;  // [P+1]
  const int64_t* byte_start; // [nbin][P+1]
  SplitOut aout[kArgCols];   // the first kArgCols output columns
  const SplitOut* out;       // [output columns] (device): the columns from kArgCols on
  int* err;
  int cbuf;                  // LDS bytes for partition-ordered utf8 staging
  int rowid_oc;              // output column of the row indices, -1 if none
};

// Column and output tables: schemas of at most kArgCols columns read them
// from the kernel arguments (WIDE = 0, no memory load), wider ones from the
// device tables (WIDE = 1).
template <int WIDE>
__device__ __forceinline__ const SplitCol& col_of(const SplitParams& p, int c) {
  return WIDE ? p.col[c] : p.acol[c];
}
template <int WIDE>
__device__ __forceinline__ const SplitOut& out_of(const SplitParams& p, int oc) {
  return WIDE ? p.out[oc] : p.aout[oc];
}

static_assert(sizeof(SplitParams) <= 4096, "kernel arguments");

__device__ __forceinline__ void load_tile_ptrs(const SplitParams& p, int batch, const void** s_bp) {
  for (int i = threadIdx.x; i < p.ncols * 3; i += blockDim.x)
    s_bp[i] = gld(&p.bufs[(int64_t)batch * p.bstride + i]);
}

// Spark hash of key column c at row r, chained through h (Murmur3Hash semantics:
// a null value leaves the running hash unchanged).  A float / double key hashes
// its bits (hashInt(floatToIntBits) / hashLong(doubleToLongBits)); -0.0 and NaN,
// whose Spark hash differs between versions and is not fixed by anything the
// reference holds, set bit 4 of *err, and the split then fails at Stop.
__device__ __forceinline__ uint32_t hash_key(const void* const* bp, int c, int t, int64_t r, uint32_t h, int* err) {
  const uint8_t* vb = (const uint8_t*)bp[c * 3 + 0];
  if (vb && !bit_get(vb, r)) return h;
  const void* v = bp[c * 3 + 2];
  switch (t) {
    case GZ_TYPE_INT64: return spark_hash_long((uint64_t)gld((const int64_t*)v + r), h);
    case GZ_TYPE_INT32:
    case GZ_TYPE_DATE32: return spark_hash_int((uint32_t)gld((const int32_t*)v + r), h);
    case GZ_TYPE_BOOL: return spark_hash_int(bit_get((const uint8_t*)v, r), h);
    case GZ_TYPE_INT8: return spark_hash_int((uint32_t)(int32_t)gld((const int8_t*)v + r), h);
    case GZ_TYPE_INT16: return spark_hash_int((uint32_t)(int32_t)gld((const int16_t*)v + r), h);
    case GZ_TYPE_FLOAT: {
      const uint32_t bits = (uint32_t)gld((const int32_t*)v + r);
      if (bits == 0x80000000u || (bits & 0x7fffffffu) > 0x7f800000u) atomicOr(err, 4);
      return spark_hash_int(bits, h);
    }
    case GZ_TYPE_DOUBLE: {
      const uint64_t bits = (uint64_t)gld((const int64_t*)v + r);
      if (bits == 0x8000000000000000ull || (bits & 0x7fffffffffffffffull) > 0x7ff0000000000000ull) atomicOr(err, 4);
      return spark_hash_long(bits, h);
    }
    case GZ_TYPE_STRING:
    case GZ_TYPE_BINARY: {
      const int32_t* off = (const int32_t*)bp[c * 3 + 1];
      int32_t o0 = gld(off + r), o1 = gld(off + r + 1);
      const uint8_t* d = (const uint8_t*)v + o0;
      int len = o1 - o0;
      int aligned = len & ~3;
      for (int i = 0; i < aligned; i += 4) {
        uint32_t w = (uint32_t)gld(d + i) | ((uint32_t)gld(d + i + 1) << 8) |
                     ((uint32_t)gld(d + i + 2) << 16) | ((uint32_t)gld(d + i + 3) << 24);
        h = mix_h1(h, mix_k1(w));
      }
      for (int i = aligned; i < len; i++) h = mix_h1(h, mix_k1((uint32_t)(int32_t)(int8_t)gld(d + i)));
      return fmix32(h, (uint32_t)len);
    }
    default: return h;
  }
}

// Partition id of row lr of a tile (generic path: any keys, any mode).
__device__ __forceinline__ int compute_pid(const SplitParams& p, const void* const* bp, const TileDesc& td,
                                           int lr) {
  int64_t r = (int64_t)td.row_begin + lr;
  if (p.mode == 0) {
    uint32_t h = p.seed;
    for (int k = 0; k < p.nkeys; k++) h = hash_key(bp, p.key_col[k], p.key_type[k], r, h, p.err);
    return pmod_hash(h, p.fm);
  }
  if (p.mode == 1) return (int)fastmod_u32((uint32_t)(td.rr + lr), p.fm);
  if (p.mode == 2) {
    int v = gld((const int32_t*)bp[2] + r);
    if (v < 0 || v >= p.P) {
      atomicOr(p.err, 1);
      return 0;
    }
    return v;
  }
  return 0;
}

// ---------------------------------------------------------------- pass 1
// KEYK = 1: one int64 key without validity (the hot case), else generic.
template <int BLOCK, int KEYK, int WIDE>
__global__ __launch_bounds__(BLOCK) void split_count_kernel(SplitParams p) {
  extern __shared__ uint32_t s_cnt[];  // [R]
  __shared__ const void* s_bp[kMaxCols * 3];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const TileDesc td = p.tiles[t];
  const int P = p.P;
  load_tile_ptrs(p, td.batch, s_bp);
  for (int i = threadIdx.x; i < p.R; i += BLOCK) s_cnt[i] = 0;
  __syncthreads();
  if (KEYK == 1) {
    // four independent rows per step
    const int64_t* key = (const int64_t*)s_bp[p.key_col[0] * 3 + 2] + td.row_begin;
    const int32_t* off = p.nbin == 1 ? (const int32_t*)s_bp[p.bin_col0 * 3 + 1] + td.row_begin : nullptr;
    for (int lr0 = threadIdx.x; lr0 < td.nrows; lr0 += 4 * BLOCK) {
      uint64_t k[4];
      int32_t o0[4], o1[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        int lr = lr0 + u * BLOCK;
        bool a = lr < td.nrows;
        k[u] = a ? (uint64_t)gld(key + lr) : 0;
        o0[u] = a && off ? gld(off + lr) : 0;
        o1[u] = a && off ? gld(off + lr + 1) : 0;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        int lr = lr0 + u * BLOCK;
        if (lr < td.nrows) {
          int pid = pmod_hash(spark_hash_long(k[u], p.seed), p.fm);
          atomicAdd(&s_cnt[pid], 1u);
          if (off) atomicAdd(&s_cnt[P + pid], (uint32_t)(o1[u] - o0[u]));
          else if (p.nbin) {
            int64_t r = (int64_t)td.row_begin + lr;
            for (int c = 0; c < p.ncols; c++) {
              int j = col_of<WIDE>(p, c).bin;
              if (j < 0) continue;
              const int32_t* of = (const int32_t*)s_bp[c * 3 + 1];
              atomicAdd(&s_cnt[P * (1 + j) + pid], (uint32_t)(gld(of + r + 1) - gld(of + r)));
            }
          }
        }
      }
    }
  } else {
    for (int lr = threadIdx.x; lr < td.nrows; lr += BLOCK) {
      int pid = compute_pid(p, s_bp, td, lr);
      atomicAdd(&s_cnt[pid], 1u);
      if (p.nbin) {
        int64_t r = (int64_t)td.row_begin + lr;
        for (int c = 0; c < p.ncols; c++) {
          int j = col_of<WIDE>(p, c).bin;
          if (j < 0) continue;
          const int32_t* off = (const int32_t*)s_bp[c * 3 + 1];
          atomicAdd(&s_cnt[P * (1 + j) + pid], (uint32_t)(gld(off + r + 1) - gld(off + r)));
        }
      }
    }
  }
  // null counts: zero bits of each validity bitmap over the tile's rows
  // (tiles start on multiples of 8 rows inside a batch).
  for (int c = 0; c < p.ncols; c++) {
    int slot = col_of<WIDE>(p, c).nullc;
    if (slot < 0) continue;
    const uint8_t* vb = (const uint8_t*)s_bp[c * 3 + 0];
    if (!vb) continue;
    uint32_t nulls = 0;
    int nbytes = (td.nrows + 7) >> 3;
    const uint8_t* base = vb + (td.row_begin >> 3);
    for (int i = threadIdx.x; i < nbytes; i += BLOCK) {
      uint32_t b = gld(base + i);
      int valid_bits = min(8, td.nrows - i * 8);
      uint32_t mask = (1u << valid_bits) - 1u;
      nulls += __popc((~b) & mask);
    }
    nulls = wave_reduce_sum(nulls);
    if (__lane_id() == 0 && nulls) atomicAdd(&s_cnt[slot], nulls);
  }
  __syncthreads();
  uint32_t* row = p.counts + t * p.R;
  for (int i = threadIdx.x; i < p.R; i += BLOCK) gst(row + i, s_cnt[i]);
}

// ---------------------------------------------------------------- scan
// (tile_scan_* in gz_scan.h)

// One block: exclusive scans of the partition totals into row_start / byte_start,
// first utf8 offset of every partition run set to 0, error on a partition whose
// bytes exceed int32 offsets, and error 8 when a binary column's bytes exceed
// the capacity of its data output (offsets past the declared data buffers).
// tab: [nbin] offsets outputs, then [nbin] data capacities.
__global__ __launch_bounds__(256) void split_layout_kernel(const uint64_t* tot, int P, int nbin,
                                                            int64_t* row_start,
                                                            int64_t* byte_start,
                                                            const int64_t* tab, int* err) {
  int32_t* const* out_off = (int32_t* const*)tab;
  const int64_t* caps = tab + nbin;
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
    if (threadIdx.x == 0) {
      dst[P] = (int64_t)carry;
      if (j >= 0 && (int64_t)carry > caps[j]) atomicOr(err, 8);
    }
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
__device__ __forceinline__ uint32_t funnel(uint32_t lo, uint32_t hi, uint32_t sh) {
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> (8 * sh));
}

// Copy one string of len bytes HBM to HBM (the path for tiles whose utf8 bytes
// do not fit the LDS buffer): byte stores up to the first 4-byte boundary of
// dst, dwords assembled from byte loads, then the tail, so exactly
// [dst, dst+len) is written.
__device__ __forceinline__ void copy_string(uint8_t* dst, const uint8_t* src, int len) {
  if (len <= 0) return;
  const int hb = min(len, (int)((4 - ((uintptr_t)dst & 3)) & 3));
  int i = 0;
  for (; i < hb; i++) gst(dst + i, gld(src + i));
  for (; i + 4 <= len; i += 4) {
    uint32_t x = (uint32_t)gld(src + i) | ((uint32_t)gld(src + i + 1) << 8) |
                 ((uint32_t)gld(src + i + 2) << 16) | ((uint32_t)gld(src + i + 3) << 24);
    gst((uint32_t*)(dst + i), x);
  }
  for (; i < len; i++) gst(dst + i, gld(src + i));
}

// Copy one string of len bytes HBM to HBM with dword loads: byte stores up to
// the first 4-byte boundary of dst, then dwords, each put together from the two
// aligned source dwords it straddles, then the tail bytes.  Only aligned dwords
// that hold at least one of the string's bytes are read.
__device__ __forceinline__ void copy_string_dw(uint8_t* dst, const uint8_t* src, int len) {
  if (len <= 0) return;
  const int head = min(len, (int)((4 - ((uintptr_t)dst & 3)) & 3));
  for (int q = 0; q < head; q++) gst(dst + q, gld(src + q));
  const int body = len - head;
  const int nd = body >> 2;
  const uint8_t* sb = src + head;
  const uint32_t sh = (uint32_t)((uintptr_t)sb & 3);
  const uint32_t* s32 = (const uint32_t*)(sb - sh);
  uint32_t* d32 = (uint32_t*)(dst + head);
  if (sh == 0) {
    for (int m = 0; m < nd; m++) gst(d32 + m, gld(s32 + m));
  } else if (nd > 0) {
    // dwords 0 .. nd of s32 cover bytes sb .. sb + 4 nd - 1
    uint32_t lo = gld(s32);
    for (int m = 0; m < nd; m++) {
      const uint32_t hi = gld(s32 + m + 1);
      gst(d32 + m, funnel(lo, hi, sh));
      lo = hi;
    }
  }
  const int tb = body & 3;
  for (int q = 0; q < tb; q++) gst((uint8_t*)(d32 + nd) + q, gld(sb + 4 * nd + q));
}

// Bytes [b0, b0 + len) of an LDS buffer to dst in HBM: byte stores up to the
// first 4-byte boundary of dst, dword stores (each assembled from two aligned LDS
// dwords), then the tail bytes.  The buffer holds 8 readable bytes past its end.
__device__ __forceinline__ void lds_to_hbm(uint8_t* dst, const uint8_t* buf, uint32_t b0, int len) {
  if (len <= 0) return;
  const int head = min(len, (int)((4 - ((uintptr_t)dst & 3)) & 3));
  for (int q = 0; q < head; q++) gst(dst + q, buf[b0 + q]);
  const int body = len - head;
  const int nd = body >> 2;
  const uint32_t sb = b0 + (uint32_t)head;
  const uint32_t sh = sb & 3;
  const uint32_t* b32 = (const uint32_t*)buf + (sb >> 2);
  uint32_t* d32 = (uint32_t*)(dst + head);
  uint32_t lo = b32[0];
  for (int m = 0; m < nd; m++) {
    const uint32_t hi = b32[m + 1];
    gst(d32 + m, funnel(lo, hi, sh));
    lo = hi;
  }
  const int tb = body & 3;
  uint8_t* dt = (uint8_t*)(d32 + nd);
  for (int q = 0; q < tb; q++) gst(dt + q, buf[sb + 4 * nd + q]);
}

// LDS layout of the scatter kernel (bytes from the start of dynamic LDS).
struct ScatterLds {
  size_t scan, pbase, bbase, brel, toff, wcnt, spid, pos, stage, cbuf, total;
};
// cbuf: bytes of the partition-ordered utf8 buffer (0 without binary columns).
__host__ __device__ inline ScatterLds scatter_layout(int TILE, int BLOCK, int P, int cbuf) {
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
  L.pbase = take(4 * (size_t)P, 4);
  L.bbase = take(8 * (size_t)P, 8);
  L.brel = take(4 * (size_t)P, 4);
  L.toff = take(4 * (size_t)(P + 1), 4);
  L.wcnt = take(2 * (size_t)NW * P, 4);
  L.spid = take(2 * (size_t)TILE, 4);
  L.pos = take(2 * (size_t)TILE, 4);
  L.stage = take((size_t)TILE * 8 + 8, 16);
  L.cbuf = take(cbuf > 0 ? (size_t)cbuf + 8 : 0, 16);
  L.total = o;
  return L;
}

// Rows of a tile are owned in the loading, ranking and validity steps by
// thread (wave w, lane l) as rows w*RW + j*64 + l, j < RPT (coalesced per wave);
// sorted slots are owned in the writing steps as j*BLOCK + tid.
//
// Every load that does not depend on the ranking is issued first: the values of
// up to kPre fixed-width columns and their validity, the offsets of the first
// utf8 column and the tile's first-row table, so one HBM round trip overlaps the
// hashing and ranking instead of one round trip per column.  The tile's utf8
// bytes are copied into LDS as one contiguous block (coalesced 16-byte loads);
// each sorted slot then writes its string from LDS.
template <int TILE, int BLOCK, int KEYK, int WIDE>
__global__ __launch_bounds__(BLOCK, 2048 / BLOCK * BLOCK / 256) void split_scatter_kernel(SplitParams p) {
  constexpr int NW = BLOCK / kWave;
  constexpr int RPT = TILE / BLOCK;  // rows per thread
  constexpr int RW = TILE / NW;      // rows per wave = RPT * 64
  static_assert(RW == RPT * kWave, "tile must be a multiple of the block");
  extern __shared__ __align__(16) uint8_t smem[];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const TileDesc td = p.tiles[t];
  const int n = td.nrows;
  const int P = p.P;
  const int tid = threadIdx.x, w = tid / kWave, lane = __lane_id();
  const ScatterLds Lo = scatter_layout(TILE, BLOCK, P, p.cbuf);
  uint64_t* s_scan = (uint64_t*)(smem + Lo.scan);
  uint32_t* s_pbase = (uint32_t*)(smem + Lo.pbase);
  int64_t* s_bbase = (int64_t*)(smem + Lo.bbase);
  uint32_t* s_brel = (uint32_t*)(smem + Lo.brel);
  uint32_t* s_toff = (uint32_t*)(smem + Lo.toff);
  uint16_t* s_wcnt = (uint16_t*)(smem + Lo.wcnt);
  uint16_t* s_spid = (uint16_t*)(smem + Lo.spid);
  uint16_t* s_pos = (uint16_t*)(smem + Lo.pos);  // sorted slot of every tile row
  uint8_t* s_stage = smem + Lo.stage;
  uint8_t* s_cbuf = smem + Lo.cbuf;
  const uint32_t* trow = p.prefix + t * p.R;  // this tile's exclusive prefixes
  // the batch's buffer table (uniform: scalar loads)
  const void* const* bp = p.bufs + (int64_t)td.batch * p.bstride;
  auto colp = [&](int c, int k) -> const uint8_t* { return (const uint8_t*)bp[c * 3 + k]; };
  const int64_t rb = td.row_begin;
  auto lrow = [&](int j) { return w * RW + j * kWave + lane; };
  // binary columns' bytes past their data outputs' capacity (set by the layout
  // kernel, earlier in the stream): the split fails and no string is written
  const bool bytes_bad = p.nbin && (gld(p.err) & 8);

  // ---- 0: loads that do not depend on the ranking
  for (int i = tid; i < NW * P; i += BLOCK) s_wcnt[i] = 0;
  for (int pp = tid; pp < P; pp += BLOCK) s_pbase[pp] = (uint32_t)(gld(p.row_start + pp) + gld(trow + pp));
  constexpr int NPRE = scatter_max_pre(RPT);
  uint64_t pv[NPRE][RPT];
  uint32_t nullm[RPT];  // bit k: the row is null in prefetched column k
#pragma unroll
  for (int j = 0; j < RPT; j++) nullm[j] = 0;
#pragma unroll
  for (int k = 0; k < NPRE; k++) {
#pragma unroll
    for (int j = 0; j < RPT; j++) pv[k][j] = 0;
    if (k < p.npre) {
      const int c = p.pre_col[k];
      const uint8_t* v = colp(c, 2);
      const uint8_t* vbm = colp(c, 0);
      const int W = col_of<WIDE>(p, c).width;
#pragma unroll
      for (int j = 0; j < RPT; j++) {
        if (lrow(j) < n) {
          const int64_t r = rb + lrow(j);
          if (W == 8) pv[k][j] = gld((const uint64_t*)v + r);
          else if (W == 4) pv[k][j] = gld((const uint32_t*)v + r);
          else if (W == 2) pv[k][j] = gld((const uint16_t*)v + r);
          else pv[k][j] = gld(v + r);
          if (vbm && !bit_get(vbm, r)) nullm[j] |= 1u << k;
        }
      }
    }
  }
  int32_t o0[RPT], o1[RPT];
  const int bc = p.bin_col0;
  if (p.nbin) {
    const int32_t* off = (const int32_t*)colp(bc, 1) + rb;
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const bool a = lrow(j) < n;
      o0[j] = a ? gld(off + lrow(j)) : 0;
      o1[j] = a ? gld(off + lrow(j) + 1) : 0;
    }
  }
  __syncthreads();

  // ---- 1: partition id and stable rank within the wave's rows
  int pid[RPT], pos[RPT];
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    if (KEYK == 1) pid[j] = pmod_hash(spark_hash_long(pv[0][j], p.seed), p.fm);
    else pid[j] = lrow(j) < n ? compute_pid(p, bp, td, lrow(j)) : 0;
  }
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    bool act = lrow(j) < n;
    uint64_t m = match_bits((uint32_t)pid[j], act, p.pbits);
    uint64_t lt = m & lanemask_lt();
    int before = act ? s_wcnt[w * P + pid[j]] : 0;
    pos[j] = before + __popcll(lt);
    if (act && lt == 0) s_wcnt[w * P + pid[j]] = (uint16_t)(before + __popcll(m));
  }
  // utf8 bytes of the first binary column: one contiguous block in LDS, from
  // the 16-byte boundary at or before the tile's first byte
  const uint8_t* cbase = nullptr;
  int64_t clen = 0;
  uint32_t cshift = 0;
  bool chars_lds = false;
  if (p.nbin) {
    const int32_t* off = (const int32_t*)colp(bc, 1);
    const int64_t cb0 = gld(off + rb), cbn = gld(off + rb + n);
    const uint8_t* first = colp(bc, 2) + cb0;
    cbase = (const uint8_t*)((uintptr_t)first & ~(uintptr_t)15);
    cshift = (uint32_t)(first - cbase) - (uint32_t)cb0;  // LDS offset of a source offset o: o + cshift
    clen = (int64_t)(first - cbase) + (cbn - cb0);
    chars_lds = clen + 16 <= (int64_t)p.cbuf;
    if (chars_lds)
      for (int64_t k = tid; k * 16 < clen; k += BLOCK) *(u32x4*)(s_cbuf + k * 16) = gld((const u32x4*)cbase + k);
  }
  __syncthreads();

  // ---- 2: per partition: exclusive scan of the wave counts (in place; a
  // segment of NW lanes per partition, lane k holding wave k) and the tile count
  {
    constexpr int GPB = BLOCK / NW;  // partitions per pass
    const int k = tid % NW, g = tid / NW;
    for (int base = 0; base < P; base += GPB) {
      const int pp = base + g;
      const uint32_t x = pp < P ? s_wcnt[k * P + pp] : 0u;
      uint32_t inc = x;
#pragma unroll
      for (int o = 1; o < NW; o <<= 1) {
        uint32_t y = __shfl_up(inc, o, NW);
        if (k >= o) inc += y;
      }
      if (pp < P) {
        s_wcnt[k * P + pp] = (uint16_t)(inc - x);
        if (k == NW - 1) s_toff[pp] = inc;  // tile count for now
      }
    }
  }
  __syncthreads();
  // tile offsets: one wave scans the partition counts (lane l owns a contiguous
  // range of partitions)
  if (w == 0) {
    const int per = (P + kWave - 1) / kWave;
    const int p0 = min(P, lane * per), p1 = min(P, p0 + per);
    uint32_t s = 0;
    for (int pp = p0; pp < p1; pp++) s += s_toff[pp];
    const uint32_t inc = wave_inclusive_scan(s);
    uint32_t run = inc - s;
    for (int pp = p0; pp < p1; pp++) {
      const uint32_t c = s_toff[pp];
      s_toff[pp] = run;
      run += c;
    }
    if (lane == kWave - 1) s_toff[P] = inc;
  }
  __syncthreads();

  // ---- 3: sorted slot of every owned row; partition of every sorted slot;
  // s_pbase[q] becomes the first segment row of partition q's run minus its
  // tile offset, so sorted slot i goes to segment row s_pbase[q] + i
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    if (lrow(j) < n) {
      pos[j] += (int)s_toff[pid[j]] + s_wcnt[w * P + pid[j]];
      s_spid[pos[j]] = (uint16_t)pid[j];
      s_pos[lrow(j)] = (uint16_t)pos[j];
    }
  }
  for (int q = tid; q < P; q += BLOCK) s_pbase[q] -= s_toff[q];
  __syncthreads();
  auto drow = [&](int i) -> uint32_t { return s_pbase[s_spid[i]] + (uint32_t)i; };

  // ---- 4a: prefetched fixed-width columns, from registers
#pragma unroll
  for (int k = 0; k < NPRE; k++) {
    if (k >= p.npre) break;
    const int c = p.pre_col[k];
    const int oc = col_of<WIDE>(p, c).out_col;
    const SplitOut& O = out_of<WIDE>(p, oc);
    if (O.valid) {
      uint32_t* ov = O.valid;
#pragma unroll
      for (int j = 0; j < RPT; j++) {
        if (lrow(j) < n && (nullm[j] >> k & 1)) {
          const uint32_t d = drow(pos[j]);
          atomicAnd(&ov[d >> 5], ~(1u << (d & 31)));
        }
      }
    }
    const int W = col_of<WIDE>(p, c).width;
    uint8_t* dst = (uint8_t*)O.values;
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      if (lrow(j) < n) {
        if (W == 8) ((uint64_t*)s_stage)[pos[j]] = pv[k][j];
        else if (W == 4) ((uint32_t*)s_stage)[pos[j]] = (uint32_t)pv[k][j];
        else if (W == 2) ((uint16_t*)s_stage)[pos[j]] = (uint16_t)pv[k][j];
        else s_stage[pos[j]] = (uint8_t)pv[k][j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int i = j * BLOCK + tid;
      if (i < n) {
        const uint64_t d = drow(i);
        if (W == 8) gst((uint64_t*)dst + d, ((uint64_t*)s_stage)[i]);
        else if (W == 4) gst((uint32_t*)dst + d, ((uint32_t*)s_stage)[i]);
        else if (W == 2) gst((uint16_t*)dst + d, ((uint16_t*)s_stage)[i]);
        else gst(dst + d, s_stage[i]);
      }
    }
    __syncthreads();
  }

  // ---- 4b: every other column
  for (int c = 0; c < p.ncols; c++) {
    const SplitCol& C = col_of<WIDE>(p, c);
    const int oc = C.out_col;
    if (oc < 0 || C.is_pre) continue;
    const int kind = C.kind;
    if (kind == K_NA) continue;
    const uint8_t* vb = colp(c, 0);
    const uint8_t* src = colp(c, 2);

    // validity: clear the bit of every null row at its destination
    const SplitOut& O = out_of<WIDE>(p, oc);
    if (vb && O.valid) {
      uint32_t* ov = O.valid;
#pragma unroll
      for (int j = 0; j < RPT; j++) {
        if (lrow(j) < n && !bit_get(vb, rb + lrow(j))) {
          uint32_t d = drow(pos[j]);
          atomicAnd(&ov[d >> 5], ~(1u << (d & 31)));
        }
      }
    }

    if (kind == K_BOOL) {
      uint32_t* ov = (uint32_t*)O.values;
#pragma unroll
      for (int j = 0; j < RPT; j++) {
        if (lrow(j) < n && !bit_get(src, rb + lrow(j))) {
          uint32_t d = drow(pos[j]);
          atomicAnd(&ov[d >> 5], ~(1u << (d & 31)));
        }
      }
    } else if (kind == K_FIXED) {
      const int W = C.width;
      const int parts = W > 8 ? W / 8 : 1;  // 16-byte values move as two 8-byte halves
      const int EW = W > 8 ? 8 : W;
      uint8_t* dst = (uint8_t*)O.values;
      for (int part = 0; part < parts; part++) {
        if (EW == 8) {
          const uint64_t* s8 = (const uint64_t*)src;
          uint64_t v[RPT];
#pragma unroll
          for (int j = 0; j < RPT; j++) v[j] = lrow(j) < n ? gld(s8 + (rb + lrow(j)) * parts + part) : 0;
#pragma unroll
          for (int j = 0; j < RPT; j++)
            if (lrow(j) < n) ((uint64_t*)s_stage)[pos[j]] = v[j];
        } else if (EW == 4) {
          const uint32_t* s4 = (const uint32_t*)src;
          uint32_t v[RPT];
#pragma unroll
          for (int j = 0; j < RPT; j++) v[j] = lrow(j) < n ? gld(s4 + rb + lrow(j)) : 0;
#pragma unroll
          for (int j = 0; j < RPT; j++)
            if (lrow(j) < n) ((uint32_t*)s_stage)[pos[j]] = v[j];
        } else if (EW == 2) {
          const uint16_t* s2 = (const uint16_t*)src;
          uint16_t v[RPT];
#pragma unroll
          for (int j = 0; j < RPT; j++) v[j] = lrow(j) < n ? gld(s2 + rb + lrow(j)) : 0;
#pragma unroll
          for (int j = 0; j < RPT; j++)
            if (lrow(j) < n) ((uint16_t*)s_stage)[pos[j]] = v[j];
        } else {
          uint8_t v[RPT];
#pragma unroll
          for (int j = 0; j < RPT; j++) v[j] = lrow(j) < n ? gld(src + rb + lrow(j)) : 0;
#pragma unroll
          for (int j = 0; j < RPT; j++)
            if (lrow(j) < n) s_stage[pos[j]] = v[j];
        }
        __syncthreads();
#pragma unroll 1
        for (int j = 0; j < RPT; j++) {
          int i = j * BLOCK + tid;
          if (i >= n) continue;
          uint64_t d = drow(i);
          if (EW == 8) gst((uint64_t*)dst + d * parts + part, ((uint64_t*)s_stage)[i]);
          else if (EW == 4) gst((uint32_t*)dst + d, ((uint32_t*)s_stage)[i]);
          else if (EW == 2) gst((uint16_t*)dst + d, ((uint16_t*)s_stage)[i]);
          else gst(dst + d, s_stage[i]);
        }
        __syncthreads();
      }
    } else if (kind == K_BINARY) {
      if (bytes_bad) continue;
      const int j_bin = C.bin;
      const bool first = c == bc;  // offsets prefetched, bytes possibly in LDS
      uint32_t* s_src = (uint32_t*)s_stage;  // [TILE] source byte offset by sorted slot
      uint32_t* s_boff = s_src + TILE;       // [TILE+1] sorted byte prefix
      {
        int32_t a0[RPT], a1[RPT];
        const int32_t* soff = (const int32_t*)colp(c, 1) + rb;
#pragma unroll
        for (int j = 0; j < RPT; j++) {
          const bool a = lrow(j) < n;
          a0[j] = first ? o0[j] : (a ? gld(soff + lrow(j)) : 0);
          a1[j] = first ? o1[j] : (a ? gld(soff + lrow(j) + 1) : 0);
        }
#pragma unroll
        for (int j = 0; j < RPT; j++) {
          if (lrow(j) < n) {
            s_src[pos[j]] = (uint32_t)a0[j];
            s_boff[pos[j]] = (uint32_t)(a1[j] - a0[j]);
          }
        }
      }
      const uint32_t* brow = trow + P * (1 + j_bin);
      for (int pp = tid; pp < P; pp += BLOCK) {
        uint32_t rel = gld(brow + pp);
        s_brel[pp] = rel;
        s_bbase[pp] = gld(p.byte_start + (int64_t)j_bin * (P + 1) + pp) + rel;
      }
      __syncthreads();
      {
        // exclusive scan of the sorted lengths; thread tid owns RPT consecutive slots
        uint32_t s = 0;
        int i0 = tid * RPT;
#pragma unroll
        for (int k = 0; k < RPT; k++)
          if (i0 + k < n) s += s_boff[i0 + k];
        uint64_t total;
        uint64_t ex = block_exclusive_scan<uint64_t>(s, s_scan, &total);
        uint32_t run = (uint32_t)ex;
#pragma unroll
        for (int k = 0; k < RPT; k++) {
          if (i0 + k < n) {
            uint32_t v = s_boff[i0 + k];
            s_boff[i0 + k] = run;
            run += v;
          }
        }
        if (tid == 0) s_boff[n] = (uint32_t)total;
      }
      __syncthreads();
      // per partition: offset value and data address adjustments relative to
      // the tile-sorted byte prefix
      uint32_t* s_badj = s_brel;  // brel - run start
      int64_t* s_dadj = s_bbase;  // bbase - run start
      for (int q = tid; q < P; q += BLOCK) {
        const uint32_t run0 = s_boff[s_toff[q]];
        s_badj[q] -= run0;
        s_dadj[q] -= (int64_t)run0;
      }
      __syncthreads();
      uint8_t* odata = O.data;
      int32_t* ooff = O.off;
      // offsets, in sorted-slot order (contiguous per partition run); offsets of
      // partition q start at index row_start + q
#pragma unroll 1
      for (int i = tid; i < n; i += BLOCK) {
        const int q = s_spid[i];
        gst(ooff + (int64_t)(s_pbase[q] + (uint32_t)i) + q + 1, (int32_t)(s_badj[q] + s_boff[i + 1]));
      }
      if (first && chars_lds) {
        // every sorted slot writes its string from the LDS copy of the tile's
        // bytes; neighbouring lanes write neighbouring bytes of a partition run
#pragma unroll 1
        for (int i = tid; i < n; i += BLOCK) {
          const uint32_t b0 = s_boff[i];
          lds_to_hbm(odata + s_dadj[s_spid[i]] + b0, s_cbuf, s_src[i] + cshift, (int)(s_boff[i + 1] - b0));
        }
      } else {
#pragma unroll 1
        for (int i = tid; i < n; i += BLOCK) {
          const uint32_t b0 = s_boff[i];
          copy_string(odata + s_dadj[s_spid[i]] + b0, src + s_src[i], (int)(s_boff[i + 1] - b0));
        }
      }
      __syncthreads();
    }
  }

  // ---- 4c: row indices (an output column without an input)
  if (p.rowid_oc >= 0) {
#pragma unroll
    for (int j = 0; j < RPT; j++)
      if (lrow(j) < n) ((uint32_t*)s_stage)[pos[j]] = (uint32_t)(td.call_row + lrow(j));
    __syncthreads();
    uint32_t* dst = (uint32_t*)out_of<WIDE>(p, p.rowid_oc).values;
    for (int i = tid; i < n; i += BLOCK) gst(dst + drow(i), ((uint32_t*)s_stage)[i]);
  }
}

// Host-side LDS bytes and block size of the scatter kernel.
static size_t scatter_lds(int TILE, int BLOCK, int P, int cbuf) { return scatter_layout(TILE, BLOCK, P, cbuf).total; }
// Partition-ordered utf8 staging: 21 bytes per row, so tiles of strings up to
// that average length go through LDS (and two 2048-row workgroups fit a CU).
static int scatter_cbuf(int TILE, int nbin) { return nbin > 0 ? 21 * TILE : 0; }
static int scatter_block(int P) { return P > 1024 ? 512 : 1024; }

// ---------------------------------------------------------------- large fan-out
// Past the LDS-staged kernels' reach (more than kMaxPartitions partitions, or
// per-tile counters that do not fit LDS) a split is a stable sort of the rows
// by partition id: every row's id and tile slot (split_pid_kernel), a radix
// sort of the (id, slot) pairs (least-significant-digit first, so each
// partition keeps input order), partition starts by binary search over the
// sorted ids, then every output column gathered in sorted order.  A slot is
// t << tile_bits | row of tile t.

// Per output column of the gather: the input column, its kind and width, and
// its output buffers.  kind K_NA with in_col -1 is the row-id column.
struct GatherCol {
  int in_col, kind, width, has_valid;
  int null_slot;  // index of its null count in the totals
  int bin, oc;    // binary column ordinal (-1 otherwise), output column
  void* values;
  uint32_t* valid;
};

struct GatherParams {
  const TileDesc* tiles;
  const void* const* bufs;  // as SplitParams::bufs
  int bstride;
  int tile_bits;
  int64_t N;
  const uint16_t* skey;  // sorted partition ids
  const uint32_t* sval;  // sorted slots
  const GatherCol* cols;
  int ncols;
  unsigned long long* tot;  // null counts at GatherCol::null_slot
};

__global__ __launch_bounds__(256) void split_pid_kernel(SplitParams p, int tile_bits, uint16_t* key, uint32_t* slot) {
  __shared__ const void* s_bp[kMaxCols * 3];
  const int64_t t = blockIdx.x;
  if (t >= p.T) return;
  const TileDesc td = p.tiles[t];
  load_tile_ptrs(p, td.batch, s_bp);
  __syncthreads();
  for (int lr = threadIdx.x; lr < td.nrows; lr += 256) {
    const int pid = compute_pid(p, s_bp, td, lr);
    const int64_t cr = (int64_t)td.call_row + lr;
    key[cr] = (uint16_t)pid;
    slot[cr] = (uint32_t)t << tile_bits | (uint32_t)lr;
  }
}

// row_start[q] = the first sorted row whose id is at least q, q in [0, P].
__global__ void split_sorted_starts_kernel(const uint16_t* skey, int64_t N, int P, int64_t* row_start) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > P) return;
  int64_t lo = 0, hi = N;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int)skey[mid] < q) lo = mid + 1;
    else hi = mid;
  }
  row_start[q] = lo;
}

__device__ __forceinline__ void gather_source(const GatherParams& g, int64_t i, int64_t* r, int* batch,
                                              int32_t* call_row) {
  const uint32_t v = g.sval[i];
  const TileDesc td = g.tiles[v >> g.tile_bits];
  const int lr = (int)(v & ((1u << g.tile_bits) - 1));
  *r = (int64_t)td.row_begin + lr;
  *batch = td.batch;
  *call_row = td.call_row + lr;
}

// Fixed-width values and row ids, four sorted rows per thread (i, i + S,
// i + 2S, i + 3S for a grid of S threads), every load of a column issued
// before its stores.
__global__ __launch_bounds__(256) void split_gather_values_kernel(GatherParams g) {
  constexpr int U = 4;
  const int64_t S = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < g.N; i0 += U * S) {
    int64_t r[U];
    int b[U];
    int32_t cr[U];
    bool act[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      act[u] = i0 + u * S < g.N;
      r[u] = 0;
      b[u] = 0;
      cr[u] = 0;
      if (act[u]) gather_source(g, i0 + u * S, &r[u], &b[u], &cr[u]);
    }
    for (int k = 0; k < g.ncols; k++) {
      const GatherCol& C = g.cols[k];
      if (C.kind == K_NA) {
        if (C.in_col < 0) {
#pragma unroll
          for (int u = 0; u < U; u++)
            if (act[u]) gst((uint32_t*)C.values + i0 + u * S, (uint32_t)cr[u]);
        }
        continue;
      }
      if (C.kind != K_FIXED) continue;
      uint64_t lo[U], hi[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        lo[u] = hi[u] = 0;
        if (!act[u]) continue;
        const uint8_t* src = (const uint8_t*)g.bufs[(int64_t)b[u] * g.bstride + C.in_col * 3 + 2];
        switch (C.width) {
          case 1: lo[u] = gld(src + r[u]); break;
          case 2: lo[u] = gld((const uint16_t*)src + r[u]); break;
          case 4: lo[u] = gld((const uint32_t*)src + r[u]); break;
          case 8: lo[u] = gld((const uint64_t*)src + r[u]); break;
          default:
            lo[u] = gld((const uint64_t*)src + 2 * r[u]);
            hi[u] = gld((const uint64_t*)src + 2 * r[u] + 1);
        }
      }
      uint8_t* dst = (uint8_t*)C.values;
#pragma unroll
      for (int u = 0; u < U; u++) {
        if (!act[u]) continue;
        const int64_t i = i0 + u * S;
        switch (C.width) {
          case 1: gst(dst + i, (uint8_t)lo[u]); break;
          case 2: gst((uint16_t*)dst + i, (uint16_t)lo[u]); break;
          case 4: gst((uint32_t*)dst + i, (uint32_t)lo[u]); break;
          case 8: gst((uint64_t*)dst + i, lo[u]); break;
          default:
            gst((uint64_t*)dst + 2 * i, lo[u]);
            gst((uint64_t*)dst + 2 * i + 1, hi[u]);
        }
      }
    }
  }
}

// Validity bitmaps and boolean values, one 32-bit word (32 sorted rows) per
// thread, written whole: bits past the last row stay set, as in the bitmaps the
// LDS-staged scatter pre-fills with ones.  Null counts are added per wave.
__global__ __launch_bounds__(256) void split_gather_bits_kernel(GatherParams g, int64_t nwords) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t w_end = (nwords + stride - 1) / stride * stride;  // every lane runs the same trips
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < w_end; w += stride) {
    const bool live = w < nwords;
    int64_t r[32];
    int b[32];
    const int64_t i0 = w * 32;
    const int rows = live ? (int)max<int64_t>(0, min<int64_t>(32, g.N - i0)) : 0;
#pragma unroll
    for (int k = 0; k < 32; k++) {
      int32_t cr;
      r[k] = 0;
      b[k] = 0;
      if (k < rows) gather_source(g, i0 + k, &r[k], &b[k], &cr);
    }
    const uint32_t live_bits = rows == 32 ? 0xFFFFFFFFu : ((1u << rows) - 1u);
    for (int c = 0; c < g.ncols; c++) {
      const GatherCol& C = g.cols[c];
      for (int part = 0; part < 2; part++) {  // 0: validity, 1: boolean values
        uint32_t* out = part == 0 ? (C.has_valid ? C.valid : nullptr) : (C.kind == K_BOOL ? (uint32_t*)C.values : nullptr);
        if (!out) continue;
        uint32_t word = ~live_bits;
        for (int k = 0; k < rows; k++) {
          const uint8_t* bm = (const uint8_t*)g.bufs[(int64_t)b[k] * g.bstride + C.in_col * 3 + (part == 0 ? 0 : 2)];
          const uint32_t bit = bm ? (uint32_t)bit_get(bm, r[k]) : 1u;
          word |= bit << k;
        }
        if (live) gst(out + w, word);
        if (part == 0) {
          const unsigned long long nulls = wave_reduce_sum((unsigned long long)__popc(~word & live_bits));
          if (__lane_id() == 0 && nulls) atomicAdd(g.tot + C.null_slot, nulls);
        }
      }
    }
  }
}

// utf8 lengths of one binary column in sorted order; len[N] = 0 so that an
// exclusive scan of N + 1 entries ends with the total.
__global__ __launch_bounds__(256) void split_gather_lens_kernel(GatherParams g, int in_col, int64_t* len) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= g.N; i += stride) {
    if (i == g.N) {
      len[i] = 0;
      continue;
    }
    int64_t r;
    int b;
    int32_t cr;
    gather_source(g, i, &r, &b, &cr);
    const int32_t* off = (const int32_t*)g.bufs[(int64_t)b * g.bstride + in_col * 3 + 1];
    len[i] = (int64_t)(gld(off + r + 1) - gld(off + r));
  }
}

// Per partition: its first byte (byte_start[q] = scanned length at its first
// row, q in [0, P]), the leading 0 of its offsets, and the 2 GiB check.
__global__ void split_sorted_bytes_kernel(const int64_t* row_start, const int64_t* gb, int P, int64_t* byte_start,
                                          int32_t* out_off, int* err) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > P) return;
  const int64_t b0 = gb[row_start[q]];
  byte_start[q] = b0;
  if (q < P) {
    out_off[row_start[q] + q] = 0;
    if (gb[row_start[q + 1]] - b0 >= 0x80000000ll) atomicOr(err, 2);
  }
}

// Offsets and bytes of one binary column, one sorted row per thread: row i of
// partition q ends at offsets[i + q + 1] (each partition's offsets start with
// its own 0), its bytes go to data + gb[i]; a row ending past the data
// output's capacity (offsets past the declared data buffers) sets error 8 and
// writes nothing.
__global__ __launch_bounds__(256) void split_gather_strings_kernel(GatherParams g, int in_col, const int64_t* gb,
                                                                   const int64_t* byte_start, int32_t* out_off,
                                                                   uint8_t* data, int64_t cap, int* err) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += stride) {
    int64_t r;
    int b;
    int32_t cr;
    gather_source(g, i, &r, &b, &cr);
    const int q = g.skey[i];
    const int64_t e = gb[i + 1];
    gst(out_off + i + q + 1, (int32_t)(e - byte_start[q]));
    if (e > cap) {
      atomicOr(err, 8);
      continue;
    }
    const int32_t* off = (const int32_t*)g.bufs[(int64_t)b * g.bstride + in_col * 3 + 1];
    const uint8_t* src = (const uint8_t*)g.bufs[(int64_t)b * g.bstride + in_col * 3 + 2];
    copy_string_dw(data + gb[i], src + gld(off + r), (int)(e - gb[i]));
  }
}

// Per input batch of a sorted segment: utf8 bytes of every binary column and
// nulls of every output column, stats[batch * (nbin + nout) + (j | nbin + oc)]
// (the partition writer's per-batch buffer sizes); one workgroup per tile.
__global__ __launch_bounds__(256) void split_batch_stats_kernel(GatherParams g, int64_t T, int nbin, int nout,
                                                                unsigned long long* stats) {
  const int64_t t = blockIdx.x;
  if (t >= T) return;
  const TileDesc td = g.tiles[t];
  unsigned long long* row = stats + (int64_t)td.batch * (nbin + nout);
  for (int k = 0; k < g.ncols; k++) {
    const GatherCol& C = g.cols[k];
    if (C.in_col < 0) continue;
    const void* const* bp = g.bufs + (int64_t)td.batch * g.bstride + C.in_col * 3;
    if (C.bin >= 0 && threadIdx.x == 0) {
      const int32_t* off = (const int32_t*)bp[1];
      const int32_t b = gld(off + td.row_begin + td.nrows) - gld(off + td.row_begin);
      if (b) atomicAdd(row + C.bin, (unsigned long long)b);
    }
    const uint8_t* vb = (const uint8_t*)bp[0];
    if (!C.has_valid || !vb) continue;
    // tiles start on multiples of 8 rows inside a batch
    unsigned long long nulls = 0;
    const int nbytes = (td.nrows + 7) >> 3;
    for (int i = threadIdx.x; i < nbytes; i += 256) {
      const uint32_t b = gld(vb + (td.row_begin >> 3) + i);
      const int bits = min(8, td.nrows - i * 8);
      nulls += (unsigned long long)__popc(~b & ((1u << bits) - 1u));
    }
    nulls = wave_reduce_sum(nulls);
    if (__lane_id() == 0 && nulls) atomicAdd(row + nbin + C.oc, nulls);
  }
}

// Rows per (input batch, partition) of a sorted segment, for the partition
// writer's batch cuts: counts[batch * P + q].
__global__ __launch_bounds__(256) void split_sorted_batch_counts_kernel(const uint16_t* skey, const uint32_t* sval,
                                                                        int64_t N, const int32_t* tile_batch,
                                                                        int tile_bits, int P, uint32_t* counts) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
    const int b = tile_batch[sval[i] >> tile_bits];
    atomicAdd(counts + (int64_t)b * P + skey[i], 1u);
  }
}

}  // namespace gz

// =============================================================== host side
using namespace gz;

struct SplitSegment {
  int64_t N = 0;
  int32_t nbatches = 0;
  int64_t T = 0;
  int R = 0;  // counters per tile row
  DBuf counts, prefix, csum, tot, row_start, byte_start, err, offtab;
  std::vector<DBuf> values, validity, offsets, data;
  std::vector<int32_t> tile_batch;  // batch of each tile
  // a large fan-out segment: sorted partition ids and tile slots; its counts
  // are rows per (batch, partition), made when first asked for
  bool sorted = false;
  int tile_bits = 0;
  DBuf skey, sval, bstats;
  // host copies, valid after gz_split_stop
  std::vector<int64_t> h_row_start, h_byte_start;
  std::vector<unsigned long long> h_null;
  bool resolved = false;
};

struct gz_ipc_plan;
static void ipc_plan_free(gz_context* ctx, gz_ipc_plan* pl);
struct gz_splitter {
  gz_context* ctx = nullptr;
  gz_ipc_plan* ipc = nullptr;  // partition writer layout (gz_split_ipc_prepare)
  int codec = GZ_COMPRESSION_UNCOMPRESSED;  // partition writer body compression
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
  std::vector<SplitCol> cols;  // per input column (nullc and is_pre set per call)
  DBuf d_cols, d_outs;         // a call's column and output tables
  std::vector<SplitSegment*> segments;
  bool sorted = false;  // large fan-out: rows sorted by partition id, then gathered
  DBuf d_tiles, d_bufs;
  bool stopped = false;
  int64_t compute_pid_ns = 0;
  // metrics: events around every partition id pass (summed at Stop), host
  // wall time of the data file's layout and assembly, and of its compression
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pid_events, pid_free;
  int64_t write_ns = 0, compress_ns = 0;
  // per-batch calls staged for the batched path (gz_split_batch)
  BatchStage stage;
};

// Elapsed time of the recorded partition id passes, added to compute_pid_ns;
// their events are kept for reuse.
static void drain_pid_events(gz_splitter* h) {
  for (auto& ev : h->pid_events) {
    float ms = 0;
    if (hipEventSynchronize(ev.second) == hipSuccess && hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess)
      h->compute_pid_ns += (int64_t)((double)ms * 1e6);
    h->pid_free.push_back(ev);
  }
  h->pid_events.clear();
  (void)hipGetLastError();
}

// A pair of events around this call's partition id pass (at most 256 pending).
static std::pair<hipEvent_t, hipEvent_t> pid_event_pair(gz_splitter* h) {
  if (h->pid_events.size() >= 256) drain_pid_events(h);
  std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
  if (!h->pid_free.empty()) {
    ev = h->pid_free.back();
    h->pid_free.pop_back();
  } else if (hipEventCreate(&ev.first) != hipSuccess || hipEventCreate(&ev.second) != hipSuccess) {
    (void)hipGetLastError();
    return {nullptr, nullptr};
  }
  return ev;
}

static int split_release_segment(gz_context* ctx, SplitSegment* s) {
  for (DBuf* b : {&s->counts, &s->prefix, &s->csum, &s->tot, &s->row_start, &s->byte_start, &s->err,
                  &s->offtab, &s->skey, &s->sval, &s->bstats})
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
  o->batch_compress_threshold = 256;
  o->hash_seed = 42;
  o->append_row_ids = 0;
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
  if (num_cols > kMaxCols) return fail(GZ_NOT_IMPLEMENTED, "more than " + std::to_string(kMaxCols) + " columns");
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
      return fail(GZ_NOT_IMPLEMENTED, "at most " + std::to_string(kMaxKeys) + " hash keys");
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
  h->tile_rows = o.tile_rows;
  if (h->tile_rows != 4096 && h->tile_rows != 2048 && h->tile_rows != 1024) h->tile_rows = 0;
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
  h->cols.assign(num_cols, SplitCol{});
  for (int c = 0; c < num_cols; c++) {
    int t = col_types[c];
    SplitCol& C = h->cols[c];
    C.kind = t == GZ_TYPE_NA ? K_NA : t == GZ_TYPE_BOOL ? K_BOOL : type_is_binary(t) ? K_BINARY : K_FIXED;
    C.width = type_width(t);
    C.nullc = -1;
    if (mode == 2 && c == 0) {
      C.out_col = -1;
      C.bin = -1;
      continue;
    }
    C.out_col = oc++;
    h->out_types.push_back(t);
    if (type_is_binary(t) && nbin == 0) b.bin_col0 = c;
    C.bin = type_is_binary(t) ? nbin++ : -1;
  }
  b.nbin = nbin;
  b.rowid_oc = -1;
  if (o.append_row_ids) {
    if (oc >= kMaxCols) {
      delete h;
      return fail(GZ_NOT_IMPLEMENTED, "more than 32 output columns");
    }
    b.rowid_oc = oc++;
    h->out_types.push_back(GZ_TYPE_INT32);
  }
  // default tile: 2048 rows (two 1024-thread workgroups per CU, also with the
  // utf8 staging buffer)
  if (h->tile_rows == 0) h->tile_rows = 2048;
  if (P > 1024) h->tile_rows = 2048;  // 512-thread scatter for large fan-out
  // Past the LDS-staged kernels' reach (the scatter's per-partition tables, or
  // the count kernel's counters: rows and utf8 bytes per partition, nulls per
  // output column) the split sorts rows by partition id and gathers them.
  if (P > kMaxPartitions ||
      scatter_lds(h->tile_rows, scatter_block(P), P, scatter_cbuf(h->tile_rows, nbin)) > 160 * 1024 ||
      sizeof(uint32_t) * ((size_t)P * (1 + nbin) + std::max(oc, 1)) > kCountLdsMax) {
    h->sorted = true;
    h->tile_rows = 4096;
  }
  *out = h;
  return GZ_OK;
}

}  // extern "C"

// Keeps a launched segment; the range splitter validates partition ids like
// the reference (splitter.cc:1511-1516), so its call waits for the check.
static int split_commit(gz_splitter* h, SplitSegment* seg) {
  h->segments.push_back(seg);
  if (h->mode == 2) {
    int err = 0;
    GZ_HIP(hipMemcpyAsync(&err, seg->err.p, sizeof(int), hipMemcpyDeviceToHost, h->ctx->stream));
    GZ_HIP(hipStreamSynchronize(h->ctx->stream));
    if (err & 1) return fail(GZ_INVALID, "Partition id is equal or greater than num_partitions");
  }
  return GZ_OK;
}

// The utf8 data output of a call is sized from the scanned byte total (one
// host round trip) when batches share data buffers (slices of one buffer, whose
// summed sizes may be many times the real size) or the declared sizes pass
// 16 GiB; otherwise the declared data buffers' sizes bound it.
static bool exact_data_size(bool shared, size_t declared) { return shared || declared > ((size_t)16 << 30); }

static int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)); }

// A segment of a large fan-out splitter (see split_pid_kernel).  The row
// outputs are allocated; the utf8 data outputs are sized and allocated here.
static int split_run_sorted(gz_splitter* h, SplitSegment* seg, const SplitParams& prm,
                            const std::vector<SplitCol>& cols, std::vector<SplitOut>& outs,
                            const std::vector<size_t>& data_bytes, const std::vector<bool>& data_shared,
                            int64_t nwords) {
  gz_context* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  const int P = h->P, nc = h->ncols_in, nbin = h->base.nbin;
  const int nout = (int)h->out_types.size();
  const int64_t N = seg->N, T = seg->T;
  int tile_bits = 0;
  while ((1 << tile_bits) < h->tile_rows) tile_bits++;
  if (T > 0 && ((uint64_t)(T - 1) >> (32 - tile_bits)) != 0)
    return fail(GZ_CAPACITY_ERROR, "too many batches in one large fan-out split call");
  seg->sorted = true;
  seg->tile_bits = tile_bits;
  DBuf key_in, val_in, tmp, glen, gb, gcols;
  auto done = [&](int code) {
    for (DBuf* b : {&key_in, &val_in, &tmp, &glen, &gb, &gcols}) ctx_free(ctx, b);
    return code;
  };
  auto hip = [&](hipError_t e, const char* what) {
    return e == hipSuccess ? GZ_OK : fail(GZ_HIP_ERROR, std::string(what) + ": " + hipGetErrorString(e));
  };
  const size_t n1 = (size_t)std::max<int64_t>(N, 1);
  int rc;
  if ((rc = ctx_alloc(ctx, 2 * n1, &key_in)) || (rc = ctx_alloc(ctx, 4 * n1, &val_in)) ||
      (rc = ctx_alloc(ctx, 2 * n1, &seg->skey)) || (rc = ctx_alloc(ctx, 4 * n1, &seg->sval)))
    return done(rc);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(2 * (int64_t)seg->R)), dim3(256), 0, st, (uint32_t*)seg->tot.p,
                     2 * (int64_t)seg->R, 0u);
  // ---- partition ids, stable sort by id
  if (T > 0) {
    KernelTimer kt(ctx, "split_count");
    const std::pair<hipEvent_t, hipEvent_t> ev = pid_event_pair(h);
    if (ev.first) {
      hipEventRecord(ev.first, st);
      h->pid_events.push_back(ev);
    }
    hipLaunchKernelGGL(split_pid_kernel, dim3((unsigned)T), dim3(256), 0, st, prm, tile_bits, (uint16_t*)key_in.p,
                       (uint32_t*)val_in.p);
    if (ev.first) hipEventRecord(ev.second, st);
  }
  {
    KernelTimer kt(ctx, "split_scan");
    if (N > 0) {
      unsigned end_bit = 1;
      while ((1 << end_bit) < P) end_bit++;
      size_t tb = 0;
      if ((rc = hip(rocprim::radix_sort_pairs(nullptr, tb, (const uint16_t*)key_in.p, (uint16_t*)seg->skey.p,
                                              (const uint32_t*)val_in.p, (uint32_t*)seg->sval.p, (size_t)N, 0u,
                                              end_bit, st),
                    "radix_sort_pairs")))
        return done(rc);
      if ((rc = ctx_alloc(ctx, std::max<size_t>(tb, 4), &tmp))) return done(rc);
      if ((rc = hip(rocprim::radix_sort_pairs(tmp.p, tb, (const uint16_t*)key_in.p, (uint16_t*)seg->skey.p,
                                              (const uint32_t*)val_in.p, (uint32_t*)seg->sval.p, (size_t)N, 0u,
                                              end_bit, st),
                    "radix_sort_pairs")))
        return done(rc);
    }
    hipLaunchKernelGGL(split_sorted_starts_kernel, dim3((unsigned)((P + 256) / 256)), dim3(256), 0, st,
                       (const uint16_t*)seg->skey.p, N, P, (int64_t*)seg->row_start.p);
  }
  // ---- gathers
  KernelTimer kt(ctx, "split_scatter");
  std::vector<GatherCol> gc;
  for (int c = 0; c < nc; c++) {
    const int oc = cols[c].out_col;
    if (oc < 0 || cols[c].kind == K_NA) continue;
    gc.push_back(GatherCol{c, cols[c].kind, cols[c].width, outs[oc].valid != nullptr, P * (1 + nbin) + oc,
                           cols[c].bin, oc, outs[oc].values, outs[oc].valid});
  }
  if (prm.rowid_oc >= 0)
    gc.push_back(GatherCol{-1, K_NA, 4, 0, P * (1 + nbin) + prm.rowid_oc, -1, prm.rowid_oc,
                           outs[prm.rowid_oc].values, nullptr});
  if ((rc = upload(ctx, &gcols, gc.data(), sizeof(GatherCol) * std::max<size_t>(gc.size(), 1)))) return done(rc);
  GatherParams g;
  g.tiles = prm.tiles;
  g.bufs = prm.bufs;
  g.bstride = prm.bstride;
  g.tile_bits = tile_bits;
  g.N = N;
  g.skey = (const uint16_t*)seg->skey.p;
  g.sval = (const uint32_t*)seg->sval.p;
  g.cols = (const GatherCol*)gcols.p;
  g.ncols = (int)gc.size();
  g.tot = (unsigned long long*)seg->tot.p;
  const int64_t nstat = (int64_t)std::max<int32_t>(seg->nbatches, 1) * (nbin + nout);
  if ((rc = ctx_alloc(ctx, sizeof(unsigned long long) * std::max<int64_t>(nstat, 1), &seg->bstats))) return done(rc);
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(2 * nstat)), dim3(256), 0, st, (uint32_t*)seg->bstats.p,
                     2 * nstat, 0u);
  if (T > 0)
    hipLaunchKernelGGL(split_batch_stats_kernel, dim3((unsigned)T), dim3(256), 0, st, g, T, nbin, nout,
                       (unsigned long long*)seg->bstats.p);
  if (N > 0) {
    hipLaunchKernelGGL(split_gather_values_kernel, dim3(grid_for((N + 3) / 4)), dim3(256), 0, st, g);
    hipLaunchKernelGGL(split_gather_bits_kernel, dim3(grid_for(nwords)), dim3(256), 0, st, g, nwords);
  }
  // utf8 columns: lengths in sorted order, their scan, per-partition starts,
  // then offsets and bytes
  if (nbin) {
    if ((rc = ctx_alloc(ctx, 8 * (n1 + 1), &glen)) || (rc = ctx_alloc(ctx, 8 * (n1 + 1), &gb))) return done(rc);
  }
  size_t sb = 0;
  for (int c = 0; c < nc; c++) {
    const int oc = cols[c].out_col, j = cols[c].bin;
    if (oc < 0 || j < 0) continue;
    int64_t* bstart = (int64_t*)seg->byte_start.p + (int64_t)j * (P + 1);
    hipLaunchKernelGGL(split_gather_lens_kernel, dim3(grid_for(N + 1)), dim3(256), 0, st, g, c, (int64_t*)glen.p);
    size_t need = 0;
    if ((rc = hip(rocprim::exclusive_scan(nullptr, need, (const int64_t*)glen.p, (int64_t*)gb.p, (int64_t)0,
                                          (size_t)N + 1, rocprim::plus<int64_t>(), st),
                  "exclusive_scan")))
      return done(rc);
    if (need > sb) {
      ctx_free(ctx, &tmp);
      if ((rc = ctx_alloc(ctx, need, &tmp))) return done(rc);
      sb = need;
    }
    if ((rc = hip(rocprim::exclusive_scan(tmp.p, need, (const int64_t*)glen.p, (int64_t*)gb.p, (int64_t)0,
                                          (size_t)N + 1, rocprim::plus<int64_t>(), st),
                  "exclusive_scan")))
      return done(rc);
    hipLaunchKernelGGL(split_sorted_bytes_kernel, dim3((unsigned)((P + 256) / 256)), dim3(256), 0, st,
                       (const int64_t*)seg->row_start.p, (const int64_t*)gb.p, P, bstart, outs[oc].off,
                       (int*)seg->err.p);
    // data: the input data buffers bound it, unless batches share them (or the
    // bound is past 16 GiB): then the scanned total, one host round trip
    size_t bytes = data_bytes[j];
    const bool exact = exact_data_size(data_shared[j], bytes);
    if (exact) {
      int64_t total = 0;
      if ((rc = hip(hipMemcpyAsync(&total, (const int64_t*)gb.p + N, sizeof(int64_t), hipMemcpyDeviceToHost, st),
                    "hipMemcpyAsync")) ||
          (rc = hip(hipStreamSynchronize(st), "hipStreamSynchronize")))
        return done(rc);
      bytes = (size_t)total;
    }
    if ((rc = ctx_alloc(ctx, std::max<size_t>(bytes, 4), &seg->data[oc]))) return done(rc);
    outs[oc].data = (uint8_t*)seg->data[oc].p;
    if (N > 0)
      hipLaunchKernelGGL(split_gather_strings_kernel, dim3(grid_for(N)), dim3(256), 0, st, g, c,
                         (const int64_t*)gb.p, (const int64_t*)bstart, outs[oc].off, outs[oc].data,
                         exact ? INT64_MAX : (int64_t)bytes, (int*)seg->err.p);
  }
  if ((rc = hip(hipGetLastError(), "split gather launch"))) return done(rc);
  // the temporaries go back to the pool, which hands them only to work queued
  // later on this stream
  return done(GZ_OK);
}

// Rows per (batch, partition) of a large fan-out segment, counts[b * P + q],
// made on first use.
static int sorted_batch_counts(gz_splitter* h, SplitSegment* s) {
  if (!s->sorted || s->counts.p) return GZ_OK;
  gz_context* ctx = h->ctx;
  const int P = h->P;
  const int64_t nb = std::max<int32_t>(s->nbatches, 1);
  int rc;
  if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * (size_t)nb * P, &s->counts))) return rc;
  DBuf tb;
  if ((rc = upload(ctx, &tb, s->tile_batch.data(), sizeof(int32_t) * std::max<size_t>(s->tile_batch.size(), 1)))) {
    ctx_free(ctx, &s->counts);
    return rc;
  }
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(nb * P)), dim3(256), 0, ctx->stream, (uint32_t*)s->counts.p,
                     nb * P, 0u);
  if (s->N > 0)
    hipLaunchKernelGGL(split_sorted_batch_counts_kernel, dim3(grid_for(s->N)), dim3(256), 0, ctx->stream,
                       (const uint16_t*)s->skey.p, (const uint32_t*)s->sval.p, s->N, (const int32_t*)tb.p,
                       s->tile_bits, P, (uint32_t*)s->counts.p);
  ctx_free(ctx, &tb);
  GZ_HIP(hipGetLastError());
  return GZ_OK;
}

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
  std::vector<bool> data_shared(h->base.nbin, false);  // some data buffer backs several batches
  std::vector<std::vector<const void*>> data_ptrs(h->base.nbin);
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
      if (type_is_binary(t)) {
        data_bytes[h->cols[c].bin] += (size_t)vsz;
        data_ptrs[h->cols[c].bin].push_back(slot[2]);
      }
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
      td.call_row = (int32_t)(N + r0);
      td.pad = 0;
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
  // batches sharing a utf8 data buffer (slices of one buffer)
  for (int b = 0; b < h->base.nbin; b++) {
    auto& v = data_ptrs[b];
    std::sort(v.begin(), v.end());
    data_shared[b] = std::adjacent_find(v.begin(), v.end()) != v.end();
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
  // ---- allocate the segment.  Counter row of a tile: rows per partition, utf8
  // bytes per (binary column, partition), nulls per output column.
  const int R = P * (1 + nbin) + std::max(nout, 1);
  seg->R = R;
  const int64_t C = (std::max<int64_t>(T, 1) + kScanChunk - 1) / kScanChunk;
  if (!h->sorted) {
    if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * R * std::max<int64_t>(T, 1), &seg->counts))) return bail(rc);
    if ((rc = ctx_alloc(ctx, sizeof(uint32_t) * R * std::max<int64_t>(T, 1), &seg->prefix))) return bail(rc);
    if ((rc = ctx_alloc(ctx, sizeof(uint64_t) * R * C, &seg->csum))) return bail(rc);
  }
  if (nbin) {
    if ((rc = ctx_alloc(ctx, sizeof(int64_t) * nbin * (P + 1), &seg->byte_start))) return bail(rc);
  }
  if ((rc = ctx_alloc(ctx, sizeof(uint64_t) * R, &seg->tot))) return bail(rc);
  if ((rc = ctx_alloc(ctx, sizeof(int64_t) * (P + 1), &seg->row_start))) return bail(rc);
  if ((rc = ctx_alloc(ctx, sizeof(int), &seg->err))) return bail(rc);
  seg->values.resize(nout);
  seg->validity.resize(nout);
  seg->offsets.resize(nout);
  seg->data.resize(nout);
  const int64_t nwords = (N + 63) / 64 * 2;  // 32-bit words of a 64-bit padded bitmap
  SplitParams prm = h->base;
  std::vector<SplitCol> cols = h->cols;  // this call's column table
  std::vector<SplitOut> outs(std::max(nout, 1), SplitOut{nullptr, nullptr, nullptr, nullptr});
  for (int c = 0; c < nc; c++) {
    int oc = cols[c].out_col;
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
    outs[oc].values = seg->values[oc].p;
    outs[oc].valid = (uint32_t*)seg->validity[oc].p;
    outs[oc].off = (int32_t*)seg->offsets[oc].p;
    outs[oc].data = (uint8_t*)seg->data[oc].p;
  }
  if (prm.rowid_oc >= 0) {
    if ((rc = ctx_alloc(ctx, (size_t)std::max<int64_t>(N, 1) * 4, &seg->values[prm.rowid_oc]))) return bail(rc);
    outs[prm.rowid_oc].values = seg->values[prm.rowid_oc].p;
  }
  // output columns of an input column are indexed by the output index
  if ((rc = upload(ctx, &h->d_tiles, tiles.data(), sizeof(TileDesc) * tiles.size()))) return bail(rc);
  if ((rc = upload(ctx, &h->d_bufs, bufs.data(), sizeof(void*) * bufs.size()))) return bail(rc);
  prm.tiles = (const TileDesc*)h->d_tiles.p;
  prm.T = T;
  prm.bufs = (const void* const*)h->d_bufs.p;
  prm.bstride = nc * 3;
  prm.R = R;
  for (int c = 0; c < nc; c++) {
    int oc = cols[c].out_col;
    cols[c].nullc = (oc >= 0 && has_validity[c] && h->types_in[c] != GZ_TYPE_NA) ? P * (1 + nbin) + oc : -1;
  }
  prm.counts = (uint32_t*)seg->counts.p;
  prm.prefix = (uint32_t*)seg->prefix.p;
  prm.row_start = (const int64_t*)seg->row_start.p;
  prm.byte_start = (const int64_t*)seg->byte_start.p;
  prm.err = (int*)seg->err.p;
  hipStream_t st = ctx->stream;
  // ---- fills (kernels, not memsets, so the sequence can be graph-captured)
  auto fill = [&](void* p, int64_t n32, uint32_t v) {
    if (!p || n32 <= 0) return;
    int blocks = (int)std::min<int64_t>((n32 + 255) / 256, 4096);
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, st, (uint32_t*)p, n32, v);
  };
  fill(seg->err.p, 1, 0);
  if (h->sorted) {
    if ((rc = split_run_sorted(h, seg, prm, cols, outs, data_bytes, data_shared, nwords))) return bail(rc);
    return split_commit(h, seg);
  }
  for (int oc = 0; oc < nout; oc++) {
    if (seg->validity[oc].p) fill(seg->validity[oc].p, nwords, 0xFFFFFFFFu);
    if (h->out_types[oc] == GZ_TYPE_BOOL) fill(seg->values[oc].p, nwords, 0xFFFFFFFFu);
  }
  const int64_t grid = (T + 7) / 8 * 8;
  // hot path: hash of one int64 key that has no validity buffer in this call
  // wide schemas read the column tables from device memory (generic hash path)
  const int wide = (nc > kArgCols || nout > kArgCols) ? 1 : 0;
  const int keyk = (!wide && h->mode == 0 && prm.nkeys == 1 && prm.key_type[0] == GZ_TYPE_INT64 &&
                    !has_validity[prm.key_col[0]]) ? 1 : 0;
  // fixed-width columns the scatter loads before ranking (the key first when the
  // int64-key path hashes it from those registers)
  prm.npre = 0;
  for (int c = 0; c < nc; c++) cols[c].is_pre = 0;
  const int max_pre = scatter_max_pre(TILE / scatter_block(P));
  auto add_pre = [&](int c) {
    if (prm.npre >= max_pre || cols[c].is_pre || cols[c].out_col < 0 || cols[c].kind != K_FIXED ||
        cols[c].width > 8)
      return;
    prm.pre_col[prm.npre++] = c;
    cols[c].is_pre = 1;
  };
  if (keyk) add_pre(prm.key_col[0]);
  for (int c = 0; c < nc; c++) add_pre(c);
  if (wide) {
    if ((rc = upload(ctx, &h->d_cols, cols.data(), sizeof(SplitCol) * cols.size()))) return bail(rc);
    prm.col = (const SplitCol*)h->d_cols.p;
  } else {
    for (int c = 0; c < nc; c++) prm.acol[c] = cols[c];
  }
  if (T > 0) {
    KernelTimer kt(ctx, "split_count");
    const std::pair<hipEvent_t, hipEvent_t> ev = pid_event_pair(h);
    if (ev.first) {
      hipEventRecord(ev.first, st);
      h->pid_events.push_back(ev);
    }
    size_t lds = sizeof(uint32_t) * R;
    if (wide)
      hipLaunchKernelGGL((split_count_kernel<256, 0, 1>), dim3((unsigned)grid), dim3(256), lds, st, prm);
    else if (keyk)
      hipLaunchKernelGGL((split_count_kernel<256, 1, 0>), dim3((unsigned)grid), dim3(256), lds, st, prm);
    else
      hipLaunchKernelGGL((split_count_kernel<256, 0, 0>), dim3((unsigned)grid), dim3(256), lds, st, prm);
    if (ev.first) hipEventRecord(ev.second, st);
  }
  {
    KernelTimer kt(ctx, "split_scan");
    if (T > 0) {
      hipLaunchKernelGGL(tile_scan_sum_kernel<kScanChunk>, dim3((unsigned)C), dim3(256), 0, st,
                         (const uint32_t*)seg->counts.p, R, T, (uint64_t*)seg->csum.p);
      hipLaunchKernelGGL(tile_scan_chunks_kernel, dim3((unsigned)R), dim3(256), 0, st,
                         (uint64_t*)seg->csum.p, R, C, (uint64_t*)seg->tot.p);
      hipLaunchKernelGGL(tile_scan_write_kernel<kScanChunk>, dim3((unsigned)C), dim3(256), 0, st,
                         (const uint32_t*)seg->counts.p, (const uint64_t*)seg->csum.p, R, T,
                         (uint32_t*)seg->prefix.p);
    } else {
      fill(seg->tot.p, 2 * (int64_t)R, 0);
    }
    // offsets outputs, then the data capacities: the declared data buffers'
    // sizes, or no limit where the data output is sized from the scanned totals
    std::vector<int64_t> offp(2 * (size_t)std::max(nbin, 1), 0);
    for (int c = 0; c < nc; c++) {
      const int j = cols[c].bin;
      if (j < 0) continue;
      offp[j] = (int64_t)(uintptr_t)outs[cols[c].out_col].off;
      offp[nbin + j] = exact_data_size(data_shared[j], data_bytes[j]) ? INT64_MAX : (int64_t)data_bytes[j];
    }
    if ((rc = upload(ctx, &seg->offtab, offp.data(), sizeof(int64_t) * offp.size()))) return bail(rc);
    hipLaunchKernelGGL(split_layout_kernel, dim3(1), dim3(256), 0, st, (const uint64_t*)seg->tot.p,
                       P, nbin, (int64_t*)seg->row_start.p, (int64_t*)seg->byte_start.p,
                       (const int64_t*)seg->offtab.p, (int*)seg->err.p);
  }
  // utf8 data outputs (exact_data_size: bound or scanned total)
  for (int c = 0; c < nc; c++) {
    int oc = cols[c].out_col;
    if (oc < 0 || !type_is_binary(h->types_in[c])) continue;
    size_t bytes = data_bytes[cols[c].bin];
    if (exact_data_size(data_shared[cols[c].bin], bytes)) {
      int64_t total = 0;
      GZ_HIP(hipMemcpyAsync(&total, (int64_t*)seg->byte_start.p + (int64_t)cols[c].bin * (P + 1) + P,
                            sizeof(int64_t), hipMemcpyDeviceToHost, st));
      GZ_HIP(hipStreamSynchronize(st));
      bytes = (size_t)total;
    }
    if ((rc = ctx_alloc(ctx, std::max<size_t>(bytes, 4), &seg->data[oc]))) return bail(rc);
    outs[oc].data = (uint8_t*)seg->data[oc].p;
  }
  if (wide) {
    if ((rc = upload(ctx, &h->d_outs, outs.data(), sizeof(SplitOut) * outs.size()))) return bail(rc);
    prm.out = (const SplitOut*)h->d_outs.p;
  } else {
    for (int oc = 0; oc < (int)outs.size(); oc++) prm.aout[oc] = outs[oc];
  }
  if (T > 0) {
    KernelTimer kt(ctx, "split_scatter");
    const int BLOCK = scatter_block(P);
    prm.cbuf = scatter_cbuf(TILE, nbin);
    size_t lds = scatter_lds(TILE, BLOCK, P, prm.cbuf);
#define GZ_SCATTER(TL, BL, KK, WD) \
  hipLaunchKernelGGL((split_scatter_kernel<TL, BL, KK, WD>), dim3((unsigned)grid), dim3(BL), lds, st, prm)
#define GZ_SCATTER_KW(TL, BL) \
  if (wide) GZ_SCATTER(TL, BL, 0, 1); else if (keyk) GZ_SCATTER(TL, BL, 1, 0); else GZ_SCATTER(TL, BL, 0, 0)
    if (BLOCK == 512) { GZ_SCATTER_KW(2048, 512); }
    else if (TILE == 4096) { GZ_SCATTER_KW(4096, 1024); }
    else if (TILE == 2048) { GZ_SCATTER_KW(2048, 1024); }
    else { GZ_SCATTER_KW(1024, 1024); }
#undef GZ_SCATTER_KW
#undef GZ_SCATTER
  }
  GZ_HIP(hipGetLastError());
  return split_commit(h, seg);
}

// Runs the staged per-batch calls as one segment.
static int split_flush(gz_splitter* h) {
  if (h->stage.empty()) return GZ_OK;
  std::vector<gz_batch> v;
  stage_view(h->stage, &v);
  const int rc = split_run(h, (int32_t)v.size(), v.data());
  stage_clear(&h->stage);  // the arena is reused by later copies, queued after this segment's kernels
  return rc;
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
  GZ_HIP(hipMemcpyAsync(s->h_null.data(), (uint64_t*)s->tot.p + (int64_t)P * (1 + nbin),
                        sizeof(unsigned long long) * std::max(nout, 1), hipMemcpyDeviceToHost, ctx->stream));
  GZ_HIP(hipMemcpyAsync(&err, s->err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  if (err & 1) return fail(GZ_INVALID, "Partition id is equal or greater than num_partitions");
  if (err & 2) return fail(GZ_CAPACITY_ERROR, "a partition's utf8 data exceeds 2 GiB");
  if (err & 8) return fail(GZ_INVALID, "utf8 offsets reach past the declared data buffer");
  if (err & 4)
    return fail(GZ_NOT_IMPLEMENTED,
                "a float or double hash partitioning key holds -0.0 or NaN (its Spark hash is version-dependent)");
  s->resolved = true;
  return GZ_OK;
}

extern "C" {

int gz_split_batch(gz_splitter* h, int64_t num_rows, const int64_t* buf_addrs,
                   const int64_t* buf_sizes, int32_t num_bufs) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (!buf_addrs || !buf_sizes) return fail(GZ_INVALID, "Native split: buf_addrs can't be null");
  if (h->stopped) return fail(GZ_INVALID, "splitter already stopped");
  gz_batch b{num_rows, buf_addrs, buf_sizes, num_bufs};
  // The batch is staged for the batched path, except where a call's own
  // result is observable: the range splitter reports a bad partition id from
  // the call that holds it, and row ids count the rows of one call.
  if (h->mode != 2 && !h->opts.append_row_ids) {
    GZ_HIP(hipSetDevice(h->ctx->device));
    GZ_TRY(stage_check(h->ncols_in, h->types_in.data(), b));
    bool staged = false;
    GZ_TRY(stage_add(h->ctx, &h->stage, b, &staged));
    if (!staged && !h->stage.empty()) {  // the arena is full: run what it holds, then stage again
      GZ_TRY(split_flush(h));
      GZ_TRY(stage_add(h->ctx, &h->stage, b, &staged));
    }
    if (staged) return GZ_OK;
  }
  GZ_TRY(split_flush(h));
  return split_run(h, 1, &b);
}

int gz_split_batches(gz_splitter* h, int32_t num_batches, const gz_batch* batches) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (num_batches < 0 || (num_batches > 0 && !batches)) return fail(GZ_INVALID, "bad batch list");
  for (int i = 0; i < num_batches; i++)
    if (!batches[i].buf_addrs || !batches[i].buf_sizes)
      return fail(GZ_INVALID, "Native split: buf_addrs can't be null");
  if (h->stopped) return fail(GZ_INVALID, "splitter already stopped");
  GZ_TRY(split_flush(h));
  return split_run(h, num_batches, batches);
}

int gz_split_stop(gz_splitter* h) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (!h->stopped) GZ_TRY(split_flush(h));
  for (auto* s : h->segments) GZ_TRY(resolve_segment(h, s));
  drain_pid_events(h);
  h->stopped = true;
  return GZ_OK;
}

int gz_split_num_partitions(gz_splitter* h) { return h ? h->P : 0; }
int gz_split_num_output_columns(gz_splitter* h) { return h ? (int)h->out_types.size() : 0; }
int gz_split_num_segments(gz_splitter* h) {
  if (!h) return 0;
  if (!h->stopped && split_flush(h) != GZ_OK) return -1;
  return (int)h->segments.size();
}

int gz_split_partition_rows(gz_splitter* h, int64_t* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  if (!h->stopped) GZ_TRY(split_flush(h));
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
      if (h->cols[c].out_col == col) in_col = c;
    int j = in_col >= 0 ? h->cols[in_col].bin : -1;
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
  if (s->sorted) {
    GZ_TRY(sorted_batch_counts(h, s));
    GZ_HIP(hipMemcpyAsync(counts, s->counts.p, sizeof(int32_t) * (size_t)s->nbatches * P, hipMemcpyDeviceToHost,
                          h->ctx->stream));
    GZ_HIP(hipStreamSynchronize(h->ctx->stream));
    return GZ_OK;
  }
  // tile-major counters: row t holds the tile's rows per partition first
  std::vector<uint32_t> tc((size_t)s->R * std::max<int64_t>(s->T, 1));
  if (s->T > 0) {
    GZ_HIP(hipMemcpyAsync(tc.data(), s->counts.p, sizeof(uint32_t) * s->R * s->T,
                          hipMemcpyDeviceToHost, h->ctx->stream));
    GZ_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  memset(counts, 0, sizeof(int32_t) * (size_t)s->nbatches * P);
  for (int64_t t = 0; t < s->T; t++) {
    int b = s->tile_batch[t];
    for (int p = 0; p < P; p++) counts[(size_t)b * P + p] += (int32_t)tc[(size_t)t * s->R + p];
  }
  return GZ_OK;
}

int64_t gz_split_compute_pid_time(gz_splitter* h) { return h ? h->compute_pid_ns : 0; }

int gz_split_release(gz_splitter* h) {
  if (!h) return GZ_OK;
  hipSetDevice(h->ctx->device);
  // No synchronization: the pool hands a freed block only to work enqueued
  // later on the context's stream, which runs after everything queued here.
  for (auto* list : {&h->pid_events, &h->pid_free})
    for (auto& ev : *list) {
      hipEventDestroy(ev.first);
      hipEventDestroy(ev.second);
    }
  for (auto* s : h->segments) split_release_segment(h->ctx, s);
  stage_release(h->ctx, &h->stage);
  if (h->ipc) ipc_plan_free(h->ctx, h->ipc);
  for (DBuf* b : {&h->d_tiles, &h->d_bufs, &h->d_cols, &h->d_outs}) ctx_free(h->ctx, b);
  delete h;
  return GZ_OK;
}

}  // extern "C"

// ======================================================= partition writer (IPC)
// Splitter::Stop's data file (PartitionWriter, splitter.cc:70-185, :395-437) with
// the default compression_type UNCOMPRESSED (shuffle/type.h:44): partition after
// partition, each non-empty partition's Arrow IPC stream (the caller's schema
// message when write_schema, the record batches cut as CacheRecordBatch cuts them
// (:462-608, the cut rule of DoSplit :889-925), the end-of-stream marker
// 0xFFFFFFFF 0x00000000); an empty partition writes nothing (length 0).
//
// The record-batch metadata (an Arrow `Message` flatbuffer with a `RecordBatch`
// header) is built on the host; the bodies are assembled on the device from the
// split segments by one kernel: byte copies (fixed-width values, utf8 bytes, the
// metadata), bit-level concatenation (validity and boolean bitmaps, which start
// at arbitrary bit positions in the segments), and rebased int32 offsets.
// Buffers follow Arrow's IPC writer: a validity buffer only when the batch has
// nulls (else a zero-length buffer entry), every buffer padded to 8 bytes.
namespace gz {

enum IpcKind : int32_t { IPC_COPY = 0, IPC_BITS = 1, IPC_OFFS = 2 };

struct IpcPiece {
  const uint8_t* src;
  int64_t a;     // BITS: source bit; OFFS: unused
  int64_t b;     // BITS: first bit in the buffer; OFFS: first row in the buffer
  int64_t n;     // BITS: bits; OFFS: rows
  int64_t base;  // OFFS: bytes of the batch before this piece
};

struct IpcTask {
  int32_t kind;
  int32_t pad;
  int64_t p0, np;  // pieces [p0, p0 + np) (COPY: one)
  int64_t dst;     // output byte offset
  int64_t n;       // COPY: bytes; BITS: bits; OFFS: rows (n + 1 entries)
  int64_t out;     // bytes written including the zero padding to a multiple of 8
  int64_t total;   // OFFS: the last entry (bytes of the batch)
};

__device__ __forceinline__ uint32_t ipc_bits32(const uint8_t* src, int64_t bit, int cnt) {
  // cnt (1..32) bits of a bitmap from an arbitrary bit position, in the low
  // bits; only the bytes holding those bits are read (a bitmap may end at the
  // byte holding its last bit)
  const uint8_t* p = src + (bit >> 3);
  const int nbytes = ((int)(bit & 7) + cnt + 7) >> 3;  // 1..5
  uint64_t w = 0;
#pragma unroll
  for (int k = 0; k < 5; k++)
    if (k < nbytes) w |= (uint64_t)gld(p + k) << (8 * k);
  w >>= (bit & 7);
  return cnt < 32 ? (uint32_t)w & ((1u << cnt) - 1) : (uint32_t)w;
}

__global__ __launch_bounds__(256) void ipc_assemble_kernel(const IpcTask* tasks, int64_t ntasks,
                                                           const IpcPiece* pieces, uint8_t* out) {
  for (int64_t ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
    const IpcTask t = tasks[ti];
    uint8_t* dst = out + t.dst;
    if (t.kind == IPC_COPY) {
      const uint8_t* src = pieces[t.p0].src;
      // widest aligned unit for [0, n), then the tail bytes and the zero padding
      const uintptr_t al = ((uintptr_t)src | (uintptr_t)dst) & 15;
      int64_t done = 0;
      if (t.n > 0 && al == 0) {
        const int64_t nv = t.n >> 4;
        for (int64_t i = threadIdx.x; i < nv; i += blockDim.x) gst((u32x4*)dst + i, gld((const u32x4*)src + i));
        done = nv << 4;
      } else if (t.n > 0 && (al & 7) == 0) {
        const int64_t nv = t.n >> 3;
        for (int64_t i = threadIdx.x; i < nv; i += blockDim.x) gst((uint64_t*)dst + i, gld((const uint64_t*)src + i));
        done = nv << 3;
      } else if (t.n > 0 && (al & 3) == 0) {
        const int64_t nv = t.n >> 2;
        for (int64_t i = threadIdx.x; i < nv; i += blockDim.x) gst((uint32_t*)dst + i, gld((const uint32_t*)src + i));
        done = nv << 2;
      }
      for (int64_t i = done + threadIdx.x; i < t.out; i += blockDim.x) gst(dst + i, i < t.n ? gld(src + i) : (uint8_t)0);
    } else if (t.kind == IPC_BITS) {
      // output word w holds bits [32w, 32w + 32); pieces are ordered by b
      const IpcPiece* pc = pieces + t.p0;
      for (int64_t w = threadIdx.x; w * 4 < t.out; w += blockDim.x) {
        const int64_t lo = 32 * w, hi = lo + 32;
        uint32_t word = 0;
        if (lo < t.n) {
          int64_t l = 0, r = t.np - 1;  // last piece starting at or before lo
          while (l < r) {
            const int64_t m = (l + r + 1) >> 1;
            if (pc[m].b <= lo) l = m; else r = m - 1;
          }
          for (int64_t k = l; k < t.np && pc[k].b < hi; k++) {
            const int64_t s0 = max(lo, pc[k].b), s1 = min(hi, pc[k].b + pc[k].n);
            if (s1 <= s0) continue;
            const int cnt = (int)(s1 - s0);
            // a segment without a validity buffer: every row valid
            uint32_t v = pc[k].src ? ipc_bits32(pc[k].src, pc[k].a + (s0 - pc[k].b), cnt) : 0xffffffffu;
            if (cnt < 32) v &= (1u << cnt) - 1;
            word |= v << (s0 - lo);
          }
          if (t.n < hi) word &= (1u << (t.n - lo)) - 1;
        }
        // the buffer's length is a multiple of 8 bytes: whole words
        gst((uint32_t*)dst + w, word);
      }
    } else {
      const IpcPiece* pc = pieces + t.p0;
      for (int64_t j = threadIdx.x; j * 4 < t.out; j += blockDim.x) {
        int32_t v = 0;
        if (j == t.n) {
          v = (int32_t)t.total;
        } else if (j < t.n) {
          int64_t l = 0, r = t.np - 1;
          while (l < r) {
            const int64_t m = (l + r + 1) >> 1;
            if (pc[m].b <= j) l = m; else r = m - 1;
          }
          const int32_t* so = (const int32_t*)pc[l].src;
          v = (int32_t)(gld(so + (j - pc[l].b)) - gld(so) + pc[l].base);
        }
        gst((int32_t*)dst + j, v);
      }
    }
  }
}

// Zero bits (nulls) among n bits of a bitmap from bit a, per piece; grid-stride
// over pieces with a wave per piece.
__global__ __launch_bounds__(256) void ipc_null_count_kernel(const IpcPiece* pieces, int64_t np,
                                                             unsigned long long* out) {
  const int lane = __lane_id();
  const int64_t wv = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / kWave;
  for (int64_t i = wv; i < np; i += nw) {
    const IpcPiece pc = pieces[i];
    unsigned long long ones = 0;
    for (int64_t o = (int64_t)lane * 32; o < pc.n; o += 32 * kWave) {
      const int cnt = (int)min<int64_t>(32, pc.n - o);
      const uint32_t v = ipc_bits32(pc.src, pc.a + o, cnt);
      ones += __popc(v);
    }
    ones = wave_reduce_sum(ones);
    if (lane == 0) out[i] = (unsigned long long)pc.n - ones;
  }
}

// Reads int32 values at a list of device addresses (the utf8 offsets at both
// ends of every span of a partition writer batch) in one launch.
__global__ void ipc_gather_i32_kernel(const int32_t* const* addr, int64_t n, int32_t* out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = gld(addr[i]);
}

// A flatbuffer built back to front (the layout the format's own builders
// produce): objects are prepended, an object is referred to by its distance
// from the end of the buffer, and every offset points forward.
class FlatBuilder {
 public:
  size_t size() const { return b_.size(); }
  void align(size_t a) {
    while (b_.size() % a) prepend_bytes("\0", 1);
  }
  void prepend_bytes(const void* p, size_t n) { b_.insert(b_.begin(), (const uint8_t*)p, (const uint8_t*)p + n); }
  template <typename T>
  void scalar(T v) {
    align(sizeof(T));
    prepend_bytes(&v, sizeof(T));
  }
  // uoffset to an object referred to by ref; returns the field's ref
  size_t offset_to(size_t ref) {
    align(4);
    const uint32_t v = (uint32_t)(b_.size() + 4 - ref);
    prepend_bytes(&v, 4);
    return b_.size();
  }
  // vector of 16-byte structs of two int64 (FieldNode, Buffer); returns its ref
  size_t vector_of_pairs(const std::vector<std::pair<int64_t, int64_t>>& v) {
    // elements 8-aligned and preceded by the uint32 length
    const size_t body = 16 * v.size();
    while ((b_.size() + body) % 8) prepend_bytes("\0", 1);
    for (size_t i = v.size(); i-- > 0;) {
      int64_t s[2] = {v[i].first, v[i].second};
      prepend_bytes(s, 16);
    }
    const uint32_t n = (uint32_t)v.size();
    prepend_bytes(&n, 4);
    return b_.size();
  }
  // tables: start, add fields (in any order), end
  void start_table() {
    fields_.clear();
    table_end_ = b_.size();
  }
  template <typename T>
  void field(int id, T v) {
    scalar(v);
    fields_.push_back({id, b_.size()});
  }
  void field_offset(int id, size_t ref) { fields_.push_back({id, offset_to(ref)}); }
  size_t end_table() {
    align(4);
    int32_t placeholder = 0;
    prepend_bytes(&placeholder, 4);
    const size_t table_ref = b_.size();
    int nf = 0;
    for (auto& f : fields_) nf = std::max(nf, f.first + 1);
    std::vector<uint16_t> vt(2 + nf, 0);
    vt[0] = (uint16_t)(2 * (2 + nf));
    vt[1] = (uint16_t)(table_ref - table_end_);
    for (auto& f : fields_) vt[2 + f.first] = (uint16_t)(table_ref - f.second);
    for (size_t i = vt.size(); i-- > 0;) prepend_bytes(&vt[i], 2);
    const size_t vt_ref = b_.size();
    const int32_t soff = (int32_t)(vt_ref - table_ref);  // vtable precedes the table
    memcpy(&b_[b_.size() - table_ref], &soff, 4);
    return table_ref;
  }
  std::vector<uint8_t> finish(size_t root) {
    // the root offset sits at the start of a buffer whose size is a multiple of 8
    while ((b_.size() + 4) % 8) prepend_bytes("\0", 1);
    offset_to(root);
    return b_;
  }

 private:
  std::vector<uint8_t> b_;
  std::vector<std::pair<int, size_t>> fields_;
  size_t table_end_ = 0;
};

// Encapsulated IPC message of a record batch: continuation token, metadata
// length, Message flatbuffer (version V5, header RecordBatch) padded so the body
// starts on a multiple of 8.
static std::vector<uint8_t> ipc_record_batch_message(int64_t rows,
                                                     const std::vector<std::pair<int64_t, int64_t>>& nodes,
                                                     const std::vector<std::pair<int64_t, int64_t>>& buffers,
                                                     int64_t body_len, bool lz4 = false) {
  FlatBuilder fb;
  size_t comp = 0;
  if (lz4) {
    fb.start_table();          // BodyCompression
    fb.field<int8_t>(1, 0);    // BodyCompressionMethod.BUFFER
    fb.field<int8_t>(0, 0);    // CompressionType.LZ4_FRAME
    comp = fb.end_table();
  }
  const size_t vbuf = fb.vector_of_pairs(buffers);
  const size_t vnode = fb.vector_of_pairs(nodes);
  fb.start_table();  // RecordBatch
  fb.field<int64_t>(0, rows);
  fb.field_offset(1, vnode);
  fb.field_offset(2, vbuf);
  if (lz4) fb.field_offset(3, comp);
  const size_t rb = fb.end_table();
  fb.start_table();  // Message
  fb.field<int64_t>(3, body_len);
  fb.field_offset(2, rb);
  fb.field<uint8_t>(1, 3);   // MessageHeader.RecordBatch
  fb.field<int16_t>(0, 4);   // MetadataVersion.V5
  const size_t msg = fb.end_table();
  std::vector<uint8_t> meta = fb.finish(msg);
  while ((meta.size() + 8) % 8) meta.push_back(0);
  std::vector<uint8_t> out(8);
  const int32_t cont = -1, len = (int32_t)meta.size();
  memcpy(out.data(), &cont, 4);
  memcpy(out.data() + 4, &len, 4);
  out.insert(out.end(), meta.begin(), meta.end());
  return out;
}

static int64_t pad8(int64_t n) { return (n + 7) / 8 * 8; }

// ---- LZ4 block encoder (partition writer with LZ4_FRAME bodies) ---------------
// The published LZ4 block format: sequences of a token (literal length : 4,
// match length - 4 : 4), length extensions (255-runs), literals, a 16-bit
// offset; a match starts at least 12 bytes before the block end, and the last
// 5 bytes are literals.  Frames are written with independent blocks (32 KiB,
// under the 64 KiB maximum the frame declares), so every block compresses on its
// own.
struct Lz4CBlock {
  const uint8_t* src;
  uint8_t* dst;  // len bytes of room: a block that does not shrink is stored
  uint32_t len;
  uint32_t pad;
};

constexpr int kLz4HashLog = 12;
constexpr uint32_t kLz4Block = 32768;  // block size: four workgroups' LDS per CU
constexpr uint32_t kLz4MaxExt = 4096;  // longest match taken at once (longer runs continue in the next sequence)

__device__ __forceinline__ uint32_t lds_rd32(const uint32_t* w, uint32_t p) {
  return __builtin_amdgcn_alignbyte(w[(p >> 2) + 1], w[p >> 2], p & 3);
}
__device__ __forceinline__ uint32_t lds_byte(const uint32_t* w, uint32_t p) { return (w[p >> 2] >> ((p & 3) * 8)) & 0xFFu; }

// One wave per block: the block is staged in LDS; 64 positions at a time
// every lane hashes its position's 4 bytes, takes the table's earlier position
// as its match candidate and extends the match; then the wave walks the 64
// candidates in order and emits a sequence at the first match at or after the
// end of the previous one (greedy, as the reference encoder's fast mode).
// csize[b] = compressed bytes, or 0 when the block is stored.
__global__ __launch_bounds__(64) void lz4_compress_kernel(const Lz4CBlock* __restrict__ blocks, int nblocks,
                                                          uint32_t* __restrict__ csize) {
  extern __shared__ uint32_t s_lz[];
  const int b = blockIdx.x;
  if (b >= nblocks) return;
  const int lane = threadIdx.x;
  const Lz4CBlock bl = blocks[b];
  const uint32_t len = bl.len;
  uint32_t* s_in = s_lz;                                 // kLz4Block + 8 bytes
  uint16_t* s_ht = (uint16_t*)(s_lz + (kLz4Block + 8) / 4);  // 2^kLz4HashLog positions
  // stage: aligned source dwords realigned into LDS (each dword loaded holds a byte of the block)
  {
    const uint64_t a = (uint64_t)bl.src;
    const uint32_t m = (uint32_t)(a & 3);
    const uint32_t* sa = (const uint32_t*)(a - m);
    const uint32_t nw = (len + 3) / 4;
    for (uint32_t i = lane; i < nw; i += kWave) {
      const uint32_t w0 = gld(sa + i);
      const uint32_t w1 = (m && (i + 1) * 4 < len + m) ? gld(sa + i + 1) : 0u;
      s_in[i] = __builtin_amdgcn_alignbyte(w1, w0, m);
    }
    for (uint32_t i = nw + lane; i < nw + 2; i += kWave) s_in[i] = 0;
    for (int i = lane; i < (1 << kLz4HashLog); i += kWave) s_ht[i] = 0;
  }
  __syncthreads();
  if (len < 13) {  // no room for a match: stored
    if (lane == 0) csize[b] = 0;
    return;
  }
  const uint32_t mflimit = len - 12;     // last position a match may start at
  const uint32_t matchlimit = len - 5;   // a match ends at or before this
  const uint32_t cap = len;              // compressed output must stay below the input size
  uint8_t* out = bl.dst;
  uint32_t anchor = 0, cur = 0, op = 0;
  bool fail = false;
  // writes a length extension of value v (v >= 0 after the 15 in the nibble) at op
  auto ext = [&](uint32_t v) {
    const uint32_t cnt = v / 255 + 1;
    for (uint32_t k = lane; k < cnt; k += kWave) gst(out + op + k, (uint8_t)(k + 1 < cnt ? 255u : v % 255u));
    op += cnt;
  };
  auto literals = [&](uint32_t from, uint32_t n) {
    for (uint32_t k = lane; k < n; k += kWave) gst(out + op + k, (uint8_t)lds_byte(s_in, from + k));
    op += n;
  };
  for (uint32_t base = 0; base <= mflimit && !fail; base += kWave) {
    const uint32_t p = base + lane;
    const bool active = p <= mflimit;
    const uint32_t v = active ? lds_rd32(s_in, p) : 0u;
    const uint32_t h = (v * 2654435761u) >> (32 - kLz4HashLog);
    const uint32_t cand = active ? s_ht[h] : 0u;
    __builtin_amdgcn_wave_barrier();  // every lane reads the table before any lane writes it
    if (active) s_ht[h] = (uint16_t)p;
    uint32_t mlen = 0;
    if (active && p >= cur && cand < p && lds_rd32(s_in, cand) == v) {
      const uint32_t lim = min(matchlimit - p, kLz4MaxExt);
      uint32_t l = 4;
      while (l + 4 <= lim) {
        const uint32_t x = lds_rd32(s_in, cand + l) ^ lds_rd32(s_in, p + l);
        if (x) {
          l += (uint32_t)__builtin_ctz(x) >> 3;
          break;
        }
        l += 4;
      }
      if (l + 4 > lim)
        while (l < lim && lds_byte(s_in, cand + l) == lds_byte(s_in, p + l)) l++;
      mlen = l >= 4 && l <= lim ? l : 0;
    }
    for (uint64_t has = __ballot(mlen >= 4); has; has &= has - 1) {
      const int l = __builtin_ctzll(has);
      const uint32_t q = base + (uint32_t)l;
      if (q < cur) continue;
      // lane l's match, read into scalar registers (v_readlane, no LDS round trip)
      const uint32_t ml = (uint32_t)__builtin_amdgcn_readlane((int)mlen, l);
      const uint32_t off = q - (uint32_t)__builtin_amdgcn_readlane((int)cand, l);
      const uint32_t lit = q - anchor;
      const uint32_t need = 1 + (lit >= 15 ? (lit - 15) / 255 + 1 : 0) + lit + 2 +
                            (ml - 4 >= 15 ? (ml - 4 - 15) / 255 + 1 : 0);
      if (op + need >= cap) {
        fail = true;
        break;
      }
      if (lane == 0) gst(out + op, (uint8_t)((min(lit, 15u) << 4) | min(ml - 4, 15u)));
      op++;
      if (lit >= 15) ext(lit - 15);
      literals(anchor, lit);
      if (lane == 0) {
        gst(out + op, (uint8_t)(off & 0xFF));
        gst(out + op + 1, (uint8_t)(off >> 8));
      }
      op += 2;
      if (ml - 4 >= 15) ext(ml - 4 - 15);
      anchor = cur = q + ml;
    }
  }
  if (!fail) {  // last literals
    const uint32_t lit = len - anchor;
    const uint32_t need = 1 + (lit >= 15 ? (lit - 15) / 255 + 1 : 0) + lit;
    if (op + need >= cap) {
      fail = true;
    } else {
      if (lane == 0) gst(out + op, (uint8_t)(min(lit, 15u) << 4));
      op++;
      if (lit >= 15) ext(lit - 15);
      literals(anchor, lit);
    }
  }
  if (lane == 0) csize[b] = fail ? 0u : op;
}

}  // namespace gz

struct gz_ipc_plan {
  int64_t total = 0;
  std::vector<int64_t> part_len;
  std::vector<IpcTask> tasks;
  std::vector<IpcPiece> pieces;
  std::vector<uint8_t> blob;  // metadata and schema bytes, copied by COPY tasks
  DBuf d_tasks, d_pieces, d_blob;
  // record batches of the uncompressed layout (what compression rewrites)
  struct BatchRec {
    int p;
    int64_t rows;
    std::vector<std::pair<int64_t, int64_t>> nodes, bufs;
    int64_t body_pos;
  };
  std::vector<BatchRec> recs;
  std::vector<uint8_t> schema;  // written first in every non-empty partition, when non-empty
  DBuf raw, comp, d_cblocks, d_csize;  // LZ4: uncompressed stream, compressed blocks
  std::vector<int64_t> raw_len;        // RawPartitionLengths: in-memory bytes of the cached batches
  int64_t batch_rows = 0;              // rows of a partition buffer (CalculateSplitBatchSize)
};

static void ipc_plan_free(gz_context* ctx, gz_ipc_plan* pl) {
  if (!pl) return;
  for (DBuf* b : {&pl->d_tasks, &pl->d_pieces, &pl->d_blob, &pl->raw, &pl->comp, &pl->d_cblocks, &pl->d_csize})
    ctx_free(ctx, b);
  delete pl;
}

// The reference's partition buffers for one partition over its input batches
// (Splitter::DoSplit, splitter.cc:876-938): a buffer holds max(batch_rows, rows
// of this input batch) rows; when the next input batch's rows do not fit, the
// filled buffer is cached as one record batch (CacheRecordBatch, :462-608) and
// is reused (prefer_spill and the new size fits) or reset and allocated again.
// Every cached batch also records which columns carry a validity buffer in
// memory (for RawPartitionLengths): a buffer allocated while the column's
// sticky input_has_null_ flag is set gets one (AllocatePartitionBuffers,
// :610-713), an input batch with nulls in the column gives one to every
// buffer it adds rows to (SplitValidityBuffer, :1179-1207), and a reused
// buffer keeps it.  The sticky flag of array slot k is set by an input batch
// with nulls in schema column array_idx[k] only when that column index is below
// the fixed-width column count (:878-886: the index compared is the schema's,
// as written there).
struct CutBatch {
  int64_t rows;
  int64_t last_batch;          // global index of the last input batch it holds
  std::vector<uint8_t> valid;  // per output column: a validity buffer in memory
};

static std::vector<CutBatch> ipc_cut(const std::vector<int32_t>& counts, int64_t bs, bool prefer_spill,
                                     const std::vector<uint8_t>& sticky,   // [batch][slot], after the batch's update
                                     const std::vector<uint8_t>& nulls,    // [batch][out col]
                                     const std::vector<int>& slot_col, int nout) {
  std::vector<CutBatch> out;
  const int ns = (int)slot_col.size();
  int64_t size = 0, base = 0, last = -1;
  std::vector<uint8_t> v(nout, 0);
  auto allocate = [&](int64_t b) {
    for (int k = 0; k < ns; k++) v[slot_col[k]] = sticky[(size_t)b * ns + k];
  };
  for (size_t b = 0; b < counts.size(); b++) {
    const int32_t cnt = counts[b];
    if (cnt <= 0) continue;
    const int64_t new_size = std::max<int64_t>(bs, cnt);
    if (size == 0) {
      size = new_size;
      allocate((int64_t)b);
    } else if (base + cnt > size) {
      if (base > 0) out.push_back({base, last, v});
      base = 0;
      if (!prefer_spill || new_size > size) {  // reset, allocate again
        size = new_size;
        allocate((int64_t)b);
      }
    }
    for (int k = 0; k < ns; k++)
      if (nulls[b * nout + slot_col[k]]) v[slot_col[k]] = 1;
    base += cnt;
    last = (int64_t)b;
  }
  if (base > 0) out.push_back({base, last, v});
  return out;
}

// Rows per (input batch, partition), nulls per (input batch, output column) and
// utf8 bytes per (input batch, binary column) of a segment, from its tile
// counter rows.
static int segment_batch_stats(gz_splitter* h, SplitSegment* s, std::vector<int32_t>& counts,
                               std::vector<uint8_t>& nulls, std::vector<int64_t>& bin_bytes) {
  const int P = h->P, nbin = h->base.nbin, nout = (int)h->out_types.size();
  counts.assign((size_t)s->nbatches * P, 0);
  nulls.assign((size_t)s->nbatches * std::max(nout, 1), 0);
  bin_bytes.assign((size_t)s->nbatches * std::max(nbin, 1), 0);
  if (s->sorted) {
    GZ_TRY(sorted_batch_counts(h, s));
    std::vector<unsigned long long> st((size_t)s->nbatches * (nbin + nout) + 1);
    if (s->nbatches > 0) {
      GZ_HIP(hipMemcpyAsync(counts.data(), s->counts.p, sizeof(int32_t) * (size_t)s->nbatches * P,
                            hipMemcpyDeviceToHost, h->ctx->stream));
      GZ_HIP(hipMemcpyAsync(st.data(), s->bstats.p, sizeof(unsigned long long) * (size_t)s->nbatches * (nbin + nout),
                            hipMemcpyDeviceToHost, h->ctx->stream));
      GZ_HIP(hipStreamSynchronize(h->ctx->stream));
    }
    for (int32_t b = 0; b < s->nbatches; b++) {
      const unsigned long long* row = &st[(size_t)b * (nbin + nout)];
      for (int j = 0; j < nbin; j++) bin_bytes[(size_t)b * nbin + j] = (int64_t)row[j];
      for (int oc = 0; oc < nout; oc++)
        if (row[nbin + oc]) nulls[(size_t)b * nout + oc] = 1;
    }
    return GZ_OK;
  }
  std::vector<uint32_t> tc((size_t)s->R * std::max<int64_t>(s->T, 1));
  if (s->T > 0) {
    GZ_HIP(hipMemcpyAsync(tc.data(), s->counts.p, sizeof(uint32_t) * s->R * s->T, hipMemcpyDeviceToHost,
                          h->ctx->stream));
    GZ_HIP(hipStreamSynchronize(h->ctx->stream));
  }
  for (int64_t t = 0; t < s->T; t++) {
    const int b = s->tile_batch[t];
    const uint32_t* row = &tc[(size_t)t * s->R];
    for (int p = 0; p < P; p++) counts[(size_t)b * P + p] += (int32_t)row[p];
    for (int j = 0; j < nbin; j++)
      for (int p = 0; p < P; p++) bin_bytes[(size_t)b * nbin + j] += row[(size_t)P * (1 + j) + p];
    for (int oc = 0; oc < nout; oc++)
      if (row[(size_t)P * (1 + nbin) + oc]) nulls[(size_t)b * nout + oc] = 1;
  }
  return GZ_OK;
}

static int ipc_prepare(gz_splitter* h, const uint8_t* schema, int64_t schema_len, bool write_schema,
                       gz_ipc_plan* plan) {
  gz_context* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  const int P = h->P;
  const int nout = (int)h->out_types.size();
  const int nbin = h->base.nbin;
  if (!h->stopped) GZ_TRY(split_flush(h));
  const int nseg = (int)h->segments.size();
  for (auto* s : h->segments) GZ_TRY(resolve_segment(h, s));
  // rows per input batch per partition, segment by segment (counts[s][b * P + p]),
  // and per input batch (all segments in order) its rows, nulls per output
  // column and utf8 bytes per binary column
  std::vector<std::vector<int32_t>> cnt(nseg);
  std::vector<uint8_t> gnull;
  std::vector<int64_t> gbin, grows;
  for (int s = 0; s < nseg; s++) {
    std::vector<uint8_t> nl;
    std::vector<int64_t> bb;
    GZ_TRY(segment_batch_stats(h, h->segments[s], cnt[s], nl, bb));
    const int nb = h->segments[s]->nbatches;
    for (int b = 0; b < nb; b++) {
      int64_t r = 0;
      for (int p = 0; p < P; p++) r += cnt[s][(size_t)b * P + p];
      grows.push_back(r);
      gnull.insert(gnull.end(), nl.begin() + (size_t)b * std::max(nout, 1),
                   nl.begin() + (size_t)b * std::max(nout, 1) + nout);
      for (int j = 0; j < nbin; j++) gbin.push_back(bb[(size_t)b * nbin + j]);
    }
  }
  const int64_t NB = (int64_t)grows.size();
  // the reference's array slots: fixed-width columns in schema order, then the
  // binary ones (null-typed columns and the row-id column have none)
  std::vector<int> slot_col;
  for (int oc = 0; oc < nout; oc++)
    if (oc != h->base.rowid_oc && h->out_types[oc] != GZ_TYPE_NA && !type_is_binary(h->out_types[oc]))
      slot_col.push_back(oc);
  const int fixed_cnt = (int)slot_col.size();
  for (int oc = 0; oc < nout; oc++)
    if (oc != h->base.rowid_oc && type_is_binary(h->out_types[oc])) slot_col.push_back(oc);
  const int ns = (int)slot_col.size();
  std::vector<uint8_t> sticky((size_t)std::max<int64_t>(NB, 1) * std::max(ns, 1), 0);
  {
    std::vector<uint8_t> run(ns, 0);
    for (int64_t b = 0; b < NB; b++)
      for (int k = 0; k < ns; k++) {
        if (slot_col[k] < fixed_cnt && gnull[(size_t)b * nout + slot_col[k]]) run[k] = 1;
        sticky[(size_t)b * ns + k] = run[k];
      }
  }
  // rows of a partition buffer: CalculateSplitBatchSize (splitter.cc:800-853),
  // the row size from the first input batch (binary columns: ceil(bytes / rows))
  int64_t bs = h->opts.buffer_size > 0 ? h->opts.buffer_size : 4096;
  if (h->opts.offheap_per_task > 0) {
    int64_t b0 = 0;
    while (b0 < NB && grows[b0] == 0) b0++;
    if (b0 < NB) {
      int64_t spr = 0;
      for (int j = 0; j < nbin; j++) {
        const int64_t bytes = gbin[(size_t)b0 * nbin + j];
        if (bytes == 0)
          return fail(GZ_NOT_IMPLEMENTED, "offheap_per_task sizing with a binary column empty in the first batch");
        spr += (bytes + grows[b0] - 1) / grows[b0];
      }
      for (int k = 0; k < ns; k++)
        if (slot_col[k] < fixed_cnt) {
          const int t = h->out_types[slot_col[k]];
          spr += type_is_binary(t) || t == GZ_TYPE_BOOL ? 0 : type_width(t);  // arrow::bit_width(t) >> 3
        }
      const int64_t pre = spr > 0 ? (h->opts.offheap_per_task / spr / P) >> 2 : bs;
      bs = std::min(pre, bs);
    }
  }
  plan->batch_rows = bs;
  // input column of each output column (range mode drops the pid column)
  std::vector<int> in_of(nout, -1);
  for (int c = 0; c < h->ncols_in; c++)
    if (h->cols[c].out_col >= 0) in_of[h->cols[c].out_col] = c;
  // ---- pass 1: batches, their pieces per segment, validity pieces for null counts
  struct Span {
    int seg;
    int64_t row0;  // partition-local row in the segment
    int64_t rows;
  };
  struct Batch {
    int p;
    int64_t rows;
    std::vector<Span> spans;
    std::vector<uint8_t> valid;  // per output column: a validity buffer in the reference's memory
  };
  std::vector<Batch> batches;
  for (int p = 0; p < P; p++) {
    std::vector<int32_t> seq;
    std::vector<std::pair<int, int64_t>> seg_rows;  // (segment, rows of p in it)
    for (int s = 0; s < nseg; s++) {
      int64_t r = 0;
      for (int32_t b = 0; b < h->segments[s]->nbatches; b++) {
        seq.push_back(cnt[s][(size_t)b * P + p]);
        r += cnt[s][(size_t)b * P + p];
      }
      seg_rows.push_back({s, r});
    }
    const std::vector<CutBatch> cuts = ipc_cut(seq, bs, h->opts.prefer_spill != 0, sticky, gnull, slot_col, nout);
    size_t si = 0;
    int64_t used = 0;  // rows of segment seg_rows[si] already taken
    for (const CutBatch& cb : cuts) {
      const int64_t n = cb.rows;
      Batch bt;
      bt.p = p;
      bt.rows = n;
      bt.valid = cb.valid;
      int64_t left = n;
      while (left > 0) {
        while (si < seg_rows.size() && used >= seg_rows[si].second) {
          si++;
          used = 0;
        }
        if (si >= seg_rows.size()) return fail(GZ_UNKNOWN_ERROR, "partition writer: row accounting");
        const int64_t take = std::min(left, seg_rows[si].second - used);
        bt.spans.push_back({seg_rows[si].first, used, take});
        used += take;
        left -= take;
      }
      batches.push_back(std::move(bt));
    }
  }
  // validity pieces of every (batch, column with validity): null counts on the device
  std::vector<IpcPiece> vp;
  std::vector<int64_t> vp_first((size_t)batches.size() * nout + 1, 0);
  for (size_t bi = 0; bi < batches.size(); bi++) {
    const Batch& bt = batches[bi];
    for (int c = 0; c < nout; c++) {
      vp_first[bi * nout + c] = (int64_t)vp.size();
      if (h->out_types[c] == GZ_TYPE_NA) continue;
      int64_t dstbit = 0;
      for (const Span& sp : bt.spans) {
        SplitSegment* sg = h->segments[sp.seg];
        if (sg->validity[c].p)
          vp.push_back({(const uint8_t*)sg->validity[c].p, sg->h_row_start[bt.p] + sp.row0, dstbit, sp.rows, 0});
        dstbit += sp.rows;
      }
    }
  }
  vp_first[batches.size() * nout] = (int64_t)vp.size();
  std::vector<unsigned long long> vnull(std::max<size_t>(vp.size(), 1), 0);
  if (!vp.empty()) {
    DBuf dvp, dnull;
    GZ_TRY(ctx_alloc(ctx, sizeof(IpcPiece) * vp.size(), &dvp));
    GZ_TRY(ctx_alloc(ctx, sizeof(unsigned long long) * vp.size(), &dnull));
    GZ_HIP(hipMemcpyAsync(dvp.p, vp.data(), sizeof(IpcPiece) * vp.size(), hipMemcpyHostToDevice, st));
    const unsigned g = (unsigned)std::min<size_t>((vp.size() + 3) / 4, 65536);
    hipLaunchKernelGGL(ipc_null_count_kernel, dim3(g), dim3(256), 0, st, (const IpcPiece*)dvp.p,
                       (int64_t)vp.size(), (unsigned long long*)dnull.p);
    GZ_HIP(hipMemcpyAsync(vnull.data(), dnull.p, sizeof(unsigned long long) * vp.size(), hipMemcpyDeviceToHost,
                          st));
    GZ_HIP(hipStreamSynchronize(st));
    ctx_free(ctx, &dvp);
    ctx_free(ctx, &dnull);
  }
  // utf8 offsets at both ends of every (batch, binary column, span), gathered at once
  std::vector<const int32_t*> oaddr;
  for (size_t b2 = 0; b2 < batches.size(); b2++) {
    const Batch& bt = batches[b2];
    for (int c = 0; c < nout; c++) {
      if (!type_is_binary(h->out_types[c])) continue;
      for (const Span& sp : bt.spans) {
        SplitSegment* sg = h->segments[sp.seg];
        const int32_t* o = (const int32_t*)sg->offsets[c].p + sg->h_row_start[bt.p] + bt.p + sp.row0;
        oaddr.push_back(o);
        oaddr.push_back(o + sp.rows);
      }
    }
  }
  std::vector<int32_t> oval(std::max<size_t>(oaddr.size(), 1), 0);
  if (!oaddr.empty()) {
    DBuf da, dv;
    GZ_TRY(ctx_alloc(ctx, sizeof(int32_t*) * oaddr.size(), &da));
    GZ_TRY(ctx_alloc(ctx, sizeof(int32_t) * oaddr.size(), &dv));
    GZ_HIP(hipMemcpyAsync(da.p, oaddr.data(), sizeof(int32_t*) * oaddr.size(), hipMemcpyHostToDevice, st));
    const unsigned g = (unsigned)std::min<size_t>((oaddr.size() + 255) / 256, 8192);
    hipLaunchKernelGGL(ipc_gather_i32_kernel, dim3(g), dim3(256), 0, st, (const int32_t* const*)da.p,
                       (int64_t)oaddr.size(), (int32_t*)dv.p);
    GZ_HIP(hipMemcpyAsync(oval.data(), dv.p, sizeof(int32_t) * oaddr.size(), hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    ctx_free(ctx, &da);
    ctx_free(ctx, &dv);
  }
  size_t ov = 0;  // next pair of oval
  // ---- pass 2: layout of the stream, metadata, tasks
  plan->part_len.assign(P, 0);
  plan->raw_len.assign(P, 0);
  plan->recs.clear();
  plan->schema.assign(write_schema && schema_len > 0 ? schema : nullptr,
                      write_schema && schema_len > 0 ? schema + schema_len : nullptr);
  plan->tasks.clear();
  plan->pieces.clear();
  plan->blob.clear();
  int64_t pos = 0;
  auto add_copy = [&](const uint8_t* src_host_blob_or_dev, bool from_blob, int64_t nbytes, int64_t out_bytes) {
    IpcTask t{};
    t.kind = IPC_COPY;
    t.p0 = (int64_t)plan->pieces.size();
    t.np = 1;
    t.dst = pos;
    t.n = nbytes;
    t.out = out_bytes;
    IpcPiece pc{};
    if (from_blob) {
      pc.src = nullptr;
      pc.a = (int64_t)plan->blob.size();  // resolved to a device address after the blob upload
      plan->blob.insert(plan->blob.end(), src_host_blob_or_dev, src_host_blob_or_dev + nbytes);
      t.pad = 1;
    } else {
      pc.src = src_host_blob_or_dev;
    }
    plan->pieces.push_back(pc);
    plan->tasks.push_back(t);
    pos += out_bytes;
  };
  size_t bi = 0;
  for (int p = 0; p < P; p++) {
    const int64_t start = pos;
    bool any = false;
    while (bi < batches.size() && batches[bi].p == p) {
      const Batch& bt = batches[bi];
      if (!any && write_schema && schema_len > 0) add_copy(schema, true, schema_len, schema_len);
      any = true;
      // field nodes and buffers (offsets relative to the body start)
      std::vector<std::pair<int64_t, int64_t>> nodes, bufs;
      struct Pending {
        int kind;
        int c;
        int64_t off, n, out;
      };
      std::vector<Pending> body;
      int64_t boff = 0;
      auto add_buf = [&](int kind, int c, int64_t n, int64_t bytes) {
        bufs.push_back({boff, bytes});
        body.push_back({kind, c, boff, n, pad8(bytes)});
        boff += pad8(bytes);
      };
      for (int c = 0; c < nout; c++) {
        const int t = h->out_types[c];
        if (t == GZ_TYPE_NA) {
          nodes.push_back({bt.rows, bt.rows});
          continue;
        }
        unsigned long long nulls = 0;
        for (int64_t k = vp_first[bi * nout + c]; k < vp_first[bi * nout + c + 1]; k++) nulls += vnull[k];
        nodes.push_back({bt.rows, (int64_t)nulls});
        if (nulls > 0) add_buf(IPC_BITS, c, bt.rows, (bt.rows + 7) / 8);
        else bufs.push_back({boff, 0});
        if (t == GZ_TYPE_BOOL) {
          add_buf(IPC_BITS, c + (1 << 16), bt.rows, (bt.rows + 7) / 8);
        } else if (type_is_binary(t)) {
          add_buf(IPC_OFFS, c, bt.rows, (bt.rows + 1) * 4);
          // bytes of the batch: from the segments' offsets (read on the host)
          add_buf(IPC_COPY, c, 0, 0);  // sized below
        } else {
          add_buf(IPC_COPY, c, bt.rows * type_width(t), bt.rows * type_width(t));
        }
      }
      // utf8 byte sizes per span: offsets at the span ends (host copies)
      std::vector<int64_t> span_bytes;  // per (binary column, span)
      for (auto& pd : body) {
        if (pd.kind != IPC_COPY || !type_is_binary(h->out_types[pd.c])) continue;
        int64_t total = 0;
        for (size_t k = 0; k < bt.spans.size(); k++, ov += 2) {
          span_bytes.push_back(oval[ov]);
          span_bytes.push_back((int64_t)oval[ov + 1] - oval[ov]);
          total += (int64_t)oval[ov + 1] - oval[ov];
        }
        pd.n = total;
      }
      // the cached batch's in-memory bytes (batch_nbytes, splitter.cc:438-453)
      {
        const int64_t n = bt.rows, bits = (n + 7) / 8;
        for (int oc : slot_col) {
          const int t = h->out_types[oc];
          plan->raw_len[p] += bt.valid[oc] ? bits : 0;
          if (t == GZ_TYPE_BOOL) {
            plan->raw_len[p] += bits;
          } else if (type_is_binary(t)) {
            int64_t bytes = 0;
            for (auto& pd : body)
              if (pd.kind == IPC_COPY && pd.c == oc) bytes = pd.n;
            plan->raw_len[p] += (n + 1) * 4 + bytes;
          } else {
            plan->raw_len[p] += n * type_width(t);
          }
        }
      }
      // recompute the buffer table and offsets now that utf8 sizes are known
      bufs.clear();
      boff = 0;
      {
        size_t k = 0;
        for (int c = 0; c < nout; c++) {
          const int t = h->out_types[c];
          if (t == GZ_TYPE_NA) continue;
          if (nodes[c].second > 0) {
            bufs.push_back({boff, (bt.rows + 7) / 8});
            body[k].off = boff;
            body[k].out = pad8((bt.rows + 7) / 8);
            boff += body[k].out;
            k++;
          } else {
            bufs.push_back({boff, 0});
          }
          const int nb = t == GZ_TYPE_BOOL ? 1 : type_is_binary(t) ? 2 : 1;
          for (int j = 0; j < nb; j++, k++) {
            Pending& pd = body[k];
            const int64_t bytes = pd.kind == IPC_BITS ? (bt.rows + 7) / 8
                                  : pd.kind == IPC_OFFS ? (bt.rows + 1) * 4
                                                        : pd.n;
            bufs.push_back({boff, bytes});
            pd.off = boff;
            pd.out = pad8(bytes);
            boff += pd.out;
          }
        }
      }
      const std::vector<uint8_t> msg = ipc_record_batch_message(bt.rows, nodes, bufs, boff);
      add_copy(msg.data(), true, (int64_t)msg.size(), (int64_t)msg.size());
      plan->recs.push_back({p, bt.rows, nodes, bufs, pos});
      // body tasks
      const int64_t body0 = pos;
      size_t sb = 0;
      for (auto& pd : body) {
        const int c = pd.c & 0xffff;
        const bool values_bits = (pd.c >> 16) != 0;
        IpcTask t{};
        t.kind = pd.kind;
        t.dst = body0 + pd.off;
        t.out = pd.out;
        t.p0 = (int64_t)plan->pieces.size();
        if (pd.kind == IPC_BITS) {
          t.n = bt.rows;
          int64_t dstbit = 0;
          for (const Span& sp : bt.spans) {
            SplitSegment* sg = h->segments[sp.seg];
            const uint8_t* bm = (const uint8_t*)(values_bits ? sg->values[c].p : sg->validity[c].p);
            plan->pieces.push_back({bm, sg->h_row_start[bt.p] + sp.row0, dstbit, sp.rows, 0});
            dstbit += sp.rows;
          }
        } else if (pd.kind == IPC_OFFS) {
          t.n = bt.rows;
          int64_t dstrow = 0, bytes = 0;
          size_t sbk = sb;
          for (const Span& sp : bt.spans) {
            SplitSegment* sg = h->segments[sp.seg];
            const int32_t* o = (const int32_t*)sg->offsets[c].p + sg->h_row_start[bt.p] + bt.p + sp.row0;
            plan->pieces.push_back({(const uint8_t*)o, 0, dstrow, sp.rows, bytes});
            dstrow += sp.rows;
            bytes += span_bytes[sbk * 2 + 1];
            sbk++;
          }
          t.total = bytes;
        } else if (type_is_binary(h->out_types[c])) {
          // utf8 bytes: one COPY task per span
          int64_t d = t.dst;
          for (const Span& sp : bt.spans) {
            SplitSegment* sg = h->segments[sp.seg];
            int64_t bstart = 0;
            {
              int in_col = in_of[c];
              const int j = in_col >= 0 ? h->cols[in_col].bin : -1;
              bstart = sg->h_byte_start[(size_t)j * (P + 1) + bt.p];
            }
            const int64_t o0 = span_bytes[sb * 2], nb2 = span_bytes[sb * 2 + 1];
            IpcTask u{};
            u.kind = IPC_COPY;
            u.p0 = (int64_t)plan->pieces.size();
            u.np = 1;
            u.dst = d;
            u.n = nb2;
            u.out = nb2;
            plan->pieces.push_back({(const uint8_t*)sg->data[c].p + bstart + o0, 0, 0, 0, 0});
            plan->tasks.push_back(u);
            d += nb2;
            sb++;
          }
          if (pd.out > pd.n) {  // zero padding of the buffer
            IpcTask z{};
            z.kind = IPC_COPY;
            z.p0 = (int64_t)plan->pieces.size();
            z.np = 1;
            z.dst = d;
            z.n = 0;
            z.out = pd.out - pd.n;
            plan->pieces.push_back({nullptr, 0, 0, 0, 0});
            plan->tasks.push_back(z);
          }
          continue;
        } else {
          // fixed-width values: one COPY task per span
          const int w = type_width(h->out_types[c]);
          int64_t d = t.dst;
          for (const Span& sp : bt.spans) {
            SplitSegment* sg = h->segments[sp.seg];
            IpcTask u{};
            u.kind = IPC_COPY;
            u.p0 = (int64_t)plan->pieces.size();
            u.np = 1;
            u.dst = d;
            u.n = sp.rows * w;
            u.out = sp.rows * w;
            plan->pieces.push_back(
                {(const uint8_t*)sg->values[c].p + (sg->h_row_start[bt.p] + sp.row0) * w, 0, 0, 0, 0});
            plan->tasks.push_back(u);
            d += sp.rows * w;
          }
          if (pd.out > pd.n) {
            IpcTask z{};
            z.kind = IPC_COPY;
            z.p0 = (int64_t)plan->pieces.size();
            z.np = 1;
            z.dst = d;
            z.n = 0;
            z.out = pd.out - pd.n;
            plan->pieces.push_back({nullptr, 0, 0, 0, 0});
            plan->tasks.push_back(z);
          }
          continue;
        }
        t.np = (int64_t)plan->pieces.size() - t.p0;
        plan->tasks.push_back(t);
      }
      pos = body0 + boff;
      bi++;
    }
    if (any) {
      static const uint8_t eos[8] = {0xff, 0xff, 0xff, 0xff, 0, 0, 0, 0};
      add_copy(eos, true, 8, 8);
    }
    plan->part_len[p] = pos - start;
  }
  plan->total = pos;
  return GZ_OK;
}

// Uploads a plan's blob, pieces and tasks and assembles its stream into dst.
static int ipc_assemble(gz_context* ctx, gz_ipc_plan* pl, void* dst) {
  GZ_TRY(upload(ctx, &pl->d_blob, pl->blob.data(), std::max<size_t>(pl->blob.size(), 1)));
  // blob pieces point into the uploaded blob
  for (auto& t : pl->tasks)
    if (t.pad == 1) {
      IpcPiece& pc = pl->pieces[t.p0];
      pc.src = (const uint8_t*)pl->d_blob.p + pc.a;
    }
  GZ_TRY(upload(ctx, &pl->d_pieces, pl->pieces.data(), sizeof(IpcPiece) * std::max<size_t>(pl->pieces.size(), 1)));
  GZ_TRY(upload(ctx, &pl->d_tasks, pl->tasks.data(), sizeof(IpcTask) * std::max<size_t>(pl->tasks.size(), 1)));
  const unsigned g = (unsigned)std::min<size_t>(pl->tasks.size(), 1 << 20);
  {
    KernelTimer kt(ctx, "split_ipc");
    hipLaunchKernelGGL(ipc_assemble_kernel, dim3(g), dim3(256), 0, ctx->stream, (const IpcTask*)pl->d_tasks.p,
                       (int64_t)pl->tasks.size(), (const IpcPiece*)pl->d_pieces.p, (uint8_t*)dst);
  }
  GZ_HIP(hipGetLastError());
  return GZ_OK;
}

// LZ4_FRAME bodies: the uncompressed stream is assembled into scratch, every
// non-empty body buffer is cut into 32 KiB blocks compressed on the device,
// and the plan is rewritten for the compressed stream: per buffer the int64
// uncompressed length, one LZ4 frame (independent blocks, no checksums; a block
// that does not shrink is stored), buffers 8-byte padded; record batch
// messages carry BodyCompression{LZ4_FRAME, BUFFER}.
static int ipc_compress_plan(gz_splitter* h, gz_ipc_plan* plan) {
  gz_context* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  if (plan->total == 0) return GZ_OK;
  GZ_TRY(ctx_alloc(ctx, (size_t)plan->total + 8, &plan->raw));
  GZ_TRY(ipc_assemble(ctx, plan, plan->raw.p));
  const uint8_t* raw = (const uint8_t*)plan->raw.p;
  constexpr int64_t kBlock = kLz4Block;
  // record batches of at most batch_compress_threshold rows are written
  // uncompressed (splitter.cc:586-590: the tiny-batch options' codec is
  // UNCOMPRESSED)
  const int64_t tiny = h->opts.batch_compress_threshold;
  std::vector<Lz4CBlock> blocks;
  int64_t coff = 0;
  for (const auto& r : plan->recs)
    if (r.rows > tiny)
      for (const auto& b : r.bufs)
      for (int64_t k = 0; k < b.second; k += kBlock) {
        const int64_t n = std::min(kBlock, b.second - k);
        blocks.push_back({raw + r.body_pos + b.first + k, (uint8_t*)(intptr_t)coff, (uint32_t)n, 0});
        coff += (n + 15) / 16 * 16;
      }
  std::vector<uint32_t> csize(std::max<size_t>(blocks.size(), 1), 0);
  if (!blocks.empty()) {
    GZ_TRY(ctx_alloc(ctx, (size_t)coff, &plan->comp));
    for (auto& b : blocks) b.dst = (uint8_t*)plan->comp.p + (intptr_t)b.dst;
    GZ_TRY(upload(ctx, &plan->d_cblocks, blocks.data(), sizeof(Lz4CBlock) * blocks.size()));
    GZ_TRY(ctx_alloc(ctx, sizeof(uint32_t) * blocks.size(), &plan->d_csize));
    const size_t lds = kLz4Block + 8 + 2 * (1 << kLz4HashLog);
    {
      KernelTimer kt(ctx, "split_ipc_lz4");
      hipLaunchKernelGGL(lz4_compress_kernel, dim3((unsigned)blocks.size()), dim3(64), lds, st,
                         (const Lz4CBlock*)plan->d_cblocks.p, (int)blocks.size(), (uint32_t*)plan->d_csize.p);
    }
    GZ_HIP(hipGetLastError());
    GZ_HIP(hipMemcpyAsync(csize.data(), plan->d_csize.p, sizeof(uint32_t) * blocks.size(), hipMemcpyDeviceToHost, st));
  }
  GZ_HIP(hipStreamSynchronize(st));
  // the compressed stream
  const int P = h->P;
  plan->tasks.clear();
  plan->pieces.clear();
  plan->blob.clear();
  plan->part_len.assign(P, 0);
  int64_t pos = 0;
  auto blob_copy = [&](const std::vector<uint8_t>& bytes, int64_t out) {
    IpcTask t{};
    t.kind = IPC_COPY;
    t.p0 = (int64_t)plan->pieces.size();
    t.np = 1;
    t.dst = pos;
    t.n = (int64_t)bytes.size();
    t.out = out;
    t.pad = 1;
    IpcPiece pc{};
    pc.a = (int64_t)plan->blob.size();
    plan->blob.insert(plan->blob.end(), bytes.begin(), bytes.end());
    plan->pieces.push_back(pc);
    plan->tasks.push_back(t);
    pos += out;
  };
  auto dev_copy = [&](const uint8_t* src, int64_t n) {
    IpcTask t{};
    t.kind = IPC_COPY;
    t.p0 = (int64_t)plan->pieces.size();
    t.np = 1;
    t.dst = pos;
    t.n = n;
    t.out = n;
    IpcPiece pc{};
    pc.src = src;
    plan->pieces.push_back(pc);
    plan->tasks.push_back(t);
    pos += n;
  };
  auto le32 = [](std::vector<uint8_t>& v, uint32_t x) {
    for (int i = 0; i < 4; i++) v.push_back((uint8_t)(x >> (8 * i)));
  };
  // frame header: magic, FLG (version 01, independent blocks), BD (64 KiB maximum
  // blocks), header checksum = second byte of XXH32(FLG, BD) with seed 0
  static const uint8_t kFrameHeader[7] = {0x04, 0x22, 0x4D, 0x18, 0x60, 0x40, 0x82};
  size_t bi = 0, ri = 0;
  for (int p = 0; p < P; p++) {
    const int64_t start = pos;
    bool any = false;
    for (; ri < plan->recs.size() && plan->recs[ri].p == p; ri++) {
      const auto& r = plan->recs[ri];
      if (!any && !plan->schema.empty()) blob_copy(plan->schema, (int64_t)plan->schema.size());
      any = true;
      if (r.rows <= tiny) {  // the uncompressed record batch: message and body as laid out
        int64_t body = 0;
        for (const auto& b : r.bufs) body = std::max<int64_t>(body, b.first + pad8(b.second));
        const std::vector<uint8_t> m = ipc_record_batch_message(r.rows, r.nodes, r.bufs, body);
        blob_copy(m, (int64_t)m.size());
        if (body) dev_copy(raw + r.body_pos, body);
        continue;
      }
      // compressed buffer sizes
      std::vector<std::pair<int64_t, int64_t>> cb;
      int64_t boff = 0;
      size_t bj = bi;
      for (const auto& b : r.bufs) {
        if (b.second == 0) {
          cb.push_back({boff, 0});
          continue;
        }
        int64_t clen = 8 + 7 + 4;
        for (int64_t k = 0; k < b.second; k += kBlock, bj++) {
          const int64_t n = std::min(kBlock, b.second - k);
          clen += 4 + (csize[bj] ? (int64_t)csize[bj] : n);
        }
        cb.push_back({boff, clen});
        boff += pad8(clen);
      }
      const std::vector<uint8_t> msg = ipc_record_batch_message(r.rows, r.nodes, cb, boff, true);
      blob_copy(msg, (int64_t)msg.size());
      for (size_t q = 0; q < r.bufs.size(); q++) {
        const auto& b = r.bufs[q];
        if (b.second == 0) continue;
        const int64_t buf0 = pos;
        std::vector<uint8_t> hdr(8);
        memcpy(hdr.data(), &b.second, 8);  // uncompressed length
        hdr.insert(hdr.end(), kFrameHeader, kFrameHeader + 7);
        for (int64_t k = 0; k < b.second; k += kBlock, bi++) {
          const int64_t n = std::min(kBlock, b.second - k);
          const bool stored = csize[bi] == 0;
          le32(hdr, stored ? (uint32_t)n | 0x80000000u : csize[bi]);
          blob_copy(hdr, (int64_t)hdr.size());
          hdr.clear();
          if (stored) dev_copy(raw + r.body_pos + b.first + k, n);
          else dev_copy(blocks[bi].dst, csize[bi]);
        }
        le32(hdr, 0);  // end mark
        blob_copy(hdr, pad8(pos + 4 - buf0) - (pos - buf0));
      }
    }
    if (any) blob_copy(std::vector<uint8_t>{0xff, 0xff, 0xff, 0xff, 0, 0, 0, 0}, 8);
    plan->part_len[p] = pos - start;
  }
  plan->total = pos;
  return GZ_OK;
}

extern "C" {

int gz_ipc_record_batch_message(int64_t num_rows, int32_t num_nodes, const int64_t* nodes, int32_t num_buffers,
                                const int64_t* buffers, int64_t body_length, uint8_t* out, int64_t out_capacity,
                                int64_t* out_length) {
  if ((num_nodes > 0 && !nodes) || (num_buffers > 0 && !buffers) || !out_length)
    return fail(GZ_INVALID, "gz_ipc_record_batch_message: null argument");
  std::vector<std::pair<int64_t, int64_t>> nv, bv;
  for (int i = 0; i < num_nodes; i++) nv.push_back({nodes[2 * i], nodes[2 * i + 1]});
  for (int i = 0; i < num_buffers; i++) bv.push_back({buffers[2 * i], buffers[2 * i + 1]});
  const std::vector<uint8_t> m = ipc_record_batch_message(num_rows, nv, bv, body_length);
  *out_length = (int64_t)m.size();
  if (out && out_capacity >= (int64_t)m.size()) memcpy(out, m.data(), m.size());
  else if (out) return fail(GZ_INVALID, "gz_ipc_record_batch_message: output too small");
  return GZ_OK;
}

int gz_split_set_compress_type(gz_splitter* h, int32_t codec) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (codec != GZ_COMPRESSION_UNCOMPRESSED && codec != GZ_COMPRESSION_LZ4_FRAME)
    return fail(GZ_NOT_IMPLEMENTED, "compression type " + std::to_string(codec) + " (UNCOMPRESSED or LZ4_FRAME)");
  h->codec = codec;
  return GZ_OK;
}

int gz_split_ipc_prepare(gz_splitter* h, const uint8_t* schema_msg, int64_t schema_len, int32_t write_schema,
                         int64_t* total_bytes, int64_t* partition_lengths) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (write_schema && schema_len > 0 && !schema_msg) return fail(GZ_INVALID, "null schema message");
  GZ_HIP(hipSetDevice(h->ctx->device));
  if (h->ipc) ipc_plan_free(h->ctx, h->ipc);
  h->ipc = new gz_ipc_plan();
  const auto t0 = std::chrono::steady_clock::now();
  int rc = ipc_prepare(h, schema_msg, schema_len, write_schema != 0, h->ipc);
  const auto t1 = std::chrono::steady_clock::now();
  if (!rc && h->codec == GZ_COMPRESSION_LZ4_FRAME) rc = ipc_compress_plan(h, h->ipc);
  const auto t2 = std::chrono::steady_clock::now();
  h->write_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
  h->compress_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(t2 - t1).count();
  if (rc) {
    ipc_plan_free(h->ctx, h->ipc);
    h->ipc = nullptr;
    return rc;
  }
  if (total_bytes) *total_bytes = h->ipc->total;
  if (partition_lengths) memcpy(partition_lengths, h->ipc->part_len.data(), sizeof(int64_t) * h->P);
  return GZ_OK;
}

int gz_split_ipc_write(gz_splitter* h, void* dst, int64_t dst_bytes) {
  if (!h) return fail(GZ_INVALID, "null splitter");
  if (!h->ipc) return fail(GZ_INVALID, "gz_split_ipc_prepare must come first");
  gz_ipc_plan* pl = h->ipc;
  if (dst_bytes < pl->total) return fail(GZ_INVALID, "destination smaller than the stream");
  if (pl->total == 0) return GZ_OK;
  if (!dst) return fail(GZ_INVALID, "null destination");
  gz_context* ctx = h->ctx;
  GZ_HIP(hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  GZ_TRY(ipc_assemble(ctx, pl, dst));
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  h->write_ns +=
      std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
  return GZ_OK;
}

int gz_split_get_metrics(gz_splitter* h, gz_split_metrics* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  drain_pid_events(h);
  out->total_compute_pid_time = h->compute_pid_ns;
  out->total_write_time = h->write_ns;
  out->total_spill_time = 0;
  out->total_compress_time = h->compress_ns;
  out->total_bytes_written = h->ipc ? h->ipc->total : 0;
  out->total_bytes_spilled = 0;
  return GZ_OK;
}

int gz_split_raw_partition_lengths(gz_splitter* h, int64_t* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  if (!h->ipc) return fail(GZ_INVALID, "gz_split_ipc_prepare must come first");
  memcpy(out, h->ipc->raw_len.data(), sizeof(int64_t) * h->P);
  return GZ_OK;
}

int gz_split_batch_rows(gz_splitter* h, int64_t* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  if (!h->ipc) return fail(GZ_INVALID, "gz_split_ipc_prepare must come first");
  *out = h->ipc->batch_rows;
  return GZ_OK;
}


}  // extern "C"
