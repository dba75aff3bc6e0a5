// Hash join build and probe for MI355X (gfx950).
//
// Build replaces HashRelationKernel / HashRelation (ext/hash_relation_kernel.cc,
// codegen/common/hash_relation.h) on an unsafeHashMap (row_wise_memory/hashMap.h);
// probe replaces ConditionedProbeResultIterator's inner / semi / anti functions
// (ext/conditioned_probe_kernel.cc:546-1173).
//
// The relation copies every build batch into a column store it owns in HBM (rows
// in batch order, then row order = the reference's insertion order).  Finish
// builds an open-addressing table of unique keys (linear probing, 64-bit CAS on
// the key, capacity 2^(ceil(log2 rows) + 1) so the load stays <= 0.5 as in
// hash_relation_kernel.cc:205-224).  Each slot keeps its key's row count and
// first row; when some key repeats, a stable sort of (slot, row) lays the rows of
// every key out contiguously in insertion order, which is the order the
// reference's bytesMap chain returns them in (hashMap.h:826-876).
//
// A probe call runs count -> tile scan -> write: per probe row the number of
// output rows (inner: the key's row count; semi: 1 on a match; anti: 1 on no
// match, with the null-aware rules of :1020-1035), an exclusive scan in row
// order, and a write pass that gathers probe and build columns into the output
// in probe-row order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <string>
#include <vector>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

constexpr int kJoinMaxCols = 32;
#ifndef GZ_WRITE01_BLOCK
#define GZ_WRITE01_BLOCK 1024
#endif
constexpr int kWrite01Block = GZ_WRITE01_BLOCK;  // threads of the at-most-one-row write pass (4096-row tiles)
// one-batch calls: 256-row workgroups, one row per thread, at most one wave of workgroups
constexpr int kSingleTile = 256;
constexpr int kSingleMaxBlocks = 64;
constexpr int kProbeInlineBufs = 48;  // buffers of up to 16 probe columns
constexpr int kLookSpins = 1 << 20;   // polls of an earlier block's count before giving up
constexpr int64_t kJEmpty = INT64_MIN;

__device__ __forceinline__ uint64_t jmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// Join key of a typed value, widened to 64 bits: integers sign/zero extended,
// booleans 0/1, floats as the bits of the double value (after Spark's
// normalize when norm is set).  *nan marks a NaN key, which never equals
// anything: the reference compares one typed key with == after its hash
// (hash_relation.h:368-383), normalised or not.
__device__ __forceinline__ int64_t join_key(const void* v, int t, int64_t r, bool* nan, int norm = 0) {
  *nan = false;
  switch (t) {
    case GZ_TYPE_BOOL: return bit_get((const uint8_t*)v, r);
    case GZ_TYPE_INT8: return ((const int8_t*)v)[r];
    case GZ_TYPE_UINT8: return ((const uint8_t*)v)[r];
    case GZ_TYPE_INT16: return ((const int16_t*)v)[r];
    case GZ_TYPE_UINT16: return ((const uint16_t*)v)[r];
    case GZ_TYPE_INT32: case GZ_TYPE_DATE32: return ((const int32_t*)v)[r];
    case GZ_TYPE_UINT32: return ((const uint32_t*)v)[r];
    case GZ_TYPE_FLOAT: {
      double d = (double)((const float*)v)[r];
      *nan = d != d;
      if (norm) d = (double)(float)norm_nan_zero(d);
      return __double_as_longlong(d);
    }
    case GZ_TYPE_DOUBLE: {
      double d = ((const double*)v)[r];
      *nan = d != d;
      if (norm) d = norm_nan_zero(d);
      return __double_as_longlong(d);
    }
    default: return ((const int64_t*)v)[r];
  }
}

// The probe reads only the 8-byte keys (the smallest table for the random
// accesses); starts and counts are read for matched rows that need them.
struct JTable {
  const int64_t* keys;     // [cap]
  const int32_t* start;    // [cap] first row (unique keys) or rowlist start
  const int32_t* count;    // [cap]
  const int32_t* rowlist;  // rows grouped by key in insertion order, or null (unique keys)
  uint64_t cap;
  int64_t special_start, special_count;  // the key equal to kJEmpty
};

// ---------------------------------------------------------------- build store
// One appended batch column: bits -> bytes for validity and booleans, values
// copied, utf8 offsets rebased to the store's int64 offsets.
struct AppendCol {
  const uint8_t* valid;  // input bitmap or null
  const void* values;
  const int32_t* offsets;
  int type, width;
  uint8_t* dvalid;       // store validity bytes (row-indexed)
  uint8_t* dvalues;      // store values (fixed: width per row; bool: 1 byte)
  int64_t* doffsets;     // store offsets (binary), entry row+1 written
  uint8_t* ddata;
  int64_t data_base;     // store byte position of this batch's first byte
};

struct AppendParams {
  int ncols;
  AppendCol col[kJoinMaxCols];
  int64_t n;
  int64_t row0;  // store row of the batch's first row
};

__global__ void join_append_kernel(AppendParams p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < p.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int64_t d = p.row0 + i;
    for (int c = 0; c < p.ncols; c++) {
      const AppendCol& a = p.col[c];
      if (a.type == GZ_TYPE_NA) continue;
      a.dvalid[d] = a.valid ? (uint8_t)bit_get(a.valid, i) : 1;
      if (a.type == GZ_TYPE_BOOL) {
        a.dvalues[d] = (uint8_t)bit_get((const uint8_t*)a.values, i);
      } else if (a.type == GZ_TYPE_STRING || a.type == GZ_TYPE_BINARY) {
        a.doffsets[d + 1] = a.data_base + (a.offsets[i + 1] - a.offsets[0]);
        if (i == 0) a.doffsets[d] = a.data_base;
      } else {
        const uint8_t* s = (const uint8_t*)a.values + i * a.width;
        uint8_t* t = a.dvalues + d * a.width;
        for (int k = 0; k < a.width; k++) t[k] = s[k];
      }
    }
  }
}

// Many batches of fixed-width columns in one launch: one workgroup per tile of
// a batch (the tile table), the batch's buffers from a device table.
struct AppendTile {
  int32_t batch, row_begin, nrows, pad;
  int64_t row0;  // store row of the tile's first row
};

struct AppendManyParams {
  int ncols;
  int type[kJoinMaxCols], width[kJoinMaxCols];
  uint8_t* dvalid[kJoinMaxCols];
  uint8_t* dvalues[kJoinMaxCols];
  const AppendTile* tiles;
  const void* const* bufs;  // [batch * ncols * 2 + col * 2 + {0 validity, 1 values}]
};

__global__ __launch_bounds__(256) void join_append_many_kernel(AppendManyParams p) {
  const AppendTile T = p.tiles[blockIdx.x];
  const void* const* b = p.bufs + (int64_t)T.batch * p.ncols * 2;
  for (int i = threadIdx.x; i < T.nrows; i += blockDim.x) {
    const int64_t r = T.row_begin + i, d = T.row0 + i;
    for (int c = 0; c < p.ncols; c++) {
      const int t = p.type[c];
      if (t == GZ_TYPE_NA) continue;
      const uint8_t* vb = (const uint8_t*)b[c * 2];
      const uint8_t* v = (const uint8_t*)b[c * 2 + 1];
      p.dvalid[c][d] = vb ? (uint8_t)bit_get(vb, r) : 1;
      const int w = p.width[c];
      if (t == GZ_TYPE_BOOL) p.dvalues[c][d] = (uint8_t)bit_get(v, r);
      else if (w == 8) ((uint64_t*)p.dvalues[c])[d] = ((const uint64_t*)v)[r];
      else if (w == 4) ((uint32_t*)p.dvalues[c])[d] = ((const uint32_t*)v)[r];
      else if (w == 2) ((uint16_t*)p.dvalues[c])[d] = ((const uint16_t*)v)[r];
      else if (w == 1) p.dvalues[c][d] = v[r];
      else
        for (int k = 0; k < w; k++) p.dvalues[c][d * w + k] = v[r * w + k];
    }
  }
}

__global__ void join_copy_bytes_kernel(uint8_t* dst, const uint8_t* src, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// ---------------------------------------------------------------- build table
__global__ void join_table_init_kernel(int64_t* keys, int32_t* start, int32_t* count, uint64_t cap) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
       s += (int64_t)gridDim.x * blockDim.x) {
    keys[s] = kJEmpty;
    start[s] = INT32_MAX;
    count[s] = 0;
  }
}

// flags: [0] has null key, [1] has duplicate key, [2] inserted keys,
//        [3] rows with key kJEmpty, [4] first such row
__global__ void join_build_kernel(const uint8_t* kvalid, const uint8_t* kvals, int ktype, int kwidth,
                                  int64_t n, int64_t* keys, int32_t* start, int32_t* count,
                                  uint64_t cap, int32_t* row_slot, int skip_dup, int codes, int norm,
                                  unsigned long long* flags) {
  const uint64_t mask = cap - 1;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    row_slot[r] = -1;
    bool fresh = false;  // this row inserted a new key
    if (!kvalid[r]) {
      flags[0] = 1;
    } else {
      bool nan;
      int64_t key;
      if (ktype == GZ_TYPE_BOOL) {
        key = kvals[r];
        nan = false;
      } else {
        key = join_key(kvals, ktype, r, &nan, norm);
      }
      // a NaN build key can never be matched; a multi-column key with a null
      // part (code -1) is never inserted (hash_relation.h:186-187)
      if (nan || (codes && key == -1)) {
      } else if (key == kJEmpty) {
        unsigned long long c = atomicAdd(&flags[3], 1ull);
        atomicMin((long long*)&flags[4], (long long)r);
        if (c >= 1) flags[1] = 1;
        row_slot[r] = -2;
      } else {
        uint64_t s = jmix64((uint64_t)key) & mask;
        for (uint64_t i = 0; i < cap; i++) {
          int64_t k = keys[s];
          if (k == kJEmpty) {
            unsigned long long prev = atomicCAS((unsigned long long*)&keys[s], (unsigned long long)kJEmpty,
                                                (unsigned long long)key);
            if ((int64_t)prev == kJEmpty) {
              fresh = true;
              k = key;
            } else {
              k = (int64_t)prev;
            }
          }
          if (k == key) {
            // the row that inserted the key owns the slot: its count and start
            // are plain stores; a later row of the key marks repeated keys,
            // whose counts and starts the finish recomputes (skip-duplicate
            // building keeps the smallest row)
            if (fresh) {
              count[s] = 1;
              if (skip_dup) atomicMin(&start[s], (int32_t)r);
              else start[s] = (int32_t)r;
            } else {
              flags[1] = 1;
              if (skip_dup) atomicMin(&start[s], (int32_t)r);
            }
            row_slot[r] = (int32_t)s;
            break;
          }
          s = (s + 1) & mask;
        }
      }
    }
    // distinct keys: one atomic per wave, by its lowest active lane
    const unsigned long long m = __ballot(fresh);
    const unsigned long long act = __ballot(1);
    if (m && (int)__lane_id() == __ffsll((long long)act) - 1) atomicAdd(&flags[2], (unsigned long long)__popcll(m));
  }
  (void)kwidth;
}

// Rows per slot of a table with repeated keys (the build counts only inserts).
__global__ void join_recount_kernel(const int32_t* row_slot, int64_t n, int32_t* count) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = row_slot[r];
    if (s >= 0) atomicAdd(&count[s], 1);
  }
}

__global__ void join_zero_i32_kernel(int32_t* p, uint64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// With skip-duplicate building (semi joins) every key keeps one row.
__global__ void join_clamp_counts_kernel(int32_t* count, uint64_t cap) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
       s += (int64_t)gridDim.x * blockDim.x)
    if (count[s] > 1) count[s] = 1;
}

// (slot, row) sort keys of the rows of repeated keys.
__global__ void join_pair_keys_kernel(const int32_t* row_slot, int64_t n, unsigned long long* out) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    int32_t s = row_slot[r];
    unsigned long long hi = s >= 0 ? (unsigned long long)s : (s == -2 ? 0x7fffffffull : 0xffffffffull);
    out[r] = (hi << 32) | (unsigned long long)r;
  }
}

__global__ void join_rowlist_kernel(const unsigned long long* sorted, int64_t n, int32_t* rowlist) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    rowlist[i] = (int32_t)(sorted[i] & 0xffffffffull);
}

// Row-major copy of the fixed-width build columns a probe outputs: per build row
// one 16-byte record, the values at their offsets and a validity bit per
// column in the last dword, so a matched row costs one random 16-byte load.
struct PayloadParams {
  int ncol;
  const uint8_t* bvalid[kJoinMaxCols];   // store validity bytes
  const uint8_t* bvalues[kJoinMaxCols];  // store values (width bytes per row)
  int width[kJoinMaxCols];
  int off[kJoinMaxCols];                 // byte offset in the record
  int bit[kJoinMaxCols];                 // validity bit: the output column index
  int64_t n;
  u32x4* out;
};

__device__ __forceinline__ u32x4 payload_record(const PayloadParams& q, int64_t r) {
  uint64_t lo = 0, hi = 0;
  uint32_t vm = 0;
  for (int c = 0; c < q.ncol; c++) {
    const int w = q.width[c], off = q.off[c];
    uint64_t v;
    const uint8_t* src = q.bvalues[c] + r * w;
    if (w == 8) v = gld((const uint64_t*)src);
    else if (w == 4) v = gld((const uint32_t*)src);
    else if (w == 2) v = gld((const uint16_t*)src);
    else v = gld(src);
    if (off < 8) lo |= v << (8 * off);
    else hi |= v << (8 * (off - 8));
    if (gld(q.bvalid[c] + r)) vm |= 1u << q.bit[c];
  }
  return u32x4{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, vm};
}

__global__ void join_payload_kernel(PayloadParams q) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < q.n;
       r += (int64_t)gridDim.x * blockDim.x)
    gst(q.out + r, payload_record(q, r));
}

// The payload of a table whose keys each name one build row, stored by slot:
// the row a key's slot keeps (start[slot]: the inserting row, or the smallest
// with skip-duplicate building) goes to record `slot`, the row of the key equal
// to the empty marker to record cap; so the probe's match code (the slot) is
// the record's index and the count pass reads no start.
__global__ void join_payload_slot_kernel(PayloadParams q, const int32_t* row_slot, const int32_t* start,
                                         int64_t special_start, uint64_t cap) {
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < q.n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int32_t s = row_slot[r];
    int64_t dst;
    if (s >= 0 && start[s] == (int32_t)r) dst = s;
    else if (s == -2 && r == special_start) dst = (int64_t)cap;
    else continue;
    gst(q.out + dst, payload_record(q, r));
  }
}

// ---------------------------------------------------------------- probe
struct PTile {
  int32_t batch;
  int32_t row_begin;
  int32_t nrows;
  int32_t pad;
};

struct OutCol {
  int side;        // 0 build, 1 probe
  int col;         // column on that side
  int type, width;
  // build store
  const uint8_t* bvalid;
  const uint8_t* bvalues;
  // output
  uint8_t* ovalid;
  uint8_t* ovalues;
  int binary;
};

struct ProbeParams {
  const PTile* tiles;
  int64_t T;
  const void* const* bufs;  // probe batches: [batch * bstride + col * 3 + k]
  int bstride;
  int key_col, key_type;
  int join_type, null_aware;
  int normalize;            // Spark's normalize on a float/double key
  int build_has_null;
  int64_t build_keys;       // keys in the table (null keys excluded)
  JTable J;
  // per probe row (tile-major): the matched slot (with match_rows: the matched
  // build row), -1 no match, -2 the kJEmpty key, -3 a null key
  int32_t* match;
  int match_rows;  // inner joins over unique keys: the count pass resolves rows
  int by_slot;     // the payload is stored by slot (its record index is the match code)
  int64_t b_special;  // payload record of the key equal to kJEmpty
  uint32_t* tile_count;
  const uint64_t* tile_off;
  int ncols;
  OutCol out[kJoinMaxCols];
  int64_t* pair_probe;      // (tile, row) of each output row when binary columns need a gather
  int32_t* pair_build;
  int write_pairs;
  int tile_rows;
  // build columns gathered from the row-major payload (null: per-column gathers)
  const u32x4* payload;
  int pay_off[kJoinMaxCols];  // byte offset of output column c in a payload record
  // a call of one small batch (join_probe_write01_kernel with SINGLE = 1): the
  // batch and its buffers ride in the arguments (bufs null), each block posts
  // its emitted-row count tagged with the call's epoch, and the row total goes
  // straight to host memory
  PTile tile0;
  uint64_t* look;       // [kSingleMaxBlocks]: epoch << 32 | rows of that block
  uint32_t epoch;
  uint32_t* host_out;   // [0] output rows, [1] set when a wait for an earlier block ran out
  const void* ibufs[kProbeInlineBufs];
};
static_assert(sizeof(ProbeParams) <= 4096, "kernel arguments");

__device__ __forceinline__ const void* pbuf(const ProbeParams& p, int batch, int col, int k) {
  const int64_t i = (int64_t)batch * p.bstride + col * 3 + k;
  return p.bufs ? p.bufs[i] : p.ibufs[i];
}

// Output rows of a probe row given its match code (see ProbeParams::match) and
// the matched slot's row count (tables with repeated keys).
__device__ __forceinline__ int out_count(const ProbeParams& p, int32_t m, int32_t slot_count) {
  const bool key_null = m == -3;
  if (p.join_type == GZ_JOIN_INNER) {
    if (m >= 0) return p.J.rowlist ? slot_count : 1;
    if (m == -2) return (int)p.J.special_count;
    return 0;
  }
  bool found = m >= 0 || m == -2;
  if (p.join_type == GZ_JOIN_SEMI) return found ? 1 : 0;
  // anti
  if (!p.null_aware) return found ? 0 : 1;
  if (p.build_keys == 0 && !p.build_has_null) return 1;
  if (p.build_has_null) return 0;
  return (!key_null && !found) ? 1 : 0;
}

// Probe keys of RPT lane-consecutive rows (row lr0 + k * S), all loads issued
// before any use; *nan marks NaN keys (never equal), bit k of the returned mask
// a non-null active row.
template <int RPT, int S>
__device__ __forceinline__ uint32_t probe_keys(const ProbeParams& p, const PTile& td, int lr0, int64_t (&key)[RPT],
                                               uint32_t* nanm) {
  const uint8_t* kv = (const uint8_t*)pbuf(p, td.batch, p.key_col, 0);
  const void* kd = pbuf(p, td.batch, p.key_col, 2);
  const int64_t rb = td.row_begin;
  const int n = td.nrows;
  uint8_t vx[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) vx[k] = kv && lr0 + k * S < n ? gld(kv + ((rb + lr0 + k * S) >> 3)) : (uint8_t)0xff;
  *nanm = 0;
#define GZ_PK(T, CONV)                                                                              \
  {                                                                                                 \
    T x[RPT];                                                                                       \
    _Pragma("unroll") for (int k = 0; k < RPT; k++) x[k] = lr0 + k * S < n ? gld((const T*)kd + rb + lr0 + k * S) : (T)0; \
    _Pragma("unroll") for (int k = 0; k < RPT; k++) key[k] = CONV(x[k]);                            \
    break;                                                                                          \
  }
#define GZ_PK_I(x) ((int64_t)(x))
  switch (p.key_type) {
    case GZ_TYPE_BOOL: {
      uint8_t x[RPT];
#pragma unroll
      for (int k = 0; k < RPT; k++) x[k] = lr0 + k * S < n ? gld((const uint8_t*)kd + ((rb + lr0 + k * S) >> 3)) : 0;
#pragma unroll
      for (int k = 0; k < RPT; k++) key[k] = (x[k] >> ((rb + lr0 + k * S) & 7)) & 1;
      break;
    }
    case GZ_TYPE_INT8: GZ_PK(int8_t, GZ_PK_I)
    case GZ_TYPE_UINT8: GZ_PK(uint8_t, GZ_PK_I)
    case GZ_TYPE_INT16: GZ_PK(int16_t, GZ_PK_I)
    case GZ_TYPE_UINT16: GZ_PK(uint16_t, GZ_PK_I)
    case GZ_TYPE_INT32: case GZ_TYPE_DATE32: GZ_PK(int32_t, GZ_PK_I)
    case GZ_TYPE_UINT32: GZ_PK(uint32_t, GZ_PK_I)
    case GZ_TYPE_FLOAT:
    case GZ_TYPE_DOUBLE: {
      // the bits of the double value; NaN keys never match
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        bool nan = false;
        key[k] = lr0 + k * S < n ? join_key(kd, p.key_type, rb + lr0 + k * S, &nan, p.normalize) : 0;
        if (nan) *nanm |= 1u << k;
      }
      break;
    }
    default: GZ_PK(int64_t, GZ_PK_I)
  }
#undef GZ_PK
#undef GZ_PK_I
  uint32_t vm = 0;
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (lr0 + k * S < n && ((vx[k] >> ((rb + lr0 + k * S) & 7)) & 1)) vm |= 1u << k;
  return vm;
}

// Match codes (see ProbeParams::match) of the tile rows lr0 + k * S: keys and
// home slots loaded together, collisions walk on.
template <int RPT, int S>
__device__ __forceinline__ void probe_match(const ProbeParams& p, const PTile& td, int lr0, int32_t (&m)[RPT]) {
  const uint64_t mask = p.J.cap - 1;
  int64_t key[RPT];
  uint32_t nanm;
  const uint32_t vm = probe_keys<RPT, S>(p, td, lr0, key, &nanm);
  uint64_t s[RPT];
  int64_t sk[RPT];
  int32_t st[RPT];  // the home slot's first row, loaded with its key (match_rows)
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    s[k] = jmix64((uint64_t)key[k]) & mask;
    const bool look = ((vm >> k) & 1) && !((nanm >> k) & 1) && key[k] != kJEmpty;
    sk[k] = look ? gld(p.J.keys + s[k]) : kJEmpty;
    st[k] = look && p.match_rows ? gld(p.J.start + s[k]) : 0;
  }
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    m[k] = -1;
    if (lr0 + k * S >= td.nrows) continue;
    if (!((vm >> k) & 1)) m[k] = -3;
    else if ((nanm >> k) & 1) m[k] = -1;
    else if (key[k] == kJEmpty) m[k] = p.J.special_count > 0 ? -2 : -1;
    else {
      int64_t v = sk[k];
      uint64_t q = s[k];
      for (uint64_t i = 0; i < p.J.cap; i++) {
        if (v == key[k]) {
          m[k] = !p.match_rows ? (int32_t)q : q == s[k] ? st[k] : gld(p.J.start + q);
          break;
        }
        if (v == kJEmpty) break;
        q = (q + 1) & mask;
        v = gld(p.J.keys + q);
      }
    }
  }
}

// Per probe row: the match code and the number of output rows; per tile the
// total.  Rows are lane-consecutive (coalesced key loads and match stores), the
// home slots of all RPT rows are loaded together and collisions walk on.
template <int BLOCK, int RPT>
__global__ __launch_bounds__(BLOCK) void join_probe_count_kernel(ProbeParams p) {
  __shared__ uint64_t s_scan[BLOCK / kWave + 1];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const PTile td = p.tiles[t];
  const int tid = threadIdx.x;
  uint64_t sum = 0;
  for (int base = 0; base < td.nrows; base += BLOCK * RPT) {
    int32_t m[RPT];
    probe_match<RPT, BLOCK>(p, td, base + tid, m);
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int lr = base + k * BLOCK + tid;
      if (lr >= td.nrows) continue;
      gst(p.match + t * p.tile_rows + lr, m[k]);
      // repeated keys: the slot's row count (inner joins only read it)
      const int32_t cs = m[k] >= 0 && p.J.rowlist && p.join_type == GZ_JOIN_INNER ? gld(p.J.count + m[k]) : 1;
      sum += (uint64_t)out_count(p, m[k], cs);
    }
  }
  uint64_t total;
  block_exclusive_scan<uint64_t>(sum, s_scan, &total);
  if (threadIdx.x == 0) p.tile_count[t] = (uint32_t)total;
}

// Value of a fixed-width output column for output row o (probe or build side).
__device__ __forceinline__ void store_value(const OutCol& oc, uint64_t o, uint64_t v, uint64_t v_hi) {
  const int w = oc.width;
  uint8_t* d = oc.ovalues + o * w;
  if (w == 8) gst((uint64_t*)d, v);
  else if (w == 4) gst((uint32_t*)d, (uint32_t)v);
  else if (w == 2) gst((uint16_t*)d, (uint16_t)v);
  else if (w == 1) gst(d, (uint8_t)v);
  else {
    gst((uint64_t*)d, v);
    gst((uint64_t*)d + 1, v_hi);
  }
}

// Write pass for tables whose probe rows emit at most one row each (inner
// joins without repeated keys, semi and anti joins).  Rows are lane-
// consecutive: wave ballots give each emitting row its place, so the stores of
// a wave land on consecutive output rows.  Build columns come from the payload
// (one 16-byte load per emitted row) when it holds them.
// SINGLE = 1: a call of one batch of at most kSingleMaxBlocks tiles probes and
// writes in this one launch: block b takes tile b, computes its match codes,
// posts its emitted-row count and sums the counts of the blocks before it (the
// lanes of one wave poll one earlier block each) for its output offset.  Blocks
// are dispatched in index order, so each wait is on a block already running;
// the polls are bounded all the same and a wait that runs out is reported.  A
// posted word carries its whole message, so its atomics are relaxed (agent-
// scope acquire / release would invalidate / write back the L2 at every poll).
template <int TILE, int BLOCK, int SINGLE>
__global__ __launch_bounds__(BLOCK) void join_probe_write01_kernel(ProbeParams p) {
  constexpr int RPT = TILE / BLOCK;
  constexpr int NW = BLOCK / kWave;
  constexpr int G = RPT < 4 ? RPT : 4;  // rows per thread whose gathers are in flight together
  static_assert(RPT % G == 0, "groups of G rounds");
  __shared__ uint32_t s_wt[RPT][NW];
  __shared__ uint32_t s_base;
  int64_t t;
  PTile td;
  if (SINGLE) {
    t = blockIdx.x;
    td = p.tile0;
    td.row_begin = (int32_t)(t * TILE);
    td.nrows = min(TILE, p.tile0.nrows - td.row_begin);
  } else {
    t = xcd_tile(blockIdx.x, gridDim.x, p.T);
    if (t < 0) return;
    td = p.tiles[t];
  }
  const int tid = threadIdx.x, w = tid / kWave, lane = __lane_id();
  const int n = td.nrows;
  const int64_t rb = td.row_begin;
  int32_t m[RPT];
  uint32_t pre[RPT];  // bit 31: the row emits; low bits: emitting rows before it in its wave and round
  if (SINGLE) {
    probe_match<RPT, BLOCK>(p, td, tid, m);
  } else {
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const int lr = k * BLOCK + tid;
      m[k] = lr < n ? gld(p.match + t * p.tile_rows + lr) : -1;
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const int lr = k * BLOCK + tid;
    const bool e = lr < n && out_count(p, m[k], 1) > 0;
    const uint64_t b = __ballot(e);
    pre[k] = (uint32_t)__popcll(b & lanemask_lt()) | (e ? 0x80000000u : 0u);
    if (lane == 0) s_wt[k][w] = (uint32_t)__popcll(b);
  }
  __syncthreads();
  uint64_t obase;
  if (SINGLE) {
    if (w == 0) {
      uint32_t rows = 0;
      for (int k = 0; k < RPT; k++)
        for (int x = 0; x < NW; x++) rows += s_wt[k][x];
      const int id = (int)t;
      if (lane == 0)
        __hip_atomic_store(p.look + id, (uint64_t)p.epoch << 32 | rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t before = 0;
      bool late = false;
      if (lane < id) {
        uint64_t v;
        int spins = 0;
        while (((v = __hip_atomic_load(p.look + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) !=
               p.epoch) {
          if (++spins > kLookSpins) {
            late = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        before = late ? 0u : (uint32_t)v;
      }
#pragma unroll
      for (int o = kWave / 2; o > 0; o >>= 1) before += __shfl_xor(before, o);
      const bool any_late = __ballot(late) != 0;
      if (lane == 0) {
        s_base = before;
        if (any_late) gst(p.host_out + 1, 1u);
        if (id == (int)gridDim.x - 1) gst(p.host_out, before + rows);
        __threadfence_system();
      }
    }
    __syncthreads();
    obase = s_base;
  } else {
    obase = p.tile_off[t];
  }
  uint32_t acc = 0;  // rows emitted by earlier rounds
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    uint32_t before = acc;
#pragma unroll
    for (int x = 0; x < NW; x++) {
      const uint32_t c = s_wt[k][x];
      if (x < w) before += c;
      acc += c;
    }
    pre[k] = (pre[k] & 0x80000000u) | (before + (pre[k] & 0x7fffffffu));
  }
  const bool inner = p.join_type == GZ_JOIN_INNER;
#pragma unroll
  for (int k0 = 0; k0 < RPT; k0 += G) {
    int32_t b[G];
    u32x4 pl[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      const int k = k0 + g;
      b[g] = -1;
      if (inner && (pre[k] >> 31))
        b[g] = m[k] < 0 ? (int32_t)p.b_special : p.match_rows || p.by_slot ? m[k] : gld(p.J.start + m[k]);
    }
#pragma unroll
    for (int g = 0; g < G; g++) pl[g] = p.payload && b[g] >= 0 ? gld(p.payload + b[g]) : u32x4{0u, 0u, 0u, 0u};
    if (p.write_pairs) {
#pragma unroll
      for (int g = 0; g < G; g++) {
        const int k = k0 + g;
        if (!(pre[k] >> 31)) continue;
        const uint64_t o = obase + (pre[k] & 0x7fffffffu);
        gst(p.pair_probe + o, (int64_t)((uint64_t)t << 32 | (uint32_t)(k * BLOCK + tid)));
        gst(p.pair_build + o, b[g]);
      }
    }
    for (int c = 0; c < p.ncols; c++) {
      const OutCol& oc = p.out[c];
      if (oc.binary) continue;
      if (oc.type == GZ_TYPE_NA) {
#pragma unroll
        for (int g = 0; g < G; g++)
          if (pre[k0 + g] >> 31) gst(oc.ovalid + obase + (pre[k0 + g] & 0x7fffffffu), (uint8_t)0);
        continue;
      }
      const int wd = oc.width;
      uint64_t v[G], vh[G];
      uint8_t ok[G];
      if (oc.side == 1) {
        const uint8_t* vb = (const uint8_t*)pbuf(p, td.batch, oc.col, 0);
        const uint8_t* vv = (const uint8_t*)pbuf(p, td.batch, oc.col, 2);
#pragma unroll
        for (int g = 0; g < G; g++) {
          const int k = k0 + g;
          const int64_t r = rb + k * BLOCK + tid;
          const bool e = pre[k] >> 31;
          vh[g] = 0;
          ok[g] = e && vb ? (uint8_t)bit_get(vb, r) : (uint8_t)1;
          if (!e) v[g] = 0;
          else if (oc.type == GZ_TYPE_BOOL) v[g] = (uint64_t)bit_get(vv, r);
          else if (wd == 8) v[g] = gld((const uint64_t*)vv + r);
          else if (wd == 4) v[g] = gld((const uint32_t*)vv + r);
          else if (wd == 2) v[g] = gld((const uint16_t*)vv + r);
          else if (wd == 1) v[g] = gld(vv + r);
          else {
            v[g] = gld((const uint64_t*)vv + 2 * r);
            vh[g] = gld((const uint64_t*)vv + 2 * r + 1);
          }
        }
      } else {
        const int off = p.pay_off[c];
#pragma unroll
        for (int g = 0; g < G; g++) {
          vh[g] = 0;
          v[g] = 0;
          ok[g] = 0;
          if (b[g] < 0) continue;  // semi / anti: build columns are null (AppendNull)
          if (off >= 0) {
            const uint64_t lo = (uint64_t)pl[g].x | ((uint64_t)pl[g].y << 32);
            const uint64_t hi = (uint64_t)pl[g].z;
            const uint64_t x = off < 8 ? (lo >> (8 * off)) | (off ? hi << (64 - 8 * off) : 0) : hi >> (8 * (off - 8));
            v[g] = wd == 8 ? x : x & ((1ull << (8 * wd)) - 1);
            ok[g] = (uint8_t)((pl[g].w >> c) & 1);
          } else {
            ok[g] = gld(oc.bvalid + b[g]);
            const uint8_t* sp = oc.bvalues + (int64_t)b[g] * wd;
            if (wd == 8) v[g] = gld((const uint64_t*)sp);
            else if (wd == 4) v[g] = gld((const uint32_t*)sp);
            else if (wd == 2) v[g] = gld((const uint16_t*)sp);
            else if (wd == 1) v[g] = gld(sp);
            else {
              v[g] = gld((const uint64_t*)sp);
              vh[g] = gld((const uint64_t*)sp + 1);
            }
          }
        }
      }
#pragma unroll
      for (int g = 0; g < G; g++) {
        const int k = k0 + g;
        if (!(pre[k] >> 31)) continue;
        const uint64_t o = obase + (pre[k] & 0x7fffffffu);
        gst(oc.ovalid + o, ok[g]);
        store_value(oc, o, v[g], vh[g]);
      }
    }
  }
}

// Write pass for inner joins over tables with repeated keys: a probe row emits
// all rows of its key in insertion order (rowlist).  Consecutive rows per
// thread, so one scan gives every output position.
template <int TILE, int BLOCK>
__global__ __launch_bounds__(BLOCK) void join_probe_write_kernel(ProbeParams p) {
  constexpr int RPT = TILE / BLOCK;  // consecutive rows per thread
  __shared__ uint64_t s_scan[BLOCK / kWave + 1];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const PTile td = p.tiles[t];
  int lr0 = threadIdx.x * RPT;
  int cnt[RPT];
  int32_t mm[RPT];
  uint64_t sum = 0;
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    int lr = lr0 + k;
    int32_t m = lr < td.nrows ? p.match[t * p.tile_rows + lr] : -1;
    mm[k] = m;
    cnt[k] = lr < td.nrows ? out_count(p, m, m >= 0 ? p.J.count[m] : 0) : 0;
    sum += (uint64_t)cnt[k];
  }
  uint64_t ex = block_exclusive_scan<uint64_t>(sum, s_scan, nullptr);
  uint64_t o = p.tile_off[t] + ex;
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    int lr = lr0 + k;
    int64_t r = (int64_t)td.row_begin + lr;
    int32_t m = mm[k];
    const int64_t st = m >= 0 ? p.J.start[m] : p.J.special_start;
    for (int j = 0; j < cnt[k]; j++, o++) {
      const int32_t b = p.J.rowlist[st + j];  // build row
      if (p.write_pairs) {
        p.pair_probe[o] = ((int64_t)t << 32) | (int64_t)lr;
        p.pair_build[o] = b;
      }
      for (int c = 0; c < p.ncols; c++) {
        const OutCol& oc = p.out[c];
        if (oc.binary || oc.type == GZ_TYPE_NA) {
          if (oc.type == GZ_TYPE_NA) oc.ovalid[o] = 0;
          continue;
        }
        const int w = oc.width;
        if (oc.side == 0) {
          oc.ovalid[o] = oc.bvalid[b];
          const uint8_t* s = oc.bvalues + (int64_t)b * w;
          uint8_t* d = oc.ovalues + o * w;
          if (w == 8) *(uint64_t*)d = *(const uint64_t*)s;
          else if (w == 4) *(uint32_t*)d = *(const uint32_t*)s;
          else for (int x = 0; x < w; x++) d[x] = s[x];
        } else {
          const uint8_t* vb = (const uint8_t*)pbuf(p, td.batch, oc.col, 0);
          const uint8_t* vv = (const uint8_t*)pbuf(p, td.batch, oc.col, 2);
          oc.ovalid[o] = vb ? (uint8_t)bit_get(vb, r) : 1;
          if (oc.type == GZ_TYPE_BOOL) {
            oc.ovalues[o] = (uint8_t)bit_get(vv, r);
          } else if (w == 8) {
            *(uint64_t*)(oc.ovalues + o * 8) = ((const uint64_t*)vv)[r];
          } else if (w == 4) {
            *(uint32_t*)(oc.ovalues + o * 4) = ((const uint32_t*)vv)[r];
          } else {
            for (int x = 0; x < w; x++) oc.ovalues[o * w + x] = vv[r * w + x];
          }
        }
      }
    }
  }
}

// utf8/binary output columns: lengths, then (after a scan) bytes.
struct BinGather {
  const PTile* tiles;
  const void* const* bufs;
  int bstride;
  int side, col;
  const int64_t* pair_probe;
  const int32_t* pair_build;
  int64_t n;
  const uint8_t* bvalid;
  const int64_t* boff;
  const uint8_t* bdata;
  uint8_t* ovalid;
  int64_t* olen;     // pass 1: lengths; after the scan: start offsets
  int32_t* ooff;     // output int32 offsets (n + 1)
  uint8_t* odata;
};

__device__ __forceinline__ void bin_src(const BinGather& g, int64_t i, const uint8_t** src, int64_t* len,
                                        uint8_t* valid) {
  if (g.side == 0) {
    int32_t b = g.pair_build[i];
    if (b < 0) { *src = nullptr; *len = 0; *valid = 0; return; }
    *valid = g.bvalid[b];
    *src = g.bdata + g.boff[b];
    *len = g.boff[b + 1] - g.boff[b];
  } else {
    int64_t pp = g.pair_probe[i];
    int t = (int)(pp >> 32), lr = (int)(pp & 0xffffffff);
    PTile td = g.tiles[t];
    int64_t r = (int64_t)td.row_begin + lr;
    const uint8_t* vb = (const uint8_t*)g.bufs[(int64_t)td.batch * g.bstride + g.col * 3 + 0];
    const int32_t* off = (const int32_t*)g.bufs[(int64_t)td.batch * g.bstride + g.col * 3 + 1];
    const uint8_t* data = (const uint8_t*)g.bufs[(int64_t)td.batch * g.bstride + g.col * 3 + 2];
    *valid = vb ? (uint8_t)bit_get(vb, r) : 1;
    *src = data + off[r];
    *len = off[r + 1] - off[r];
  }
}

__global__ void join_bin_len_kernel(BinGather g) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t* s;
    int64_t len;
    uint8_t v;
    bin_src(g, i, &s, &len, &v);
    g.olen[i] = len;
    g.ovalid[i] = v;
  }
}

__global__ void join_bin_copy_kernel(BinGather g) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t* s;
    int64_t len;
    uint8_t v;
    bin_src(g, i, &s, &len, &v);
    int64_t st = g.olen[i];
    g.ooff[i] = (int32_t)st;
    if (i == g.n - 1) g.ooff[g.n] = (int32_t)(st + len);
    for (int64_t x = 0; x < len; x++) g.odata[st + x] = s[x];
  }
}

__global__ void join_fill_i64_kernel(int64_t* p, int64_t n, int64_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

}  // namespace gz

// =============================================================== host side
using namespace gz;

struct ColStore {
  int type = 0, width = 0;
  DBuf valid, values, offsets, data;
  int64_t data_bytes = 0, data_cap = 0;
};

struct gz_relation {
  gz_context* ctx = nullptr;
  int ncols = 0;
  std::vector<int> types;
  int key_col = 0, builder_type = 1;
  bool normalize = false;  // Spark's normalize on the (one) float/double key
  std::vector<ColStore> cols;
  int64_t n = 0, cap_rows = 0;
  // table
  DBuf keys, start, count, rowlist, row_slot, flags, sort_a, sort_b, sort_tmp;
  uint64_t cap = 0;
  bool finished = false, has_null = false, has_dups = false;
  int64_t nkeys = 0, special_start = 0, special_count = 0;
  // several key columns (or a utf8 key): rows are keyed by a dictionary code
  // appended as the last column
  KeyDict* kd = nullptr;
  std::vector<int> kcols, otypes;  // kcols in the dictionary's key order
  std::vector<int> kperm;          // dictionary position -> the caller's key index
  int ocols = 0;
  KeyScratch ksc;
  DBuf d_atiles, d_abufs;  // tile and buffer tables of a batched append
};

static int rel_grow(gz_relation* r, int64_t need) {
  if (need <= r->cap_rows) return GZ_OK;
  gz_context* ctx = r->ctx;
  int64_t nc = std::max<int64_t>(need, std::max<int64_t>(r->cap_rows * 2, 4096));
  for (auto& c : r->cols) {
    if (c.type == GZ_TYPE_NA) continue;
    DBuf nv, nval, noff;
    GZ_TRY(ctx_alloc(ctx, (size_t)nc, &nv));
    if (r->n) GZ_HIP(hipMemcpyAsync(nv.p, c.valid.p, (size_t)r->n, hipMemcpyDeviceToDevice, ctx->stream));
    ctx_free(ctx, &c.valid);
    c.valid = nv;
    if (type_is_binary(c.type)) {
      GZ_TRY(ctx_alloc(ctx, sizeof(int64_t) * (nc + 1), &noff));
      if (r->n)
        GZ_HIP(hipMemcpyAsync(noff.p, c.offsets.p, sizeof(int64_t) * (r->n + 1), hipMemcpyDeviceToDevice,
                              ctx->stream));
      else
        GZ_HIP(hipMemsetAsync(noff.p, 0, sizeof(int64_t), ctx->stream));
      ctx_free(ctx, &c.offsets);
      c.offsets = noff;
    } else {
      int w = c.type == GZ_TYPE_BOOL ? 1 : c.width;
      GZ_TRY(ctx_alloc(ctx, (size_t)nc * w, &nval));
      if (r->n) GZ_HIP(hipMemcpyAsync(nval.p, c.values.p, (size_t)r->n * w, hipMemcpyDeviceToDevice, ctx->stream));
      ctx_free(ctx, &c.values);
      c.values = nval;
    }
  }
  r->cap_rows = nc;
  return GZ_OK;
}

static int rel_append_one(gz_relation* r, const gz_batch& bt) {
  gz_context* ctx = r->ctx;
  if (r->finished) return fail(GZ_INVALID, "relation already finished");
  if (bt.num_rows < 0) return fail(GZ_INVALID, "bad num_rows");
  if (bt.num_rows == 0) return GZ_OK;
  // ArrayItemIndex limits (array_item_index.h:24-29) do not apply: rows are int32
  if (r->n + bt.num_rows > INT32_MAX) return fail(GZ_CAPACITY_ERROR, "build side exceeds 2^31 rows");
  int need = 0;
  for (int c = 0; c < r->ncols; c++) need += type_num_bufs(r->types[c]);
  if (bt.num_bufs < need) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
  GZ_TRY(rel_grow(r, r->n + bt.num_rows));
  AppendParams ap;
  memset(&ap, 0, sizeof(ap));
  ap.ncols = r->ncols;
  ap.n = bt.num_rows;
  ap.row0 = r->n;
  int k = 0;
  for (int c = 0; c < r->ncols; c++) {
    ColStore& cs = r->cols[c];
    AppendCol& a = ap.col[c];
    a.type = cs.type;
    a.width = cs.width;
    a.valid = (bt.buf_addrs[k] && bt.buf_sizes[k] > 0) ? (const uint8_t*)bt.buf_addrs[k] : nullptr;
    if (a.valid && bt.buf_sizes[k] < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
    k++;
    if (type_is_binary(cs.type)) {
      if (bt.buf_sizes[k] < (bt.num_rows + 1) * 4) return fail(GZ_INVALID, "offsets buffer too small");
      a.offsets = (const int32_t*)bt.buf_addrs[k++];
      // bytes used by the batch: read both ends of its offsets
      int32_t ends[2];
      GZ_HIP(hipMemcpyAsync(&ends[0], a.offsets, 4, hipMemcpyDeviceToHost, ctx->stream));
      GZ_HIP(hipMemcpyAsync(&ends[1], a.offsets + bt.num_rows, 4, hipMemcpyDeviceToHost, ctx->stream));
      GZ_HIP(hipStreamSynchronize(ctx->stream));
      int64_t nb = (int64_t)ends[1] - ends[0];
      a.values = (const void*)bt.buf_addrs[k];
      if (cs.data_bytes + nb > cs.data_cap) {
        int64_t ncap = std::max<int64_t>(cs.data_bytes + nb, cs.data_cap * 2 + 4096);
        DBuf nd;
        GZ_TRY(ctx_alloc(ctx, (size_t)ncap, &nd));
        if (cs.data_bytes)
          GZ_HIP(hipMemcpyAsync(nd.p, cs.data.p, (size_t)cs.data_bytes, hipMemcpyDeviceToDevice, ctx->stream));
        ctx_free(ctx, &cs.data);
        cs.data = nd;
        cs.data_cap = ncap;
      }
      if (nb > 0 && !a.values) return fail(GZ_INVALID, "missing values buffer");
      if (nb > 0)
        GZ_HIP(hipMemcpyAsync((uint8_t*)cs.data.p + cs.data_bytes, (const uint8_t*)a.values + ends[0], (size_t)nb,
                              hipMemcpyDeviceToDevice, ctx->stream));
      a.data_base = cs.data_bytes;
      cs.data_bytes += nb;
      a.doffsets = (int64_t*)cs.offsets.p;
    } else {
      int64_t need_b = cs.type == GZ_TYPE_BOOL ? (bt.num_rows + 7) / 8 : bt.num_rows * cs.width;
      if (cs.type != GZ_TYPE_NA && bt.buf_sizes[k] < need_b) return fail(GZ_INVALID, "values buffer too small");
      a.values = (const void*)bt.buf_addrs[k];
      a.dvalues = (uint8_t*)cs.values.p;
    }
    k++;
    a.dvalid = (uint8_t*)cs.valid.p;
    // (utf8 bytes may be absent when every string is empty)
    if (cs.type != GZ_TYPE_NA && !type_is_binary(cs.type) && !a.values) return fail(GZ_INVALID, "missing values buffer");
  }
  unsigned blocks = (unsigned)std::min<int64_t>((bt.num_rows + 255) / 256, 2048);
  {
    KernelTimer kt(ctx, "join_append");
    hipLaunchKernelGGL(join_append_kernel, dim3(blocks), dim3(256), 0, ctx->stream, ap);
  }
  GZ_HIP(hipGetLastError());
  r->n += bt.num_rows;
  return GZ_OK;
}

// Many batches of a relation without utf8 columns: checked on the host, copied
// by one launch over a tile table.
static int rel_append_many(gz_relation* r, int32_t nb, const gz_batch* batches) {
  gz_context* ctx = r->ctx;
  if (r->finished) return fail(GZ_INVALID, "relation already finished");
  const int nc = r->ncols;
  int need = 0;
  for (int c = 0; c < nc; c++) need += type_num_bufs(r->types[c]);
  std::vector<const void*> bufs((size_t)std::max(nb, 1) * nc * 2, nullptr);
  std::vector<AppendTile> tiles;
  int64_t row0 = r->n;
  for (int b = 0; b < nb; b++) {
    const gz_batch& bt = batches[b];
    if (bt.num_rows < 0) return fail(GZ_INVALID, "bad num_rows");
    if (bt.num_rows == 0) continue;
    if (!bt.buf_addrs || !bt.buf_sizes) return fail(GZ_INVALID, "buf_addrs can't be null");
    if (row0 + bt.num_rows > INT32_MAX) return fail(GZ_CAPACITY_ERROR, "build side exceeds 2^31 rows");
    if (bt.num_bufs < need) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
    for (int c = 0, k = 0; c < nc; c++, k += 2) {
      const int t = r->types[c];
      const void* vb = (bt.buf_addrs[k] && bt.buf_sizes[k] > 0) ? (const void*)bt.buf_addrs[k] : nullptr;
      if (vb && bt.buf_sizes[k] < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
      if (t != GZ_TYPE_NA) {
        const int64_t need_b = t == GZ_TYPE_BOOL ? (bt.num_rows + 7) / 8 : bt.num_rows * type_width(t);
        if (!bt.buf_addrs[k + 1]) return fail(GZ_INVALID, "missing values buffer");
        if (bt.buf_sizes[k + 1] < need_b) return fail(GZ_INVALID, "values buffer too small");
      }
      bufs[((size_t)b * nc + c) * 2] = vb;
      bufs[((size_t)b * nc + c) * 2 + 1] = (const void*)bt.buf_addrs[k + 1];
    }
    for (int64_t r0 = 0; r0 < bt.num_rows; r0 += 1024)
      tiles.push_back(AppendTile{b, (int32_t)r0, (int32_t)std::min<int64_t>(1024, bt.num_rows - r0), 0, row0 + r0});
    row0 += bt.num_rows;
  }
  if (tiles.empty()) return GZ_OK;
  GZ_TRY(rel_grow(r, row0));
  AppendManyParams ap;
  memset(&ap, 0, sizeof(ap));
  ap.ncols = nc;
  for (int c = 0; c < nc; c++) {
    ColStore& cs = r->cols[c];
    ap.type[c] = cs.type;
    ap.width[c] = cs.width;
    ap.dvalid[c] = (uint8_t*)cs.valid.p;
    ap.dvalues[c] = (uint8_t*)cs.values.p;
  }
  GZ_TRY(upload(ctx, &r->d_atiles, tiles.data(), sizeof(AppendTile) * tiles.size()));
  GZ_TRY(upload(ctx, &r->d_abufs, bufs.data(), sizeof(void*) * bufs.size()));
  ap.tiles = (const AppendTile*)r->d_atiles.p;
  ap.bufs = (const void* const*)r->d_abufs.p;
  {
    KernelTimer kt(ctx, "join_append");
    hipLaunchKernelGGL(join_append_many_kernel, dim3((unsigned)tiles.size()), dim3(256), 0, ctx->stream, ap);
  }
  GZ_HIP(hipGetLastError());
  r->n = row0;
  return GZ_OK;
}

extern "C" {

int gz_relation_make(gz_context* ctx, int32_t num_cols, const int32_t* col_types, int32_t key_col,
                     int32_t builder_type, gz_relation** out) {
  if (!ctx || !out || (num_cols > 0 && !col_types)) return fail(GZ_INVALID, "gz_relation_make: null argument");
  if (num_cols > kJoinMaxCols) return fail(GZ_NOT_IMPLEMENTED, "more than 32 build columns");
  if (key_col < 0 || key_col >= num_cols) return fail(GZ_INVALID, "key column out of range");
  if (builder_type != 1 && builder_type != 3)
    return fail(GZ_NOT_IMPLEMENTED, "builder_type " + std::to_string(builder_type) + " is not supported");
  for (int c = 0; c < num_cols; c++)
    if (!type_supported(col_types[c]))
      return fail(GZ_NOT_IMPLEMENTED, "column type " + std::to_string(col_types[c]));
  int kt = col_types[key_col];
  if (type_is_binary(kt) || kt == GZ_TYPE_DECIMAL128 || kt == GZ_TYPE_NA)
    return fail(GZ_NOT_IMPLEMENTED, "HashRelation Evaluate doesn't support single key type " + std::to_string(kt));
  auto* r = new gz_relation();
  r->ctx = ctx;
  r->ncols = num_cols;
  r->types.assign(col_types, col_types + num_cols);
  r->key_col = key_col;
  r->builder_type = builder_type;
  r->cols.resize(num_cols);
  for (int c = 0; c < num_cols; c++) {
    r->cols[c].type = col_types[c];
    r->cols[c].width = type_width(col_types[c]);
  }
  *out = r;
  return GZ_OK;
}

int gz_relation_append_batches(gz_relation* r, int32_t num_batches, const gz_batch* batches);

int gz_relation_append(gz_relation* r, int64_t num_rows, const int64_t* buf_addrs,
                       const int64_t* buf_sizes, int32_t num_bufs) {
  if (!r) return fail(GZ_INVALID, "null relation");
  if (!buf_addrs || !buf_sizes) return fail(GZ_INVALID, "buf_addrs can't be null");
  gz_batch b{num_rows, buf_addrs, buf_sizes, num_bufs};
  if (r->kd) return gz_relation_append_batches(r, 1, &b);
  return rel_append_one(r, b);
}

int gz_relation_append_batches(gz_relation* r, int32_t num_batches, const gz_batch* batches) {
  if (!r) return fail(GZ_INVALID, "null relation");
  std::vector<int64_t> kaddrs, ksizes;
  std::vector<gz_batch> kbatches;
  if (r->kd) {
    // AppendKeyColumn with UnsafeRow keys (hash_relation.h:176-198): a row with
    // a null key is never inserted, so it gets a code no probe row carries
    if (r->finished) return fail(GZ_INVALID, "relation already finished");
    const int64_t* codes = nullptr;
    int64_t rows = 0;
    GZ_TRY(kd_encode(r->kd, r->ctx, &r->ksc, r->ocols, r->otypes.data(), r->kcols.data(), num_batches, batches, 1,
                     &codes, &rows));
    // one utf8 key: a null key is a null code, recorded as the relation's null
    // key (InsertNull, hash_relation.h:244-270)
    with_code_column(r->ocols, r->otypes.data(), num_batches, batches, codes, kaddrs, ksizes, kbatches,
                     r->kcols.size() == 1 ? r->kcols[0] : -1);
    batches = kbatches.data();
  }
  bool fixed = num_batches > 1;
  for (int c = 0; c < r->ncols && fixed; c++)
    if (type_is_binary(r->types[c])) fixed = false;
  if (fixed) return rel_append_many(r, num_batches, batches);
  // reserve once so the column store grows a single time
  int64_t total = r->n;
  for (int i = 0; i < num_batches; i++) total += batches[i].num_rows;
  GZ_TRY(rel_grow(r, total));
  for (int i = 0; i < num_batches; i++) GZ_TRY(rel_append_one(r, batches[i]));
  return GZ_OK;
}

int gz_relation_finish(gz_relation* r) {
  if (!r) return fail(GZ_INVALID, "null relation");
  if (r->finished) return GZ_OK;
  gz_context* ctx = r->ctx;
  hipStream_t st = ctx->stream;
  GZ_HIP(hipSetDevice(ctx->device));
  const int64_t n = r->n;
  uint64_t cap = 128;
  if (n > 32) {
    cap = 1;
    while (cap < (uint64_t)n) cap <<= 1;
    cap <<= 1;  // 2^(ceil(log2 rows) + 1)
  }
  r->cap = cap;
  GZ_TRY(ctx_alloc(ctx, sizeof(int64_t) * cap, &r->keys));
  GZ_TRY(ctx_alloc(ctx, sizeof(int32_t) * cap, &r->start));
  GZ_TRY(ctx_alloc(ctx, sizeof(int32_t) * cap, &r->count));
  GZ_TRY(ctx_alloc(ctx, sizeof(int32_t) * std::max<int64_t>(n, 1), &r->row_slot));
  GZ_TRY(ctx_alloc(ctx, sizeof(unsigned long long) * 8, &r->flags));
  unsigned gc = (unsigned)std::min<uint64_t>((cap + 255) / 256, 8192);
  hipLaunchKernelGGL(join_table_init_kernel, dim3(gc), dim3(256), 0, st, (int64_t*)r->keys.p,
                     (int32_t*)r->start.p, (int32_t*)r->count.p, cap);
  unsigned long long hflags[8] = {0, 0, 0, 0, (unsigned long long)INT64_MAX, 0, 0, 0};
  GZ_HIP(hipMemcpyAsync(r->flags.p, hflags, sizeof(hflags), hipMemcpyHostToDevice, st));
  const ColStore& kc = r->cols[r->key_col];
  if (n > 0) {
    KernelTimer kt(ctx, "join_build");
    unsigned gb = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(join_build_kernel, dim3(gb), dim3(256), 0, st, (const uint8_t*)kc.valid.p,
                       (const uint8_t*)kc.values.p, kc.type, kc.width, n, (int64_t*)r->keys.p,
                       (int32_t*)r->start.p, (int32_t*)r->count.p, cap, (int32_t*)r->row_slot.p,
                       r->builder_type == 3 ? 1 : 0, r->kd ? 1 : 0, r->normalize ? 1 : 0,
                       (unsigned long long*)r->flags.p);
  }
  GZ_HIP(hipGetLastError());
  GZ_HIP(hipMemcpyAsync(hflags, r->flags.p, sizeof(hflags), hipMemcpyDeviceToHost, st));
  GZ_HIP(hipStreamSynchronize(st));
  r->has_null = hflags[0] != 0;
  r->nkeys = (int64_t)hflags[2] + (hflags[3] ? 1 : 0);
  r->special_count = (int64_t)hflags[3];
  r->special_start = (int64_t)hflags[4];
  bool dups = hflags[1] != 0;
  if (r->builder_type == 3) {
    // InsertSkipDup keeps the first row of every key
    hipLaunchKernelGGL(join_clamp_counts_kernel, dim3(gc), dim3(256), 0, st, (int32_t*)r->count.p, cap);
    if (r->special_count > 1) r->special_count = 1;
    dups = false;
  }
  r->has_dups = dups;
  if (dups) {
    // the build counted inserts only: count every row of its slot
    hipLaunchKernelGGL(join_zero_i32_kernel, dim3(gc), dim3(256), 0, st, (int32_t*)r->count.p, cap);
    {
      unsigned gr = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
      hipLaunchKernelGGL(join_recount_kernel, dim3(gr), dim3(256), 0, st, (const int32_t*)r->row_slot.p, n,
                         (int32_t*)r->count.p);
    }
    // stable (slot, row) order of all keyed rows, then each slot's start in it
    GZ_TRY(ctx_alloc(ctx, sizeof(unsigned long long) * n, &r->sort_a));
    GZ_TRY(ctx_alloc(ctx, sizeof(unsigned long long) * n, &r->sort_b));
    GZ_TRY(ctx_alloc(ctx, sizeof(int32_t) * n, &r->rowlist));
    unsigned gb = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(join_pair_keys_kernel, dim3(gb), dim3(256), 0, st, (const int32_t*)r->row_slot.p, n,
                       (unsigned long long*)r->sort_a.p);
    size_t tmp = 0;
    GZ_HIP(rocprim::radix_sort_keys(nullptr, tmp, (unsigned long long*)r->sort_a.p,
                                    (unsigned long long*)r->sort_b.p, (size_t)n, 0, 64, st));
    GZ_TRY(ctx_ensure(ctx, &r->sort_tmp, tmp));
    GZ_HIP(rocprim::radix_sort_keys(r->sort_tmp.p, tmp, (unsigned long long*)r->sort_a.p,
                                    (unsigned long long*)r->sort_b.p, (size_t)n, 0, 64, st));
    hipLaunchKernelGGL(join_rowlist_kernel, dim3(gb), dim3(256), 0, st, (const unsigned long long*)r->sort_b.p, n,
                       (int32_t*)r->rowlist.p);
    // start[slot] = exclusive scan of counts in slot order (the sort order)
    size_t tmp2 = 0;
    GZ_HIP(rocprim::exclusive_scan(nullptr, tmp2, (const int32_t*)r->count.p, (int32_t*)r->start.p, 0,
                                   (size_t)cap, rocprim::plus<int32_t>(), st));
    GZ_TRY(ctx_ensure(ctx, &r->sort_tmp, tmp2));
    GZ_HIP(rocprim::exclusive_scan(r->sort_tmp.p, tmp2, (const int32_t*)r->count.p, (int32_t*)r->start.p, 0,
                                   (size_t)cap, rocprim::plus<int32_t>(), st));
    // the kJEmpty key sorts after every slot: its rows start after all slot rows
    int64_t slot_rows = 0;
    if (r->special_count) {
      int32_t last[2];
      GZ_HIP(hipMemcpyAsync(&last[0], (int32_t*)r->start.p + cap - 1, 4, hipMemcpyDeviceToHost, st));
      GZ_HIP(hipMemcpyAsync(&last[1], (int32_t*)r->count.p + cap - 1, 4, hipMemcpyDeviceToHost, st));
      GZ_HIP(hipStreamSynchronize(st));
      slot_rows = (int64_t)last[0] + last[1];
      r->special_start = slot_rows;
    }
    ctx_free(ctx, &r->sort_a);
    ctx_free(ctx, &r->sort_b);
  }
  GZ_HIP(hipGetLastError());
  r->finished = true;
  return GZ_OK;
}

int gz_relation_num_keys(gz_relation* r, int64_t* out) {
  if (!r || !out) return fail(GZ_INVALID, "null argument");
  if (!r->finished) return fail(GZ_INVALID, "relation not finished");
  *out = r->nkeys;
  return GZ_OK;
}

int gz_relation_has_null_key(gz_relation* r, int32_t* out) {
  if (!r || !out) return fail(GZ_INVALID, "null argument");
  if (!r->finished) return fail(GZ_INVALID, "relation not finished");
  *out = r->has_null ? 1 : 0;
  return GZ_OK;
}

int gz_relation_release(gz_relation* r) {
  if (!r) return GZ_OK;
  gz_context* ctx = r->ctx;
  kd_release(r->kd);
  kd_scratch_free(ctx, &r->ksc);
  ctx_free(ctx, &r->d_atiles);
  ctx_free(ctx, &r->d_abufs);
  for (auto& c : r->cols)
    for (DBuf* b : {&c.valid, &c.values, &c.offsets, &c.data}) ctx_free(ctx, b);
  for (DBuf* b : {&r->keys, &r->start, &r->count, &r->rowlist, &r->row_slot, &r->flags, &r->sort_a,
                  &r->sort_b, &r->sort_tmp})
    ctx_free(ctx, b);
  delete r;
  return GZ_OK;
}

}  // extern "C"

struct gz_probe {
  gz_relation* rel = nullptr;
  gz_context* ctx = nullptr;  // stream and memory of the probe (the relation's by default)
  int nprobe = 0;
  std::vector<int> ptypes;
  int key_col = 0, join_type = 0, null_aware = 0;
  std::vector<int> sides, idx;
  int tile_rows = 4096;
  DBuf d_tiles, d_bufs, match, tile_count, tile_off, scan_tmp, pair_probe, pair_build;
  // one-batch calls: per-block counts tagged with the call's epoch, and the
  // pinned host words the kernel writes the row total to
  DBuf look;
  uint32_t epoch = 0;
  uint32_t* host_out = nullptr;
  uint32_t* host_out_dev = nullptr;
  bool used = false;  // a process call ran (its buffers came from ctx)
  // row-major payload of the fixed-width build output columns (built at the
  // first probe call when they fit a 16-byte record)
  DBuf payload;
  bool payload_ready = false;
  std::vector<int> pay_off;  // per output column: byte offset in the record, -1 not in it
  int pay_by_slot = -1;      // the payload is stored by slot (-1: not decided yet)
  struct Out {
    int type = 0, width = 0;
    DBuf values, valid, offsets, data, lens, bits;  // bits: Arrow bitmap, packed on request
  };
  std::vector<Out> outs;
  int64_t out_rows = 0;
  // several key columns: probe rows are looked up in the relation's dictionary
  KeyDict* kd = nullptr;
  std::vector<int> kcols, otypes;
  int ocols = 0;
  KeyScratch ksc;
};

// Probe batches of a multi-key probe with their code column appended.
static int probe_codes(gz_probe* h, int32_t nb, const gz_batch* batches, std::vector<int64_t>& addrs,
                       std::vector<int64_t>& sizes, std::vector<gz_batch>& out) {
  h->used = true;
  const int64_t* codes = nullptr;
  int64_t rows = 0;
  GZ_TRY(kd_encode(h->kd, h->ctx, &h->ksc, h->ocols, h->otypes.data(), h->kcols.data(), nb, batches, 2, &codes,
                   &rows));
  with_code_column(h->ocols, h->otypes.data(), nb, batches, codes, addrs, sizes, out,
                   h->kcols.size() == 1 ? h->kcols[0] : -1);
  return GZ_OK;
}

// Payload layout of a probe handle's build output columns: fixed-width ones of
// at most 8 bytes packed by decreasing width into 12 bytes; off[c] is output
// column c's byte offset, -1 when not in the record.  False (every off -1) when
// there is no such column or they do not fit.
static bool payload_layout(const gz_probe* h, std::vector<int>* off) {
  const int nres = (int)h->sides.size();
  off->assign(nres, -1);
  std::vector<int> order;
  auto width = [&](int c) { return h->outs[c].type == GZ_TYPE_BOOL ? 1 : type_width(h->outs[c].type); };
  for (int c = 0; c < nres; c++) {
    const int t = h->outs[c].type;
    if (h->sides[c] == 0 && !type_is_binary(t) && t != GZ_TYPE_NA && width(c) <= 8) order.push_back(c);
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return width(a) > width(b); });
  int o = 0;
  for (int c : order) {
    const int wd = width(c);
    o = (o + wd - 1) / wd * wd;
    if (o + wd > 12) {
      off->assign(nres, -1);
      return false;
    }
    (*off)[c] = o;
    o += wd;
  }
  return !order.empty();
}

static int probe_run(gz_probe* h, int32_t nb, const gz_batch* batches, int64_t* out_rows,
                     int64_t* out_batch_rows) {
  gz_relation* r = h->rel;
  gz_context* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  GZ_HIP(hipSetDevice(ctx->device));
  h->used = true;
  const int TILE = h->tile_rows;
  const int nc = h->nprobe;
  std::vector<PTile> tiles;
  std::vector<const void*> bufs((size_t)std::max(nb, 1) * nc * 3, nullptr);
  for (int bi = 0; bi < nb; bi++) {
    const gz_batch& bt = batches[bi];
    if (bt.num_rows < 0 || bt.num_rows > INT32_MAX) return fail(GZ_INVALID, "bad num_rows");
    int need = 0;
    for (int c = 0; c < nc; c++) need += type_num_bufs(h->ptypes[c]);
    if (bt.num_bufs < need) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
    int k = 0;
    for (int c = 0; c < nc; c++) {
      const void** slot = &bufs[((size_t)bi * nc + c) * 3];
      slot[0] = (bt.buf_addrs[k] && bt.buf_sizes[k] > 0) ? (const void*)bt.buf_addrs[k] : nullptr;
      if (slot[0] && bt.buf_sizes[k] < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
      k++;
      if (type_is_binary(h->ptypes[c])) {
        if (bt.buf_sizes[k] < (bt.num_rows + 1) * 4) return fail(GZ_INVALID, "offsets buffer too small");
        slot[1] = (const void*)bt.buf_addrs[k++];
      }
      int t = h->ptypes[c];
      int64_t need_b = t == GZ_TYPE_BOOL ? (bt.num_rows + 7) / 8 : bt.num_rows * type_width(t);
      if (t != GZ_TYPE_NA && !type_is_binary(t) && bt.buf_sizes[k] < need_b)
        return fail(GZ_INVALID, "values buffer too small");
      slot[2] = (const void*)bt.buf_addrs[k++];
      // utf8 bytes may be absent when every string is empty
      if (h->ptypes[c] != GZ_TYPE_NA && bt.num_rows > 0 && !slot[2] && !(type_is_binary(t) && bt.buf_sizes[k - 1] <= 0))
        return fail(GZ_INVALID, "missing values buffer");
    }
    for (int64_t r0 = 0; r0 < bt.num_rows; r0 += TILE) {
      PTile t;
      t.batch = bi;
      t.row_begin = (int32_t)r0;
      t.nrows = (int32_t)std::min<int64_t>(TILE, bt.num_rows - r0);
      t.pad = 0;
      tiles.push_back(t);
    }
  }
  const int64_t T = (int64_t)tiles.size();
  ProbeParams p;
  memset(&p, 0, sizeof(p));
  p.T = T;
  p.bstride = nc * 3;
  p.key_col = h->key_col;
  p.key_type = h->ptypes[h->key_col];
  p.join_type = h->join_type;
  p.null_aware = h->null_aware;
  p.normalize = r->normalize ? 1 : 0;
  p.build_has_null = r->has_null;
  p.build_keys = r->nkeys;
  p.tile_rows = TILE;
  p.J.keys = (const int64_t*)r->keys.p;
  p.J.start = (const int32_t*)r->start.p;
  p.J.count = (const int32_t*)r->count.p;
  p.J.rowlist = r->has_dups ? (const int32_t*)r->rowlist.p : nullptr;
  p.J.cap = r->cap;
  p.J.special_start = r->special_start;
  p.J.special_count = r->special_count;
  // an inner join over unique keys whose build outputs all fit the payload
  // reads the payload by slot: its count pass keeps the slot as match code
  if (h->pay_by_slot < 0) {
    std::vector<int> off;
    bool all = payload_layout(h, &off) && p.join_type == GZ_JOIN_INNER && !p.J.rowlist;
    for (size_t c = 0; c < h->sides.size() && all; c++) {
      const int t = h->outs[c].type;
      if (type_is_binary(t) || (h->sides[c] == 0 && t != GZ_TYPE_NA && off[c] < 0)) all = false;
    }
    h->pay_by_slot = all ? 1 : 0;
  }
  p.by_slot = h->pay_by_slot;
  p.match_rows = p.join_type == GZ_JOIN_INNER && !p.J.rowlist && !p.by_slot ? 1 : 0;
  p.b_special = p.by_slot ? (int64_t)r->cap : r->special_start;
  std::vector<uint32_t> tc(std::max<int64_t>(T, 1), 0);
  bool tc_device = false;  // the per-tile output counts are on the device only
  uint64_t total = 0;
  // a call of one small batch whose probe rows emit at most one row each,
  // without utf8 outputs: one launch probes and writes (outputs sized for the
  // batch), with no uploads and no copy back
  bool any_bin_out = false;
  for (auto& o : h->outs) any_bin_out |= type_is_binary(o.type);
  const int64_t rows0 = nb == 1 ? batches[0].num_rows : 0;
  const int64_t sblocks = (rows0 + kSingleTile - 1) / kSingleTile;
  const bool at_most_one = p.join_type != GZ_JOIN_INNER || !p.J.rowlist;
  const bool single = nb == 1 && rows0 > 0 && sblocks <= kSingleMaxBlocks && nc * 3 <= kProbeInlineBufs &&
                      at_most_one && !any_bin_out;
  if (single) {
    if (!h->look.p) {
      GZ_TRY(ctx_ensure(ctx, &h->look, sizeof(uint64_t) * kSingleMaxBlocks));
      hipLaunchKernelGGL(join_zero_i32_kernel, dim3(1), dim3(256), 0, st, (int32_t*)h->look.p,
                         (uint64_t)(2 * kSingleMaxBlocks));
      h->epoch = 0;
    }
    if (!h->host_out) {
      GZ_HIP(hipHostMalloc((void**)&h->host_out, 2 * sizeof(uint32_t), hipHostMallocDefault));
      GZ_HIP(hipHostGetDevicePointer((void**)&h->host_out_dev, h->host_out, 0));
    }
    if (++h->epoch == 0) {  // a fresh epoch never matches a count a block posted before
      hipLaunchKernelGGL(join_zero_i32_kernel, dim3(1), dim3(256), 0, st, (int32_t*)h->look.p,
                         (uint64_t)(2 * kSingleMaxBlocks));
      h->epoch = 1;
    }
    p.tile0 = PTile{0, 0, (int32_t)rows0, 0};
    p.look = (uint64_t*)h->look.p;
    p.epoch = h->epoch;
    p.host_out = h->host_out_dev;
    for (int i = 0; i < nc * 3; i++) p.ibufs[i] = bufs[i];
    p.bufs = nullptr;
    p.tiles = nullptr;
    ((volatile uint32_t*)h->host_out)[0] = 0xffffffffu;
    ((volatile uint32_t*)h->host_out)[1] = 0;
    total = (uint64_t)rows0;  // output rows the columns are sized for
  } else if (T > 0) {
    GZ_TRY(upload(ctx, &h->d_tiles, tiles.data(), sizeof(PTile) * T));
    GZ_TRY(upload(ctx, &h->d_bufs, bufs.data(), sizeof(void*) * bufs.size()));
    GZ_TRY(ctx_ensure(ctx, &h->match, sizeof(int32_t) * T * TILE));
    GZ_TRY(ctx_ensure(ctx, &h->tile_count, sizeof(uint32_t) * T));
    GZ_TRY(ctx_ensure(ctx, &h->tile_off, sizeof(uint64_t) * (T + 1)));
    p.tiles = (const PTile*)h->d_tiles.p;
    p.bufs = (const void* const*)h->d_bufs.p;
    p.match = (int32_t*)h->match.p;
    p.tile_count = (uint32_t*)h->tile_count.p;
    p.tile_off = (const uint64_t*)h->tile_off.p;
    int64_t grid = (T + 7) / 8 * 8;
    {
      KernelTimer kt(ctx, "join_probe_count");
      hipLaunchKernelGGL((join_probe_count_kernel<256, 8>), dim3((unsigned)grid), dim3(256), 0, st, p);
    }
    // tile offsets: exclusive scan (uint32 counts -> uint64 offsets)
    size_t tmp = 0;
    GZ_HIP(rocprim::exclusive_scan(nullptr, tmp, (const uint32_t*)h->tile_count.p, (uint64_t*)h->tile_off.p,
                                   (uint64_t)0, (size_t)T + 0, rocprim::plus<uint64_t>(), st));
    GZ_TRY(ctx_ensure(ctx, &h->scan_tmp, tmp));
    GZ_HIP(rocprim::exclusive_scan(h->scan_tmp.p, tmp, (const uint32_t*)h->tile_count.p, (uint64_t*)h->tile_off.p,
                                   (uint64_t)0, (size_t)T, rocprim::plus<uint64_t>(), st));
    // the output size: the last tile's offset plus its count (the per-tile
    // counts come back only when rows per batch are asked for, after the
    // write pass is queued)
    uint64_t last_off = 0;
    uint32_t last_cnt = 0;
    GZ_HIP(hipMemcpyAsync(&last_off, (const uint64_t*)h->tile_off.p + (T - 1), sizeof(uint64_t),
                          hipMemcpyDeviceToHost, st));
    GZ_HIP(hipMemcpyAsync(&last_cnt, (const uint32_t*)h->tile_count.p + (T - 1), sizeof(uint32_t),
                          hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    total = last_off + last_cnt;
    tc_device = true;
  }
  // outputs
  bool any_binary = false;
  const int nres = (int)h->sides.size();
  p.ncols = nres;
  for (int c = 0; c < nres; c++) {
    auto& o = h->outs[c];
    int t = o.type;
    OutCol& oc = p.out[c];
    oc.side = h->sides[c];
    oc.col = h->idx[c];
    oc.type = t;
    oc.width = t == GZ_TYPE_BOOL ? 1 : type_width(t);
    oc.binary = type_is_binary(t);
    any_binary |= oc.binary;
    GZ_TRY(ctx_ensure(ctx, &o.valid, std::max<uint64_t>(total, 1)));
    if (oc.binary) {
      GZ_TRY(ctx_ensure(ctx, &o.offsets, sizeof(int32_t) * (total + 1)));
      GZ_TRY(ctx_ensure(ctx, &o.lens, sizeof(int64_t) * std::max<uint64_t>(total, 1)));
    } else if (t != GZ_TYPE_NA) {
      GZ_TRY(ctx_ensure(ctx, &o.values, (size_t)std::max<uint64_t>(total, 1) * oc.width));
    }
    oc.ovalid = (uint8_t*)o.valid.p;
    oc.ovalues = (uint8_t*)o.values.p;
    if (oc.side == 0) {
      const ColStore& cs = r->cols[oc.col];
      oc.bvalid = (const uint8_t*)cs.valid.p;
      oc.bvalues = (const uint8_t*)cs.values.p;
    }
  }
  if (any_binary) {
    GZ_TRY(ctx_ensure(ctx, &h->pair_probe, sizeof(int64_t) * std::max<uint64_t>(total, 1)));
    GZ_TRY(ctx_ensure(ctx, &h->pair_build, sizeof(int32_t) * std::max<uint64_t>(total, 1)));
    p.pair_probe = (int64_t*)h->pair_probe.p;
    p.pair_build = (int32_t*)h->pair_build.p;
    p.write_pairs = 1;
  }
  // build columns from the payload: fixed-width columns packed by decreasing
  // width into 12 bytes (a validity bit each in the last dword), one record
  // per build row, or per slot (pay_by_slot)
  if (!h->payload_ready && r->n > 0) {
    const bool fits = payload_layout(h, &h->pay_off);
    if (fits) {
      PayloadParams q;
      memset(&q, 0, sizeof(q));
      for (int c = 0; c < nres; c++) {
        if (h->pay_off[c] < 0) continue;
        q.bvalid[q.ncol] = p.out[c].bvalid;
        q.bvalues[q.ncol] = p.out[c].bvalues;
        q.width[q.ncol] = p.out[c].width;
        q.off[q.ncol] = h->pay_off[c];
        q.bit[q.ncol] = c;
        q.ncol++;
      }
      q.n = r->n;
      unsigned gb = (unsigned)std::min<int64_t>((r->n + 255) / 256, 8192);
      if (h->pay_by_slot) {
        GZ_TRY(ctx_ensure(ctx, &h->payload, sizeof(u32x4) * (r->cap + 1)));
        q.out = (u32x4*)h->payload.p;
        hipLaunchKernelGGL(join_payload_slot_kernel, dim3(gb), dim3(256), 0, st, q, (const int32_t*)r->row_slot.p,
                           (const int32_t*)r->start.p, r->special_start, (uint64_t)r->cap);
      } else {
        GZ_TRY(ctx_ensure(ctx, &h->payload, sizeof(u32x4) * r->n));
        q.out = (u32x4*)h->payload.p;
        hipLaunchKernelGGL(join_payload_kernel, dim3(gb), dim3(256), 0, st, q);
      }
    }
    h->payload_ready = true;
  }
  bool use_payload = false;
  for (int c = 0; c < nres && c < (int)h->pay_off.size(); c++) {
    p.pay_off[c] = h->pay_off[c];
    use_payload |= h->pay_off[c] >= 0;
  }
  for (int c = (int)h->pay_off.size(); c < nres; c++) p.pay_off[c] = -1;
  p.payload = use_payload ? (const u32x4*)h->payload.p : nullptr;
  if (single) {
    {
      KernelTimer kt(ctx, "join_probe_write");
      hipLaunchKernelGGL((join_probe_write01_kernel<kSingleTile, kSingleTile, 1>), dim3((unsigned)sblocks),
                         dim3(kSingleTile), 0, st, p);
    }
    GZ_HIP(hipGetLastError());
    GZ_HIP(hipStreamSynchronize(st));
    const uint32_t rows = ((volatile uint32_t*)h->host_out)[0];
    if (((volatile uint32_t*)h->host_out)[1] || rows > (uint64_t)rows0)
      return fail(GZ_HIP_ERROR, "probe: a block's wait for the blocks before it ran out");
    total = rows;
    tc[0] = rows;
  } else if (T > 0 && total > 0) {
    int64_t grid = (T + 7) / 8 * 8;
    KernelTimer kt(ctx, "join_probe_write");
    if (at_most_one) {
      if (TILE == 4096)
        hipLaunchKernelGGL((join_probe_write01_kernel<4096, kWrite01Block, 0>), dim3((unsigned)grid),
                           dim3(kWrite01Block), 0, st, p);
      else
        hipLaunchKernelGGL((join_probe_write01_kernel<1024, 256, 0>), dim3((unsigned)grid), dim3(256), 0, st, p);
    } else if (TILE == 4096) {
      hipLaunchKernelGGL((join_probe_write_kernel<4096, 256>), dim3((unsigned)grid), dim3(256), 0, st, p);
    } else {
      hipLaunchKernelGGL((join_probe_write_kernel<1024, 256>), dim3((unsigned)grid), dim3(256), 0, st, p);
    }
  }
  if (any_binary && total > 0) {
    for (int c = 0; c < nres; c++) {
      if (!p.out[c].binary) continue;
      auto& o = h->outs[c];
      BinGather g;
      memset(&g, 0, sizeof(g));
      g.tiles = p.tiles;
      g.bufs = p.bufs;
      g.bstride = p.bstride;
      g.side = p.out[c].side;
      g.col = p.out[c].col;
      g.pair_probe = p.pair_probe;
      g.pair_build = p.pair_build;
      g.n = (int64_t)total;
      if (g.side == 0) {
        const ColStore& cs = r->cols[g.col];
        g.bvalid = (const uint8_t*)cs.valid.p;
        g.boff = (const int64_t*)cs.offsets.p;
        g.bdata = (const uint8_t*)cs.data.p;
      }
      g.ovalid = (uint8_t*)o.valid.p;
      g.olen = (int64_t*)o.lens.p;
      g.ooff = (int32_t*)o.offsets.p;
      unsigned gb = (unsigned)std::min<uint64_t>((total + 255) / 256, 8192);
      hipLaunchKernelGGL(join_bin_len_kernel, dim3(gb), dim3(256), 0, st, g);
      size_t tmp = 0;
      GZ_HIP(rocprim::exclusive_scan(nullptr, tmp, (const int64_t*)o.lens.p, (int64_t*)o.lens.p, (int64_t)0,
                                     (size_t)total, rocprim::plus<int64_t>(), st));
      GZ_TRY(ctx_ensure(ctx, &h->scan_tmp, tmp));
      // in-place scan is not allowed: scan into the pair buffer's space is not
      // reusable either, so scan via a second buffer
      DBuf scanned;
      GZ_TRY(ctx_alloc(ctx, sizeof(int64_t) * total, &scanned));
      GZ_HIP(rocprim::exclusive_scan(h->scan_tmp.p, tmp, (const int64_t*)o.lens.p, (int64_t*)scanned.p, (int64_t)0,
                                     (size_t)total, rocprim::plus<int64_t>(), st));
      // total bytes = last start + last length
      int64_t last[2];
      GZ_HIP(hipMemcpyAsync(&last[0], (int64_t*)scanned.p + total - 1, 8, hipMemcpyDeviceToHost, st));
      GZ_HIP(hipMemcpyAsync(&last[1], (int64_t*)o.lens.p + total - 1, 8, hipMemcpyDeviceToHost, st));
      GZ_HIP(hipStreamSynchronize(st));
      int64_t bytes = last[0] + last[1];
      if (bytes >= INT32_MAX) {
        ctx_free(ctx, &scanned);
        return fail(GZ_CAPACITY_ERROR, "utf8 output exceeds int32 offsets");
      }
      GZ_TRY(ctx_ensure(ctx, &o.data, (size_t)std::max<int64_t>(bytes, 1)));
      g.odata = (uint8_t*)o.data.p;
      g.olen = (int64_t*)scanned.p;
      hipLaunchKernelGGL(join_bin_copy_kernel, dim3(gb), dim3(256), 0, st, g);
      GZ_HIP(hipStreamSynchronize(st));
      ctx_free(ctx, &scanned);
    }
  } else if (any_binary) {
    for (int c = 0; c < nres; c++)
      if (p.out[c].binary) GZ_HIP(hipMemsetAsync(h->outs[c].offsets.p, 0, 4, st));
  }
  GZ_HIP(hipGetLastError());
  h->out_rows = (int64_t)total;
  if (out_rows) *out_rows = (int64_t)total;
  if (out_batch_rows) {
    if (tc_device) {
      GZ_HIP(hipMemcpyAsync(tc.data(), h->tile_count.p, sizeof(uint32_t) * T, hipMemcpyDeviceToHost, st));
      GZ_HIP(hipStreamSynchronize(st));
    }
    for (int bi = 0; bi < nb; bi++) out_batch_rows[bi] = 0;
    for (int64_t t = 0; t < T; t++) out_batch_rows[tiles[t].batch] += tc[t];
  }
  return GZ_OK;
}

namespace gz {

int relation_view(gz_relation* r, RelationView* v) {
  if (!r || !r->finished) return fail(GZ_INVALID, "relation not finished");
  if (r->ncols > 32) return fail(GZ_NOT_IMPLEMENTED, "relation with more than 32 columns");
  if (r->kd) return fail(GZ_NOT_IMPLEMENTED, "the fused stage joins on one key column");
  memset(v, 0, sizeof(*v));
  v->ctx = r->ctx;
  v->keys = (const int64_t*)r->keys.p;
  v->start = (const int32_t*)r->start.p;
  v->count = (const int32_t*)r->count.p;
  v->rowlist = r->has_dups ? (const int32_t*)r->rowlist.p : nullptr;
  v->cap = r->cap;
  v->special_start = r->special_start;
  v->special_count = r->special_count;
  v->nkeys = r->nkeys;
  v->rows = r->n;
  v->key_col = r->key_col;
  v->ncols = r->ncols;
  for (int c = 0; c < r->ncols; c++) {
    v->types[c] = r->cols[c].type;
    v->valid[c] = (const uint8_t*)r->cols[c].valid.p;
    v->values[c] = (const void*)r->cols[c].values.p;
  }
  return GZ_OK;
}

}  // namespace gz

extern "C" {

int gz_probe_make(gz_relation* r, int32_t num_probe_cols, const int32_t* probe_types, int32_t probe_key_col,
                  int32_t join_type, int32_t null_aware_anti, int32_t num_result_cols,
                  const int32_t* result_sides, const int32_t* result_indices, gz_probe** out) {
  if (!r || !out || (num_probe_cols > 0 && !probe_types)) return fail(GZ_INVALID, "gz_probe_make: null argument");
  if (!r->finished) return fail(GZ_INVALID, "the hash relation must be finished before probing");
  if (join_type != GZ_JOIN_INNER && join_type != GZ_JOIN_SEMI && join_type != GZ_JOIN_ANTI)
    return fail(GZ_NOT_IMPLEMENTED, "ConditionedProbeArraysTypedImpl only support join type: InnerJoin, "
                                    "SemiJoin, AntiJoin");
  if (num_probe_cols > kJoinMaxCols || num_result_cols > kJoinMaxCols)
    return fail(GZ_NOT_IMPLEMENTED, "more than 32 columns");
  if (probe_key_col < 0 || probe_key_col >= num_probe_cols) return fail(GZ_INVALID, "probe key out of range");
  int pk = probe_types[probe_key_col], bk = r->types[r->key_col];
  if (pk != bk) return fail(GZ_TYPE_ERROR, "probe and build key types differ");
  auto* h = new gz_probe();
  h->rel = r;
  h->ctx = r->ctx;
  h->nprobe = num_probe_cols;
  h->ptypes.assign(probe_types, probe_types + num_probe_cols);
  h->key_col = probe_key_col;
  h->join_type = join_type;
  h->null_aware = null_aware_anti;
  for (int c = 0; c < num_result_cols; c++) {
    int s = result_sides[c], i = result_indices[c];
    if ((s == 0 && (i < 0 || i >= r->ncols)) || (s == 1 && (i < 0 || i >= num_probe_cols)) || (s != 0 && s != 1)) {
      delete h;
      return fail(GZ_INVALID, "result column out of range");
    }
    h->sides.push_back(s);
    h->idx.push_back(i);
    gz_probe::Out o;
    o.type = s == 0 ? r->types[i] : probe_types[i];
    o.width = type_width(o.type);
    if (o.type == GZ_TYPE_DECIMAL128 && s == 0) {
      // 16-byte build values are gathered as bytes
    }
    h->outs.push_back(o);
  }
  *out = h;
  return GZ_OK;
}

int gz_probe_process(gz_probe* h, int64_t num_rows, const int64_t* buf_addrs, const int64_t* buf_sizes,
                     int32_t num_bufs, int64_t* out_rows) {
  if (!h) return fail(GZ_INVALID, "null probe");
  if (!buf_addrs || !buf_sizes) return fail(GZ_INVALID, "buf_addrs can't be null");
  gz_batch b{num_rows, buf_addrs, buf_sizes, num_bufs};
  if (h->kd) {
    std::vector<int64_t> a, s;
    std::vector<gz_batch> kb;
    GZ_TRY(probe_codes(h, 1, &b, a, s, kb));
    return probe_run(h, 1, kb.data(), out_rows, nullptr);
  }
  return probe_run(h, 1, &b, out_rows, nullptr);
}

int gz_probe_process_batches(gz_probe* h, int32_t num_batches, const gz_batch* batches, int64_t* out_rows,
                             int64_t* out_batch_rows) {
  if (!h) return fail(GZ_INVALID, "null probe");
  if (num_batches < 0 || (num_batches > 0 && !batches)) return fail(GZ_INVALID, "bad batch list");
  if (h->kd) {
    std::vector<int64_t> a, s;
    std::vector<gz_batch> kb;
    GZ_TRY(probe_codes(h, num_batches, batches, a, s, kb));
    return probe_run(h, num_batches, kb.data(), out_rows, out_batch_rows);
  }
  return probe_run(h, num_batches, batches, out_rows, out_batch_rows);
}

int gz_probe_set_context(gz_probe* h, gz_context* ctx) {
  if (!h || !ctx) return fail(GZ_INVALID, "null argument");
  if (h->used) return fail(GZ_INVALID, "a probe's context is set before its first process call");
  if (ctx->device != h->rel->ctx->device) return fail(GZ_INVALID, "the probe's context is on another device");
  GZ_HIP(hipSetDevice(ctx->device));
  GZ_HIP(hipStreamSynchronize(h->rel->ctx->stream));
  h->ctx = ctx;
  return GZ_OK;
}

int gz_probe_result_column(gz_probe* h, int32_t col, void** values, void** valid_bytes, void** offsets,
                           void** data) {
  if (!h) return fail(GZ_INVALID, "null probe");
  if (col < 0 || col >= (int)h->outs.size()) return fail(GZ_INDEX_ERROR, "result column out of range");
  auto& o = h->outs[col];
  if (values) *values = o.values.p;
  if (valid_bytes) *valid_bytes = o.valid.p;
  if (offsets) *offsets = o.offsets.p;
  if (data) *data = o.data.p;
  return GZ_OK;
}

int gz_probe_result_bitmap(gz_probe* h, int32_t col, void** bitmap) {
  if (!h || !bitmap) return fail(GZ_INVALID, "null argument");
  if (col < 0 || col >= (int)h->outs.size()) return fail(GZ_INDEX_ERROR, "result column out of range");
  auto& o = h->outs[col];
  GZ_TRY(ctx_ensure(h->ctx, &o.bits, (size_t)std::max<int64_t>((h->out_rows + 63) / 64 * 8, 8)));
  GZ_TRY(pack_validity(h->ctx, (const uint8_t*)o.valid.p, h->out_rows, (uint8_t*)o.bits.p));
  *bitmap = o.bits.p;
  return GZ_OK;
}

int gz_probe_release(gz_probe* h) {
  if (!h) return GZ_OK;
  gz_context* ctx = h->ctx;
  kd_scratch_free(ctx, &h->ksc);
  for (auto& o : h->outs)
    for (DBuf* b : {&o.values, &o.valid, &o.offsets, &o.data, &o.lens, &o.bits}) ctx_free(ctx, b);
  for (DBuf* b : {&h->d_tiles, &h->d_bufs, &h->match, &h->tile_count, &h->tile_off, &h->scan_tmp,
                  &h->pair_probe, &h->pair_build, &h->payload, &h->look})
    ctx_free(ctx, b);
  if (h->host_out) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipHostFree(h->host_out);
  }
  delete h;
  return GZ_OK;
}

int gz_relation_make_keys(gz_context* ctx, int32_t num_cols, const int32_t* col_types, int32_t num_keys,
                          const int32_t* key_cols, int32_t builder_type, gz_relation** out) {
  if (!ctx || !out || (num_cols > 0 && !col_types) || (num_keys > 0 && !key_cols))
    return fail(GZ_INVALID, "gz_relation_make_keys: null argument");
  if (num_keys < 1) return fail(GZ_INVALID, "no key column");
  for (int j = 0; j < num_keys; j++)
    if (key_cols[j] < 0 || key_cols[j] >= num_cols) return fail(GZ_INVALID, "key column out of range");
  if (num_keys == 1 && col_types[key_cols[0]] != GZ_TYPE_STRING)
    return gz_relation_make(ctx, num_cols, col_types, key_cols[0], builder_type, out);
  if (num_cols + 1 > kJoinMaxCols) return fail(GZ_NOT_IMPLEMENTED, "more than 31 build columns with several keys");
  // the dictionary's key order (one utf8 key last); probe keys and
  // normalisation flags, given in the caller's order, follow it
  std::vector<int32_t> kt, ext(col_types, col_types + num_cols);
  for (int j = 0; j < num_keys; j++) kt.push_back(col_types[key_cols[j]]);
  const std::vector<int> perm = kd_key_order(num_keys, kt.data());
  kt.clear();
  for (int j = 0; j < num_keys; j++) kt.push_back(col_types[key_cols[perm[j]]]);
  ext.push_back(GZ_TYPE_INT64);
  KeyDict* kd = nullptr;
  GZ_TRY(kd_make(ctx, num_keys, kt.data(), &kd));
  gz_relation* r = nullptr;
  int rc = gz_relation_make(ctx, num_cols + 1, ext.data(), num_cols, builder_type, &r);
  if (rc) {
    kd_release(kd);
    return rc;
  }
  r->kd = kd;
  r->kperm = perm;
  for (int j = 0; j < num_keys; j++) r->kcols.push_back(key_cols[perm[j]]);
  r->otypes.assign(col_types, col_types + num_cols);
  r->ocols = num_cols;
  *out = r;
  return GZ_OK;
}

int gz_relation_normalize_keys(gz_relation* r, int32_t num_keys, const int32_t* normalize) {
  if (!r || (num_keys > 0 && !normalize)) return fail(GZ_INVALID, "gz_relation_normalize_keys: null argument");
  if (r->n > 0 || r->finished) return fail(GZ_INVALID, "key normalisation is set before the first append");
  const int K = r->kd ? (int)r->kcols.size() : 1;
  if (num_keys != K) return fail(GZ_INVALID, "the relation has " + std::to_string(K) + " key columns");
  uint32_t mask = 0;
  for (int i = 0; i < K; i++) {  // dictionary position i holds the caller's key j
    const int j = r->kd ? r->kperm[i] : i;
    if (!normalize[j]) continue;
    const int t = r->kd ? r->otypes[r->kcols[i]] : r->types[r->key_col];
    if (t != GZ_TYPE_FLOAT && t != GZ_TYPE_DOUBLE)
      return fail(GZ_TYPE_ERROR, "only float and double keys are normalised (key " + std::to_string(j) + ")");
    mask |= 1u << i;
  }
  if (r->kd) return kd_set_normalize(r->kd, mask);
  r->normalize = mask != 0;
  return GZ_OK;
}

int gz_probe_make_keys(gz_relation* r, int32_t num_probe_cols, const int32_t* probe_types, int32_t num_keys,
                       const int32_t* probe_key_cols, int32_t join_type, int32_t null_aware_anti,
                       int32_t num_result_cols, const int32_t* result_sides, const int32_t* result_indices,
                       gz_probe** out) {
  if (!r || !out || (num_probe_cols > 0 && !probe_types) || (num_keys > 0 && !probe_key_cols))
    return fail(GZ_INVALID, "gz_probe_make_keys: null argument");
  if (!r->kd) {
    if (num_keys != 1) return fail(GZ_INVALID, "the relation was built with one key column");
    return gz_probe_make(r, num_probe_cols, probe_types, probe_key_cols[0], join_type, null_aware_anti,
                         num_result_cols, result_sides, result_indices, out);
  }
  if (num_keys != (int)r->kcols.size()) return fail(GZ_INVALID, "probe and build key counts differ");
  for (int i = 0; i < num_keys; i++) {  // dictionary position i: the caller's key kperm[i]
    const int pc = probe_key_cols[r->kperm[i]];
    if (pc < 0 || pc >= num_probe_cols) return fail(GZ_INVALID, "probe key out of range");
    if (probe_types[pc] != r->otypes[r->kcols[i]]) return fail(GZ_TYPE_ERROR, "probe and build key types differ");
  }
  // Several keys: the reference's UnsafeRow anti probe (conditioned_probe_kernel.cc:957-982)
  // never reads is_null_aware_anti_join_: a probe row is kept unless its key
  // row exists, and a key row with a null part never does (the build skips
  // such rows).  So the null-aware anti join over several keys is the plain
  // one; one utf8 key keeps the one-key null-aware rules (:1016-1031).
  const int na = num_keys == 1 ? null_aware_anti : 0;
  for (int c = 0; c < num_result_cols; c++) {
    if (!result_sides || !result_indices) break;
    const int lim = result_sides[c] == 0 ? r->ocols : num_probe_cols;  // the appended code column is internal
    if (result_indices[c] < 0 || result_indices[c] >= lim) return fail(GZ_INVALID, "result column out of range");
  }
  if (num_probe_cols + 1 > kJoinMaxCols) return fail(GZ_NOT_IMPLEMENTED, "more than 31 probe columns with several keys");
  std::vector<int32_t> ext(probe_types, probe_types + num_probe_cols);
  ext.push_back(GZ_TYPE_INT64);
  gz_probe* h = nullptr;
  GZ_TRY(gz_probe_make(r, num_probe_cols + 1, ext.data(), num_probe_cols, join_type, na, num_result_cols,
                       result_sides, result_indices, &h));
  h->kd = r->kd;
  for (int i = 0; i < num_keys; i++) h->kcols.push_back(probe_key_cols[r->kperm[i]]);
  h->otypes.assign(probe_types, probe_types + num_probe_cols);
  h->ocols = num_probe_cols;
  *out = h;
  return GZ_OK;
}

}  // extern "C"

namespace gz {
// Probe on another context (stream and memory) than the relation's: a resident
// relation serves probes from many streams; its build must be complete.
void probe_set_context(gz_probe* h, gz_context* ctx) { h->ctx = ctx; }
int relation_synchronize(gz_relation* r) {
  GZ_HIP(hipStreamSynchronize(r->ctx->stream));
  return GZ_OK;
}
}  // namespace gz
