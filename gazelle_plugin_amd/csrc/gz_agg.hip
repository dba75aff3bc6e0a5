// Hash aggregate for MI355X (gfx950).
//
// Replaces HashAggregateKernel + HashAggregateResultIterator<number>
// (ext/hash_aggregate_kernel.cc:51-928) with its accumulators (ext/actions_impl.cc).
//
// Groups live in an open-addressing table in HBM: int64 keys inserted with a
// 64-bit CAS, linear probing, and per group the first global row that carried
// the key (atomicMin), which reproduces the reference's first-seen group order
// (SparseHashMap ids, sparse_hash_map.h:39-47) at finish.  Aggregate states are
// stored as 8-byte "lanes" (structure of arrays), one per distinct (operation,
// input column): sum / non-null count / min / max ..., updated with global
// atomics (FP64 adds with the hardware f64 atomic; integer and ordered-float
// min/max with 64-bit atomics).  Sums stay FP64 / int64.
//
// Low-cardinality instances use an LDS first-level table per workgroup: rows
// aggregate with LDS atomics, and each workgroup merges its table into the HBM
// table once per instance.  An instance whose LDS table fills spills rows to
// the HBM path, so the result never depends on the placement.
//
// An aggregate *set* holds many independent instances (one GROUP BY each); one
// launch sequence processes the tiles of all of them.  The single-aggregate C
// ABI (gz_agg_*) is a set of one instance.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <string>
#include <vector>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

constexpr int kAggMaxCols = 32;
constexpr int kMaxLanes = 24;
constexpr int kMaxActions = 32;
constexpr int64_t kEmpty = INT64_MIN;
constexpr int kHllBits = 11;
constexpr int kHllRegs = 1 << kHllBits;
constexpr int kMaxUsed = 8;        // used input columns of a partition record
constexpr int kMaxFanout = 4096;   // partitions per instance on the partitioned path
constexpr int kSmallFanout = 256;  // fan-outs up to this share a scatter launch with a small LDS footprint

// Floating min/max keep one input value per group, as the reference's
// accumulators do (MinAction / MaxAction, ext/actions_impl.cc:905-936,
// :1504-1539): L_MINF / L_MAXF order the non-NaN values with -0.0 and 0.0 as one
// key; L_FIRSTZ (row * 2 + sign of the first zero) gives a zero result the sign
// of the first zero seen; the NaN a result keeps is named by a row lane
// (L_FIRSTNAN, L_LASTNAN, L_GLASTNAN) and its bits are copied into L_NANBITS by
// a fix-up pass over the rows (agg_fix_kernel), since no 64-bit atomic can
// carry both a row index and a NaN payload.
enum LaneOp : int {
  L_CNT = 0, L_NN, L_NN2, L_SUMI, L_SUMF, L_SUMCNT, L_MINI, L_MAXI, L_MINU, L_MAXU,
  L_MINF, L_MAXF, L_HASNAN, L_FIRSTNN, L_FIRSTNAN,
  L_FIRSTZ,    // min over zero rows of row * 2 + sign bit
  L_LASTNAN,   // max over NaN rows of the row (max with NaN check: every NaN overwrites)
  L_GLASTNAN,  // no group-by key, max with NaN check: the first NaN of the last batch holding one
  L_NANBITS    // bits of the NaN named by lane col2 (-1: the group's first row), written by the fix-up
};

struct Lane {
  int op, col, col2;
};

struct AggInst {
  int64_t* base;          // keys [S], first_row [S], lanes [S][nl] with S = cap + 2 (a slot's lanes contiguous)
  uint64_t cap;           // power of two; slot cap = null key, cap + 1 = key equal to kEmpty
  unsigned int* counter;  // keys inserted
  int* overflow;
  int32_t fan_log2;       // partitioned placement: log2 of the partition count (0 = off)
  int32_t bucket_base;    // first global partition of the instance
  int32_t nl;             // lanes per slot
  int64_t cnt_off;        // partitioned: the instance's [partition][tile] record counts start here
  int32_t cnt_tiles;      // ... its partition tiles (a count row per partition)
  int32_t tile0;          // ... and its first partition tile
  int32_t sub_log2;       // 0: one hash table over all slots; else partition-structured: partition b
                          // (fan_log2 high hash bits) owns slots [b << sub_log2, (b + 1) << sub_log2),
                          // probed from the low hash bits within them (an LDS table's layout)
};

// The HBM table's first-row word and lanes of slot s: keys and first rows are
// arrays of their own (probes and occupancy scans read them coalesced); a
// slot's lanes are one contiguous run, so a group's state is one or two cache
// lines for its atomics and for the gathers of the finish.
__device__ __forceinline__ int64_t* tab_first(const AggInst& I, int64_t s) { return I.base + (int64_t)I.cap + 2 + s; }
__device__ __forceinline__ int64_t* tab_lanes(const AggInst& I, int64_t s) {
  return I.base + 2 * ((int64_t)I.cap + 2) + s * I.nl;
}

struct AggTile {
  int32_t batch;
  int32_t row_begin;
  int32_t nrows;
  int32_t inst;
  int64_t row_base;  // instance-global index of the tile's first row
};

struct AggParams {
  const AggTile* tiles;
  int64_t T;
  const void* const* bufs;
  int bstride;
  int key_col, key_type;
  int nl;
  Lane lanes[kMaxLanes];
  int col_type[kAggMaxCols];
  const AggInst* inst;
  int tiles_per_block;
  int lcap;
  int lprobe;                 // LDS table slots a key tries before going to the HBM table
  // partitioned placement: the used input columns of a record
  int nu;                     // used columns (lane inputs other than the key)
  int ucol[kMaxUsed];         // column of used slot k
  int uslot[kAggMaxCols];     // used slot of a column
  int64_t rstride;            // records: word w of record q at recs[w * rstride + q]
  int64_t tile_base;          // partition tiles: index of tiles[0] among all partition tiles
  int* fix_flag;              // set once a NaN reaches a lane whose bits a result keeps (null: no such lane)
};

__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ __forceinline__ const void* abuf(const AggParams& p, int batch, int col, int k) {
  return p.bufs[(int64_t)batch * p.bstride + col * 3 + k];
}
__device__ __forceinline__ bool is_valid(const AggParams& p, int batch, int col, int64_t r) {
  const uint8_t* vb = (const uint8_t*)abuf(p, batch, col, 0);
  return !vb || bit_get(vb, r);
}
__device__ __forceinline__ int64_t read_i64(const void* v, int t, int64_t r) {
  switch (t) {
    case GZ_TYPE_BOOL: return bit_get((const uint8_t*)v, r);
    case GZ_TYPE_INT8: return ((const int8_t*)v)[r];
    case GZ_TYPE_UINT8: return ((const uint8_t*)v)[r];
    case GZ_TYPE_INT16: return ((const int16_t*)v)[r];
    case GZ_TYPE_UINT16: return ((const uint16_t*)v)[r];
    case GZ_TYPE_INT32: case GZ_TYPE_DATE32: return ((const int32_t*)v)[r];
    case GZ_TYPE_UINT32: return ((const uint32_t*)v)[r];
    default: return ((const int64_t*)v)[r];
  }
}
__device__ __forceinline__ double read_f64(const void* v, int t, int64_t r) {
  if (t == GZ_TYPE_DOUBLE) return ((const double*)v)[r];
  if (t == GZ_TYPE_FLOAT) return (double)((const float*)v)[r];
  if (t == GZ_TYPE_UINT64) return (double)((const uint64_t*)v)[r];
  return (double)read_i64(v, t, r);
}
// Order-preserving map of a non-NaN double to uint64; -0.0 maps as 0.0 (the
// reference's `<` / `>` never prefer one zero to the other).
__device__ __forceinline__ uint64_t ordered(double d) {
  uint64_t b = (uint64_t)__double_as_longlong(d);
  if ((b << 1) == 0) b = 0;
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
constexpr int64_t kCanonNaN = 0x7ff8000000000000ll;
__device__ __forceinline__ bool word_is_nan(int64_t w) { return (w & 0x7fffffffffffffffll) > 0x7ff0000000000000ll; }
// A float group key as the table holds it (SparseHashMap<double>,
// sparse_hash_map.h:86-160; phmap Hash<double>, phmap_utils.h:251-258): the
// double's bits with -0.0 as 0.0 and every NaN as one NaN.
__device__ __forceinline__ int64_t norm_float_key(int64_t w) {
  if ((w << 1) == 0) return 0;
  return word_is_nan(w) ? kCanonNaN : w;
}
__device__ __forceinline__ void note_nan(const AggParams& p) {
  if (p.fix_flag && *(volatile int*)p.fix_flag == 0) *(volatile int*)p.fix_flag = 1;
}
// Value bits as stored (a float's 32 bits zero-extended): the bits a min / max
// result or a float group key comes out with.
__device__ __forceinline__ int64_t raw_bits(const void* v, int t, int64_t r) {
  if (t == GZ_TYPE_FLOAT) return (int64_t)((const uint32_t*)v)[r];
  return ((const int64_t*)v)[r];
}
__device__ __forceinline__ bool raw_is_nan(int64_t raw, int t) {
  if (t == GZ_TYPE_FLOAT) return (raw & 0x7fffffff) > 0x7f800000;
  return word_is_nan(raw);
}
// Ordinal of a NaN row for L_GLASTNAN: the latest batch first, then the
// earliest row in it (batch start < 2^39 rows, rows per batch < 2^24).
__device__ __forceinline__ int64_t glast_ord(int64_t grow, int64_t r) {
  return ((grow - r) << 24) | (int64_t)(0xffffff - r);
}

__device__ __forceinline__ int64_t lane_init(int op) {
  switch (op) {
    case L_MINI: return INT64_MAX;
    case L_MAXI: return INT64_MIN;
    case L_MINU: case L_MINF: return -1;  // all ones
    case L_FIRSTNN: case L_FIRSTNAN: case L_FIRSTZ: return INT64_MAX;
    case L_LASTNAN: case L_GLASTNAN: return -1;
    default: return 0;
  }
}

// Group key of row r (widened to int64); false for a null key.
__device__ __forceinline__ bool row_key(const AggParams& p, int batch, int64_t r, int64_t* key) {
  if (p.key_col < 0) {
    *key = 0;
    return true;
  }
  if (!is_valid(p, batch, p.key_col, r)) return false;
  const void* v = abuf(p, batch, p.key_col, 2);
  if (p.key_type == GZ_TYPE_DOUBLE || p.key_type == GZ_TYPE_FLOAT) {
    const int64_t w = p.key_type == GZ_TYPE_DOUBLE ? ((const int64_t*)v)[r]
                                                   : __double_as_longlong((double)((const float*)v)[r]);
    if (word_is_nan(w)) note_nan(p);
    *key = norm_float_key(w);
    return true;
  }
  *key = read_i64(v, p.key_type, r);
  return true;
}

// ---- batched row loads: RPT rows lr0 + k * S (k < RPT) of a tile starting at
// input row r0, rows with lr0 + k * S >= n inactive.  Every load is issued
// before any value is used (one HBM round trip for all RPT rows); the type
// switch is uniform.
//
// Record word of a column: integers widened as read_i64, floats as double bits,
// 8-byte types raw (the same words the scatter writes and RecRow decodes).
template <int RPT, int S>
__device__ __forceinline__ void load_words(const void* vb, int ty, int64_t r0, int lr0, int n, int64_t (&v)[RPT]) {
#define GZ_LOADW(T, CONV)                                                        \
  {                                                                              \
    const T* s = (const T*)vb + r0 + lr0;                                        \
    T x[RPT];                                                                    \
    _Pragma("unroll") for (int k = 0; k < RPT; k++) x[k] = lr0 + k * S < n ? gld(s + k * S) : (T)0; \
    _Pragma("unroll") for (int k = 0; k < RPT; k++) v[k] = CONV(x[k]);          \
    break;                                                                       \
  }
#define GZ_SEXT(x) ((int64_t)(x))
#define GZ_F2W(x) (__double_as_longlong((double)(x)))
  switch (ty) {
    case GZ_TYPE_BOOL: {
      const uint8_t* s = (const uint8_t*)vb;
      uint8_t x[RPT];
#pragma unroll
      for (int k = 0; k < RPT; k++) x[k] = lr0 + k * S < n ? gld(s + ((r0 + lr0 + k * S) >> 3)) : (uint8_t)0;
#pragma unroll
      for (int k = 0; k < RPT; k++) v[k] = (x[k] >> ((r0 + lr0 + k * S) & 7)) & 1;
      break;
    }
    case GZ_TYPE_INT8: GZ_LOADW(int8_t, GZ_SEXT)
    case GZ_TYPE_UINT8: GZ_LOADW(uint8_t, GZ_SEXT)
    case GZ_TYPE_INT16: GZ_LOADW(int16_t, GZ_SEXT)
    case GZ_TYPE_UINT16: GZ_LOADW(uint16_t, GZ_SEXT)
    case GZ_TYPE_INT32: case GZ_TYPE_DATE32: GZ_LOADW(int32_t, GZ_SEXT)
    case GZ_TYPE_UINT32: GZ_LOADW(uint32_t, GZ_SEXT)
    case GZ_TYPE_FLOAT: GZ_LOADW(float, GZ_F2W)
    default: GZ_LOADW(int64_t, GZ_SEXT)
  }
#undef GZ_LOADW
#undef GZ_SEXT
#undef GZ_F2W
}

// Validity of RPT rows in two steps, so the bitmap loads can be issued before
// other loads and used after them: valid_bytes issues the loads, valid_mask
// gives bit k set when row k is active and valid (no bitmap: every active row).
template <int RPT, int S>
__device__ __forceinline__ void valid_bytes(const uint8_t* vb, int64_t r0, int lr0, int n, uint8_t (&x)[RPT]) {
#pragma unroll
  for (int k = 0; k < RPT; k++) x[k] = vb && lr0 + k * S < n ? gld(vb + ((r0 + lr0 + k * S) >> 3)) : (uint8_t)0xff;
}
template <int RPT, int S>
__device__ __forceinline__ uint32_t valid_mask(const uint8_t (&x)[RPT], int64_t r0, int lr0, int n) {
  uint32_t m = 0;
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (lr0 + k * S < n && ((x[k] >> ((r0 + lr0 + k * S) & 7)) & 1)) m |= 1u << k;
  return m;
}

// Group keys of RPT rows and their validity mask (a null key: bit clear).
template <int RPT, int S>
__device__ __forceinline__ uint32_t load_keys(const AggParams& p, int batch, int64_t r0, int lr0, int n,
                                              int64_t (&key)[RPT]) {
  uint8_t vx[RPT];
  if (p.key_col < 0) {
#pragma unroll
    for (int k = 0; k < RPT; k++) key[k] = 0;
    valid_bytes<RPT, S>(nullptr, r0, lr0, n, vx);
    return valid_mask<RPT, S>(vx, r0, lr0, n);
  }
  const void* const* bt = p.bufs + (int64_t)batch * p.bstride + p.key_col * 3;
  valid_bytes<RPT, S>((const uint8_t*)bt[0], r0, lr0, n, vx);
  load_words<RPT, S>(bt[2], p.key_type, r0, lr0, n, key);
  const uint32_t vm = valid_mask<RPT, S>(vx, r0, lr0, n);
  if (p.key_type == GZ_TYPE_DOUBLE || p.key_type == GZ_TYPE_FLOAT) {  // load_words gives a float's double bits
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      if (((vm >> k) & 1) && word_is_nan(key[k])) note_nan(p);
      key[k] = norm_float_key(key[k]);
    }
  }
  return vm;
}

// Where a key's probe runs in an HBM table: slots base + ((s + i) & mask).
__device__ __forceinline__ void probe_range(const AggInst& I, int64_t key, uint64_t* base, uint64_t* mask,
                                            uint64_t* s) {
  const uint64_t h = fmix64((uint64_t)key);
  if (I.sub_log2) {
    *base = (I.fan_log2 ? h >> (64 - I.fan_log2) : 0) << I.sub_log2;
    *mask = (1ull << I.sub_log2) - 1;
  } else {
    *base = 0;
    *mask = I.cap - 1;
  }
  *s = h & *mask;
}

// Slot of a key in an HBM table (insert if new); -1 when the table (or the
// key's partition of a structured table) is full.
__device__ __forceinline__ int64_t table_slot(const AggInst& I, bool valid, int64_t key) {
  if (!valid) return (int64_t)I.cap;
  if (key == kEmpty) return (int64_t)I.cap + 1;
  int64_t* keys = I.base;
  uint64_t base, mask, s;
  probe_range(I, key, &base, &mask, &s);
  for (uint64_t i = 0; i <= mask; i++) {
    const uint64_t q = base + s;
    int64_t k = keys[q];
    if (k == key) return (int64_t)q;
    if (k == kEmpty) {
      unsigned long long prev = atomicCAS((unsigned long long*)&keys[q], (unsigned long long)kEmpty,
                                          (unsigned long long)key);
      if ((int64_t)prev == kEmpty) {
        atomicAdd(I.counter, 1u);
        return (int64_t)q;
      }
      if ((int64_t)prev == key) return (int64_t)q;
    }
    s = (s + 1) & mask;
  }
  atomicOr(I.overflow, 1);
  return -1;
}

// Slot of a key already in an HBM table; -1 when absent.
__device__ __forceinline__ int64_t table_find(const AggInst& I, bool valid, int64_t key) {
  if (!valid) return (int64_t)I.cap;
  if (key == kEmpty) return (int64_t)I.cap + 1;
  uint64_t base, mask, s;
  probe_range(I, key, &base, &mask, &s);
  for (uint64_t i = 0; i <= mask; i++) {
    const int64_t k = I.base[base + s];
    if (k == key) return (int64_t)(base + s);
    if (k == kEmpty) return -1;
    s = (s + 1) & mask;
  }
  return -1;
}

__device__ __forceinline__ void min_i64(int64_t* p, int64_t v) {
  if (v < *p) atomicMin((long long*)p, (long long)v);
}
__device__ __forceinline__ void max_i64(int64_t* p, int64_t v) {
  if (v > *p) atomicMax((long long*)p, (long long)v);
}
__device__ __forceinline__ void min_u64(int64_t* p, uint64_t v) {
  if (v < (uint64_t)*p) atomicMin((unsigned long long*)p, (unsigned long long)v);
}
__device__ __forceinline__ void max_u64(int64_t* p, uint64_t v) {
  if (v > (uint64_t)*p) atomicMax((unsigned long long*)p, (unsigned long long)v);
}

// Row accessors for apply_row: a row of the input Arrow buffers, or a partition
// record (the row's used inputs widened to 8 bytes: integers as read_i64,
// floats as double bits, uint64 raw, and a null mask).
struct ArrowRow {
  const AggParams& p;
  int batch;
  int64_t r;
  __device__ __forceinline__ bool valid(int col) const { return is_valid(p, batch, col, r); }
  __device__ __forceinline__ int64_t i64(int col) const {
    return read_i64(abuf(p, batch, col, 2), p.col_type[col], r);
  }
  __device__ __forceinline__ double f64(int col) const {
    return read_f64(abuf(p, batch, col, 2), p.col_type[col], r);
  }
  __device__ __forceinline__ uint64_t u64(int col) const { return ((const uint64_t*)abuf(p, batch, col, 2))[r]; }
  __device__ __forceinline__ int64_t raw64(int col) const { return ((const int64_t*)abuf(p, batch, col, 2))[r]; }
  __device__ __forceinline__ int64_t rib() const { return r; }  // row in its batch
};

// Records are stored word-major (word w of record q at rec[w * rstride]); the
// first two used words are loaded ahead with the key, the others on use.
struct RecRow {
  const AggParams& p;
  int64_t v0, v1;      // used words 0 and 1
  const int64_t* rec;  // &recs[q]
  uint32_t nulls;      // bit k: used column k is null
  __device__ __forceinline__ int64_t word(int col) const {
    const int u = p.uslot[col];
    return u == 0 ? v0 : u == 1 ? v1 : gld(rec + (int64_t)(2 + u) * p.rstride);
  }
  __device__ __forceinline__ bool valid(int col) const { return !((nulls >> p.uslot[col]) & 1); }
  __device__ __forceinline__ int64_t i64(int col) const { return word(col); }
  __device__ __forceinline__ double f64(int col) const {
    const int64_t x = word(col);
    const int t = p.col_type[col];
    if (t == GZ_TYPE_DOUBLE || t == GZ_TYPE_FLOAT) return __longlong_as_double(x);
    if (t == GZ_TYPE_UINT64) return (double)(uint64_t)x;
    return (double)x;
  }
  __device__ __forceinline__ uint64_t u64(int col) const { return (uint64_t)word(col); }
  __device__ __forceinline__ int64_t raw64(int col) const { return word(col); }
  __device__ __forceinline__ int64_t rib() const { return 0; }  // records carry no batch row (no L_GLASTNAN here)
};

// Apply one row to one lane's state at ptr.
template <class Row>
__device__ __forceinline__ void apply_lane(const AggParams& p, const Lane ln, int64_t* ptr, const Row& row,
                                           int64_t grow) {
  switch (ln.op) {
    case L_CNT: atomicAdd((unsigned long long*)ptr, 1ull); break;
    case L_NN:
      if (row.valid(ln.col)) atomicAdd((unsigned long long*)ptr, 1ull);
      break;
    case L_NN2:
      if (row.valid(ln.col) && row.valid(ln.col2)) atomicAdd((unsigned long long*)ptr, 1ull);
      break;
    case L_SUMI:
      if (row.valid(ln.col)) atomicAdd((unsigned long long*)ptr, (unsigned long long)row.i64(ln.col));
      break;
    case L_SUMF:
      if (row.valid(ln.col)) unsafeAtomicAdd((double*)ptr, row.f64(ln.col));
      break;
    case L_SUMCNT:
      if (row.valid(ln.col)) atomicAdd((unsigned long long*)ptr, (unsigned long long)row.raw64(ln.col2));
      break;
    case L_MINI:
    case L_MAXI: {
      if (!row.valid(ln.col)) break;
      int64_t v = row.i64(ln.col);
      if (ln.op == L_MINI) min_i64(ptr, v); else max_i64(ptr, v);
      break;
    }
    case L_MINU:
    case L_MAXU: {
      if (!row.valid(ln.col)) break;
      uint64_t v = row.u64(ln.col);
      if (ln.op == L_MINU) min_u64(ptr, v); else max_u64(ptr, v);
      break;
    }
    case L_MINF:
    case L_MAXF: {
      if (!row.valid(ln.col)) break;
      double d = row.f64(ln.col);
      if (d != d) break;
      if (ln.op == L_MINF) min_u64(ptr, ordered(d)); else max_u64(ptr, ordered(d));
      break;
    }
    case L_HASNAN: {
      if (!row.valid(ln.col)) break;
      double d = row.f64(ln.col);
      if (d != d && *ptr == 0) atomicOr((unsigned long long*)ptr, 1ull);
      break;
    }
    case L_FIRSTNN:
      if (row.valid(ln.col) && grow < *ptr) atomicMin((long long*)ptr, (long long)grow);
      break;
    case L_FIRSTNAN: {
      if (!row.valid(ln.col)) break;
      double d = row.f64(ln.col);
      if (d != d) {
        note_nan(p);
        if (grow < *ptr) atomicMin((long long*)ptr, (long long)grow);
      }
      break;
    }
    case L_FIRSTZ: {
      if (!row.valid(ln.col)) break;
      const int64_t w = __double_as_longlong(row.f64(ln.col));
      if ((w << 1) == 0) {
        const int64_t z = grow * 2 + (int64_t)((uint64_t)w >> 63);
        if (z < *ptr) atomicMin((long long*)ptr, (long long)z);
      }
      break;
    }
    case L_LASTNAN:
    case L_GLASTNAN: {
      if (!row.valid(ln.col)) break;
      double d = row.f64(ln.col);
      if (d != d) {
        note_nan(p);
        const int64_t o = ln.op == L_LASTNAN ? grow : glast_ord(grow, row.rib());
        if (o > *ptr) atomicMax((long long*)ptr, (long long)o);
      }
      break;
    }
    case L_NANBITS: break;  // written by the fix-up pass
  }
}

// Apply one row to a slot's state: its first-row word and its lanes, lane l
// at lanes[l * lstride].  Works on HBM and on LDS (the atomics are
// generic-address).
template <class Row>
__device__ __forceinline__ void apply_row(const AggParams& p, int64_t* fr, int64_t* lanes, int64_t lstride,
                                          const Row& row, int64_t grow) {
  if (grow < *fr) atomicMin((long long*)fr, (long long)grow);
  for (int l = 0; l < p.nl; l++) apply_lane(p, p.lanes[l], lanes + l * lstride, row, grow);
}
// ... to slot `slot` of an HBM table
template <class Row>
__device__ __forceinline__ void apply_row_hbm(const AggParams& p, const AggInst& I, int64_t slot, const Row& row,
                                              int64_t grow) {
  apply_row(p, tab_first(I, slot), tab_lanes(I, slot), 1, row, grow);
}
// ... to slot `slot` of an LDS table laid out as word w of slot s at tab[w * LS + s]
template <class Row>
__device__ __forceinline__ void apply_row_lds(const AggParams& p, int64_t* tab, int64_t LS, int64_t slot,
                                              const Row& row, int64_t grow) {
  apply_row(p, tab + LS + slot, tab + 2 * LS + slot, LS, row, grow);
}

// Merge one LDS slot's lanes into an HBM slot.  The atomics return nothing
// and nothing waits on them: no HBM round trip per lane.
__device__ __forceinline__ void merge_slot(const AggParams& p, const AggInst& I, int64_t gslot, const int64_t* l,
                                           int64_t LS, int64_t lslot) {
  const int64_t fr = l[LS + lslot];
  if (fr != INT64_MAX) atomicMin((long long*)tab_first(I, gslot), (long long)fr);
  int64_t* glanes = tab_lanes(I, gslot);
  for (int k = 0; k < p.nl; k++) {
    int64_t v = l[(2 + k) * LS + lslot];
    int64_t* ptr = glanes + k;
    switch (p.lanes[k].op) {
      case L_CNT: case L_NN: case L_NN2: case L_SUMI: case L_SUMCNT:
        if (v) atomicAdd((unsigned long long*)ptr, (unsigned long long)v);
        break;
      case L_SUMF: {
        double d = __longlong_as_double(v);
        if (d != 0.0) unsafeAtomicAdd((double*)ptr, d);
        break;
      }
      case L_MINI: if (v != INT64_MAX) atomicMin((long long*)ptr, (long long)v); break;
      case L_MAXI: if (v != INT64_MIN) atomicMax((long long*)ptr, (long long)v); break;
      case L_MINU: case L_MINF: if (v != -1) atomicMin((unsigned long long*)ptr, (unsigned long long)v); break;
      case L_MAXU: case L_MAXF: if (v != 0) atomicMax((unsigned long long*)ptr, (unsigned long long)v); break;
      case L_HASNAN: if (v) atomicOr((unsigned long long*)ptr, 1ull); break;
      case L_FIRSTNN: case L_FIRSTNAN: case L_FIRSTZ:
        if (v != INT64_MAX) atomicMin((long long*)ptr, (long long)v);
        break;
      case L_LASTNAN: case L_GLASTNAN:
        if (v != -1) atomicMax((long long*)ptr, (long long)v);
        break;
      case L_NANBITS: break;
    }
  }
}

// Application of RPT records to the lanes of an LDS table with no per-record
// branch: one switch per lane per call, and every record with a slot updates
// every lane (a null input contributes the lane's neutral value).  A record
// without a slot of its own (past the item's end, or its LDS probe failed)
// carries slot `none` and issues no atomic: its lane is masked off, since the
// atomics of a wave's idle lanes on one shared slot would serialise.  Results
// equal apply_row's: adding 0 or 0.0 and min/max/or with the neutral value
// change nothing (a sum lane is never -0.0: it starts at +0.0).  Records of at
// most two used columns (every operand in registers, no memory access but the
// LDS atomics).  w1: record word 1 (first row | nulls << 48); u0/u1: used
// words 0 and 1.
template <int RPT>
__device__ __forceinline__ void lds_apply_records(const AggParams& p, int64_t* tab, int64_t LS, const int (&slot)[RPT],
                                                  int none, const int64_t (&w1)[RPT], const int64_t (&u0)[RPT],
                                                  const int64_t (&u1)[RPT]) {
  int64_t grow[RPT];
  uint32_t nulls[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    grow[k] = (int64_t)((uint64_t)w1[k] & ((1ull << 48) - 1));
    nulls[k] = (uint32_t)((uint64_t)w1[k] >> 48);
    if (slot[k] != none) atomicMin((long long*)(tab + LS + slot[k]), (long long)grow[k]);
  }
  for (int l = 0; l < p.nl; l++) {
    const Lane ln = p.lanes[l];
    int64_t* lp = tab + (2 + l) * LS;
    const int u = ln.col >= 0 ? p.uslot[ln.col] : 0;
    const int u2 = ln.col2 >= 0 ? p.uslot[ln.col2] : 0;
    const int ty = ln.col >= 0 ? p.col_type[ln.col] : 0;
    auto word = [&](int uu, int k) -> int64_t { return uu == 0 ? u0[k] : u1[k]; };
    auto ok = [&](int uu, int k) -> bool { return !((nulls[k] >> uu) & 1); };
    // the input as double, converted for all RPT records before any use (a
    // uniform branch on the type, no per-record branch)
    double dv[RPT];
    if (ty == GZ_TYPE_DOUBLE || ty == GZ_TYPE_FLOAT) {
#pragma unroll
      for (int k = 0; k < RPT; k++) dv[k] = __longlong_as_double(word(u, k));
    } else if (ty == GZ_TYPE_UINT64) {
#pragma unroll
      for (int k = 0; k < RPT; k++) dv[k] = (double)(uint64_t)word(u, k);
    } else {
#pragma unroll
      for (int k = 0; k < RPT; k++) dv[k] = (double)word(u, k);
    }
    auto f64 = [&](int k) -> double { return dv[k]; };
#define GZ_EACH(STMT) _Pragma("unroll") for (int k = 0; k < RPT; k++) { if (slot[k] == none) continue; unsigned long long* up = (unsigned long long*)(lp + slot[k]); long long* sp = (long long*)(lp + slot[k]); (void)up; (void)sp; STMT; }
    switch (ln.op) {
      case L_CNT: GZ_EACH(atomicAdd(up, 1ull)) break;
      case L_NN: GZ_EACH(atomicAdd(up, ok(u, k) ? 1ull : 0ull)) break;
      case L_NN2: GZ_EACH(atomicAdd(up, ok(u, k) && ok(u2, k) ? 1ull : 0ull)) break;
      case L_SUMI: GZ_EACH(atomicAdd(up, ok(u, k) ? (unsigned long long)word(u, k) : 0ull)) break;
      case L_SUMF: GZ_EACH(unsafeAtomicAdd((double*)up, ok(u, k) ? f64(k) : 0.0)) break;
      case L_SUMCNT: GZ_EACH(atomicAdd(up, ok(u, k) ? (unsigned long long)word(u2, k) : 0ull)) break;
      case L_MINI: GZ_EACH(atomicMin(sp, ok(u, k) ? (long long)word(u, k) : INT64_MAX)) break;
      case L_MAXI: GZ_EACH(atomicMax(sp, ok(u, k) ? (long long)word(u, k) : INT64_MIN)) break;
      case L_MINU: GZ_EACH(atomicMin(up, ok(u, k) ? (unsigned long long)word(u, k) : ~0ull)) break;
      case L_MAXU: GZ_EACH(atomicMax(up, ok(u, k) ? (unsigned long long)word(u, k) : 0ull)) break;
      case L_MINF: GZ_EACH(const double d = f64(k); atomicMin(up, ok(u, k) && d == d ? ordered(d) : ~0ull)) break;
      case L_MAXF: GZ_EACH(const double d = f64(k); atomicMax(up, ok(u, k) && d == d ? ordered(d) : 0ull)) break;
      case L_HASNAN: GZ_EACH(const double d = f64(k); atomicOr(up, ok(u, k) && d != d ? 1ull : 0ull)) break;
      case L_FIRSTNN: GZ_EACH(atomicMin(sp, ok(u, k) ? (long long)grow[k] : INT64_MAX)) break;
      case L_FIRSTNAN: GZ_EACH(const double d = f64(k); atomicMin(sp, ok(u, k) && d != d ? (long long)grow[k] : INT64_MAX)) break;
      case L_FIRSTZ:
        GZ_EACH(const int64_t w = __double_as_longlong(f64(k));
                atomicMin(sp, ok(u, k) && (w << 1) == 0 ? (long long)(grow[k] * 2 + (int64_t)((uint64_t)w >> 63))
                                                        : INT64_MAX))
        break;
      case L_LASTNAN: GZ_EACH(const double d = f64(k); atomicMax(sp, ok(u, k) && d != d ? (long long)grow[k] : -1ll)) break;
      default: break;  // L_GLASTNAN (no group-by key: never partitioned), L_NANBITS (fix-up pass)
    }
    if (ln.op == L_FIRSTNAN || ln.op == L_LASTNAN) {
      bool any = false;
#pragma unroll
      for (int k = 0; k < RPT; k++) any = any || (ok(u, k) && dv[k] != dv[k]);
      if (any) note_nan(p);
    }
#undef GZ_EACH
  }
}

// ---------------------------------------------------------------- kernels
__global__ void fill_u32_kernel_agg(uint32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// Fill of every instance's table: keys empty, first rows max, lanes initial.
__global__ void agg_init_kernel(const AggInst* inst, int ninst, int nl, const int* lane_ops) {
  int i = blockIdx.y;
  if (i >= ninst) return;
  const AggInst I = inst[i];
  int64_t S = (int64_t)I.cap + 2;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < S;
       s += (int64_t)gridDim.x * blockDim.x) {
    I.base[s] = kEmpty;
    *tab_first(I, s) = INT64_MAX;
    int64_t* ln = tab_lanes(I, s);
    for (int l = 0; l < nl; l++) ln[l] = lane_init(lane_ops[l]);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *I.counter = 0;
    *I.overflow = 0;
  }
}

// Reset of every instance's table in one launch: index g of the concatenated
// reset space belongs to the instance whose range [prefix[i], prefix[i+1])
// holds it.  A partition-structured table's range holds only its two special
// slots: its partitions are reset per partition once the work items are known
// (agg_part_reset_kernel), or stored whole by the work item that owns them.
__global__ void agg_init_all_kernel(const AggInst* inst, int ninst, const int64_t* prefix, int64_t total,
                                    int nl, const int* lane_ops) {
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const AggInst I = inst[lo];
    const int64_t s = (I.sub_log2 ? (int64_t)I.cap : 0) + (g - prefix[lo]);
    I.base[s] = kEmpty;
    *tab_first(I, s) = INT64_MAX;
    int64_t* ln = tab_lanes(I, s);
    for (int l = 0; l < nl; l++) ln[l] = lane_init(lane_ops[l]);
    if (g == prefix[lo]) {
      *I.counter = 0;
      *I.overflow = 0;
    }
  }
}

// HBM path: one workgroup per tile.
__global__ __launch_bounds__(256) void agg_update_kernel(AggParams p) {
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const AggTile td = p.tiles[t];
  const AggInst I = p.inst[td.inst];
  for (int lr = threadIdx.x; lr < td.nrows; lr += blockDim.x) {
    int64_t r = (int64_t)td.row_begin + lr;
    int64_t key;
    bool valid = row_key(p, td.batch, r, &key);
    int64_t slot = table_slot(I, valid, key);
    if (slot < 0) continue;
    apply_row_hbm(p, I, slot, ArrowRow{p, td.batch, r}, td.row_base + lr);
  }
}


// LDS path: persistent workgroups over static chunks of consecutive tiles; an
// LDS table of lcap slots (+2 special) per workgroup, merged into HBM whenever
// the instance changes and at the end.
__device__ void lds_merge(const AggParams& p, const AggInst& I, int64_t* l, int64_t LS) {
  for (int64_t s = threadIdx.x; s < (int64_t)p.lcap + 2; s += blockDim.x) {
    bool special = s >= p.lcap;
    int64_t key = l[s];
    if (!special && key == kEmpty) continue;
    if (special && l[LS + s] == INT64_MAX) continue;
    int64_t g;
    if (special) g = (int64_t)I.cap + (s - p.lcap);
    else g = table_slot(I, true, key);
    if (g < 0) continue;
    merge_slot(p, I, g, l, LS, s);
  }
}

// An LDS table whose groups no other workgroup updates (all rows of their
// keys went through this table) written into partition b's slots of a
// structured HBM table, whose layout it has: slot s of the LDS table is slot
// (b << sub_log2) + s, so every group is stored in place with no atomics; the
// special slots go to cap and cap + 1.  The partition's slots are not reset
// beforehand (no other workgroup touches them): every slot's key and first row
// are stored (an empty slot's first row, INT64_MAX, marks it unoccupied for the
// collection), and the lanes of every group; an empty slot's lanes are never
// read.
__device__ void lds_write_image(const AggParams& p, const AggInst& I, int b, const int64_t* l, int64_t LS) {
  __shared__ unsigned s_keys;
  if (threadIdx.x == 0) s_keys = 0;
  __syncthreads();
  const int64_t base = (int64_t)b << I.sub_log2;
  unsigned mine = 0;
  for (int64_t s = threadIdx.x; s < (int64_t)p.lcap + 2; s += blockDim.x) {
    const bool special = s >= p.lcap;
    if (special && l[LS + s] == INT64_MAX) continue;
    const int64_t key = l[s];
    const int64_t g = special ? (int64_t)I.cap + (s - p.lcap) : base + s;
    *tab_first(I, g) = l[LS + s];
    if (!special) {
      I.base[g] = key;
      if (key == kEmpty) continue;
      mine++;
    }
    int64_t* ln = tab_lanes(I, g);
    for (int k = 0; k < p.nl; k++) ln[k] = l[(2 + k) * LS + s];
  }
  if (mine) atomicAdd(&s_keys, mine);
  __syncthreads();
  if (threadIdx.x == 0 && s_keys) atomicAdd(I.counter, s_keys);
}

// The same table written by claiming one slot per group (a key whose LDS probe
// failed already sits in the HBM table, so the image would be wrong), then
// plain stores of its first row and lanes over the slot's initial values.
__device__ void lds_writeout(const AggParams& p, const AggInst& I, const int64_t* l, int64_t LS) {
  for (int64_t s = threadIdx.x; s < (int64_t)p.lcap + 2; s += blockDim.x) {
    const bool special = s >= p.lcap;
    const int64_t key = l[s];
    if (!special && key == kEmpty) continue;
    if (special && l[LS + s] == INT64_MAX) continue;
    const int64_t g = special ? (int64_t)I.cap + (s - p.lcap) : table_slot(I, true, key);
    if (g < 0) continue;
    *tab_first(I, g) = l[LS + s];
    int64_t* ln = tab_lanes(I, g);
    for (int k = 0; k < p.nl; k++) ln[k] = l[(2 + k) * LS + s];
  }
}

__device__ void lds_reset(const AggParams& p, int64_t* l, int64_t LS) {
  for (int64_t s = threadIdx.x; s < LS; s += blockDim.x) {
    l[s] = kEmpty;
    l[LS + s] = INT64_MAX;
    for (int k = 0; k < p.nl; k++) l[(2 + k) * LS + s] = lane_init(p.lanes[k].op);
  }
}

__global__ __launch_bounds__(256) void agg_update_lds_kernel(AggParams p) {
  extern __shared__ int64_t s_tab[];
  const int64_t LS = p.lcap + 3;  // stride: lcap slots, 2 special, 1 spare
  const uint64_t lmask = (uint64_t)p.lcap - 1;
  int64_t t0 = (int64_t)blockIdx.x * p.tiles_per_block;
  int64_t t1 = min(p.T, t0 + p.tiles_per_block);
  if (t0 >= t1) return;
  int cur = -1;
  for (int64_t t = t0; t < t1; t++) {
    const AggTile td = p.tiles[t];
    if (td.inst != cur) {
      if (cur >= 0) {
        __syncthreads();
        lds_merge(p, p.inst[cur], s_tab, LS);
      }
      __syncthreads();
      lds_reset(p, s_tab, LS);
      __syncthreads();
      cur = td.inst;
    }
    const AggInst I = p.inst[td.inst];
    for (int lr = threadIdx.x; lr < td.nrows; lr += blockDim.x) {
      int64_t r = (int64_t)td.row_begin + lr;
      int64_t grow = td.row_base + lr;
      int64_t key;
      bool valid = row_key(p, td.batch, r, &key);
      int64_t ls = -1;
      if (!valid) ls = p.lcap;
      else if (key == kEmpty) ls = p.lcap + 1;
      else {
        uint64_t s = fmix64((uint64_t)key) & lmask;
        for (int i = 0; i < p.lprobe; i++) {  // short probe: a full LDS table spills to HBM
          int64_t k = s_tab[s];
          if (k == key) { ls = (int64_t)s; break; }
          if (k == kEmpty) {
            unsigned long long prev = atomicCAS((unsigned long long*)&s_tab[s],
                                                (unsigned long long)kEmpty, (unsigned long long)key);
            if ((int64_t)prev == kEmpty || (int64_t)prev == key) { ls = (int64_t)s; break; }
          }
          s = (s + 1) & lmask;
        }
      }
      if (ls >= 0) {
        apply_row_lds(p, s_tab, LS, ls, ArrowRow{p, td.batch, r}, grow);
      } else {
        int64_t g = table_slot(I, valid, key);
        if (g >= 0) apply_row_hbm(p, I, g, ArrowRow{p, td.batch, r}, grow);
      }
    }
  }
  __syncthreads();
  lds_merge(p, p.inst[cur], s_tab, LS);
}

// ---------------------------------------------------------------- fix-up
// The bits of the NaN a result keeps: after a call's update kernels, every
// valid NaN row of a column with an L_NANBITS lane looks up its group and, if
// its ordinal is the one the group's row lane kept, stores its bits.  Each
// group's row lane names one row, so one store per lane and group lands.  The
// pass returns at once unless a NaN reached such a lane (the flag is sticky).
__global__ __launch_bounds__(256) void agg_fix_kernel(AggParams p) {
  if (*(volatile int*)p.fix_flag == 0) return;
  const int64_t t = blockIdx.x;
  if (t >= p.T) return;
  const AggTile td = p.tiles[t];
  const AggInst I = p.inst[td.inst];
  for (int lr = threadIdx.x; lr < td.nrows; lr += blockDim.x) {
    const int64_t r = (int64_t)td.row_begin + lr, grow = td.row_base + lr;
    int64_t key;
    const bool kvalid = row_key(p, td.batch, r, &key);
    int64_t slot = -2;  // looked up on first need
    for (int l = 0; l < p.nl; l++) {
      const Lane ln = p.lanes[l];
      if (ln.op != L_NANBITS || !is_valid(p, td.batch, ln.col, r)) continue;
      const int ty = p.col_type[ln.col];
      const int64_t raw = raw_bits(abuf(p, td.batch, ln.col, 2), ty, r);
      if (!raw_is_nan(raw, ty)) continue;
      if (slot == -2) slot = table_find(I, kvalid, key);
      if (slot < 0) break;
      int64_t want = grow, have;
      if (ln.col2 < 0) {
        have = *tab_first(I, slot);
      } else {
        if (p.lanes[ln.col2].op == L_GLASTNAN) want = glast_ord(grow, r);
        have = tab_lanes(I, slot)[ln.col2];
      }
      if (have == want) tab_lanes(I, slot)[l] = raw;
    }
  }
}

// A batch of an aggregate without a group-by key.
struct GBatch {
  int32_t batch, inst;
  int64_t nrows, base;  // rows; instance-global index of its first row
};

// Integer min/max without a group-by key merge arrow MinMax per batch
// (MinAction::GetMinResult / MaxAction::GetMaxResult, ext/actions_impl.cc:995-1011,
// :1605-1621): a batch whose input is all null yields a null scalar of value 0,
// which is merged once an earlier batch has set the state (a non-null row before
// the batch's first row).  One workgroup per batch.
__global__ __launch_bounds__(256) void agg_allnull_kernel(AggParams p, const GBatch* gb, int nb) {
  __shared__ int s_any;
  const int b = blockIdx.x;
  if (b >= nb) return;
  const GBatch g = gb[b];
  const AggInst I = p.inst[g.inst];
  for (int l = 0; l < p.nl; l++) {
    const Lane ln = p.lanes[l];
    if (ln.op != L_MINI && ln.op != L_MAXI && ln.op != L_MINU) continue;  // unsigned max with 0 changes nothing
    const uint8_t* vb = (const uint8_t*)abuf(p, g.batch, ln.col, 0);
    if (!vb || g.nrows == 0) continue;  // no bitmap: every row valid
    __syncthreads();
    if (threadIdx.x == 0) s_any = 0;
    __syncthreads();
    const int64_t nbytes = (g.nrows + 7) / 8;
    for (int64_t i = threadIdx.x; i < nbytes; i += blockDim.x) {
      int m = vb[i];
      if (i == nbytes - 1 && (g.nrows & 7)) m &= (1 << (g.nrows & 7)) - 1;
      if (m) s_any = 1;
    }
    __syncthreads();
    if (threadIdx.x == 0 && !s_any) {
      int fl = -1;
      for (int k = 0; k < p.nl; k++)
        if (p.lanes[k].op == L_FIRSTNN && p.lanes[k].col == ln.col) fl = k;
      const int64_t slot = table_find(I, true, 0);
      if (fl >= 0 && slot >= 0 && tab_lanes(I, slot)[fl] < g.base) {
        int64_t* ptr = tab_lanes(I, slot) + l;
        if (ln.op == L_MINI) atomicMin((long long*)ptr, 0ll);
        else if (ln.op == L_MAXI) atomicMax((long long*)ptr, 0ll);
        else atomicMin((unsigned long long*)ptr, 0ull);
      }
    }
  }
}

// ---------------------------------------------------------------- fused join
// Join -> aggregate in one pass (whole-stage codegen's conditionedProbeArraysInner
// feeding hashAggregateArrays, whole_stage_codegen_kernel.cc:48-330): probe rows
// look up the build relation's hash table, and every joined (probe row, build
// row) pair updates the aggregate directly; the join output is never written.
// The aggregate's input columns are join result columns, each read from the
// probe batch or from the relation's row-indexed store.  The group key is the
// join key, so a group's first row is the first probe row with that key: the
// groups come out in the order of the join output, as the reference's.
struct JoinIn {
  int side;  // 0 build (relation store), 1 probe batch
  int idx;   // column on that side
  int type;
  const uint8_t* bvalid;  // build side: validity bytes
  const uint8_t* bvalues; // build side: values (bool: one byte per row)
};

struct JoinTile {
  int32_t batch, row_begin, nrows, pad;
  int64_t row_base;  // global index of the tile's first probe row
};

struct JoinAggParams {
  const JoinTile* tiles;
  int64_t T;
  const void* const* bufs;  // probe batches: [batch * bstride + col * 3 + k]
  int bstride;
  int pkey, key_type;       // probe key column and its type
  const int64_t* jkeys;
  const int32_t* jstart;
  const int32_t* jcount;
  const int32_t* jrowlist;  // null: unique keys
  uint64_t jcap;
  int64_t special_start, special_count;  // the key equal to kEmpty
  int tiles_per_block;
  JoinIn in[kAggMaxCols];
};

struct JoinRow {
  const JoinAggParams& q;
  int batch;
  int64_t r;   // probe row in its batch
  int64_t br;  // build row
  __device__ __forceinline__ const void* pb(int col, int k) const {
    return q.bufs[(int64_t)batch * q.bstride + q.in[col].idx * 3 + k];
  }
  __device__ __forceinline__ bool valid(int col) const {
    const JoinIn& c = q.in[col];
    if (c.side == 0) return c.bvalid[br] != 0;
    const uint8_t* vb = (const uint8_t*)pb(col, 0);
    return !vb || bit_get(vb, r);
  }
  __device__ __forceinline__ int64_t i64(int col) const {
    const JoinIn& c = q.in[col];
    if (c.side == 0) return c.type == GZ_TYPE_BOOL ? (int64_t)c.bvalues[br] : read_i64(c.bvalues, c.type, br);
    return read_i64(pb(col, 2), c.type, r);
  }
  __device__ __forceinline__ double f64(int col) const {
    const JoinIn& c = q.in[col];
    if (c.side == 0) return c.type == GZ_TYPE_BOOL ? (double)c.bvalues[br] : read_f64(c.bvalues, c.type, br);
    return read_f64(pb(col, 2), c.type, r);
  }
  __device__ __forceinline__ uint64_t u64(int col) const {
    const JoinIn& c = q.in[col];
    return c.side == 0 ? ((const uint64_t*)c.bvalues)[br] : ((const uint64_t*)pb(col, 2))[r];
  }
  __device__ __forceinline__ int64_t raw64(int col) const {
    const JoinIn& c = q.in[col];
    return c.side == 0 ? ((const int64_t*)c.bvalues)[br] : ((const int64_t*)pb(col, 2))[r];
  }
  __device__ __forceinline__ int64_t rib() const { return 0; }  // a join aggregate always has a key
};

// A joined row's used input columns loaded once into registers (integers
// widened as read_i64, floats as double bits, 8-byte types raw), so the lane
// updates read no memory: every input load of the pair is issued before the
// first update.
struct RegRow {
  const AggParams& p;
  int64_t w0, w1, w2, w3, w4, w5, w6, w7;
  uint32_t nulls;  // bit k: used column k is null
  __device__ __forceinline__ int64_t word(int col) const {
    // a select chain (a switch becomes a lookup table in scratch memory)
    const int u = p.uslot[col];
    const int64_t a = u & 1 ? w1 : w0, b = u & 1 ? w3 : w2, c = u & 1 ? w5 : w4, d = u & 1 ? w7 : w6;
    const int64_t e = u & 2 ? b : a, f = u & 2 ? d : c;
    return u & 4 ? f : e;
  }
  __device__ __forceinline__ bool valid(int col) const { return !((nulls >> p.uslot[col]) & 1); }
  __device__ __forceinline__ int64_t i64(int col) const { return word(col); }
  __device__ __forceinline__ double f64(int col) const {
    const int64_t x = word(col);
    const int t = p.col_type[col];
    if (t == GZ_TYPE_DOUBLE || t == GZ_TYPE_FLOAT) return __longlong_as_double(x);
    if (t == GZ_TYPE_UINT64) return (double)(uint64_t)x;
    return (double)x;
  }
  __device__ __forceinline__ uint64_t u64(int col) const { return (uint64_t)word(col); }
  __device__ __forceinline__ int64_t raw64(int col) const { return word(col); }
  __device__ __forceinline__ int64_t rib() const { return 0; }  // a join aggregate always has a key
};

// Word of a join input column for RegRow; *ok = not null.
__device__ __forceinline__ int64_t join_word(const JoinAggParams& q, int col, int batch, int64_t r, int64_t br,
                                             bool* ok) {
  const JoinIn& c = q.in[col];
  const void* v;
  int64_t row;
  if (c.side == 0) {
    *ok = c.bvalid[br] != 0;
    v = c.bvalues;
    row = br;
    if (c.type == GZ_TYPE_BOOL) return (int64_t)c.bvalues[br];
  } else {
    const uint8_t* vb = (const uint8_t*)q.bufs[(int64_t)batch * q.bstride + c.idx * 3];
    *ok = !vb || bit_get(vb, r);
    v = q.bufs[(int64_t)batch * q.bstride + c.idx * 3 + 2];
    row = r;
  }
  if (c.type == GZ_TYPE_DOUBLE) return ((const int64_t*)v)[row];
  if (c.type == GZ_TYPE_FLOAT) return __double_as_longlong((double)((const float*)v)[row]);
  return read_i64(v, c.type, row);
}

// Persistent workgroups over consecutive probe tiles, an LDS first-level group
// table each (as the plain aggregate's LDS path: a key that finds no LDS slot
// within a short probe goes to the HBM table), merged into HBM at the end.
__global__ __launch_bounds__(256) void join_agg_update_kernel(AggParams p, JoinAggParams q) {
  extern __shared__ int64_t s_tab[];
  const int64_t LS = p.lcap + 3;
  const uint64_t lmask = (uint64_t)p.lcap - 1;
  const int64_t t0 = (int64_t)blockIdx.x * q.tiles_per_block;
  const int64_t t1 = min(q.T, t0 + q.tiles_per_block);
  if (t0 >= t1) return;
  lds_reset(p, s_tab, LS);
  __syncthreads();
  const AggInst I = p.inst[0];
  const uint64_t jmask = q.jcap - 1;
  for (int64_t t = t0; t < t1; t++) {
    const JoinTile td = q.tiles[t];
    const uint8_t* kv = (const uint8_t*)q.bufs[(int64_t)td.batch * q.bstride + q.pkey * 3];
    const void* kd = q.bufs[(int64_t)td.batch * q.bstride + q.pkey * 3 + 2];
    for (int lr = threadIdx.x; lr < td.nrows; lr += blockDim.x) {
      const int64_t r = (int64_t)td.row_begin + lr;
      if (kv && !bit_get(kv, r)) continue;  // a null key joins nothing
      const int64_t key = read_i64(kd, q.key_type, r);
      int64_t st, cnt;
      if (key == kEmpty) {
        if (q.special_count == 0) continue;
        st = q.special_start;
        cnt = q.jrowlist ? q.special_count : 1;
      } else {
        uint64_t s = fmix64((uint64_t)key) & jmask;
        int64_t m = -1;
        for (uint64_t i = 0; i < q.jcap; i++) {
          const int64_t k = gld(q.jkeys + s);
          if (k == key) {
            m = (int64_t)s;
            break;
          }
          if (k == kEmpty) break;
          s = (s + 1) & jmask;
        }
        if (m < 0) continue;
        st = gld(q.jstart + m);
        cnt = q.jrowlist ? gld(q.jcount + m) : 1;
      }
      const int64_t grow = td.row_base + lr;
      int64_t ls = -1;
      if (key == kEmpty) {
        ls = p.lcap + 1;
      } else {
        uint64_t s = fmix64((uint64_t)key) & lmask;
        for (int i = 0; i < p.lprobe; i++) {  // short probe: a full LDS table spills to HBM
          const int64_t k = s_tab[s];
          if (k == key) {
            ls = (int64_t)s;
            break;
          }
          if (k == kEmpty) {
            const unsigned long long prev =
                atomicCAS((unsigned long long*)&s_tab[s], (unsigned long long)kEmpty, (unsigned long long)key);
            if ((int64_t)prev == kEmpty || (int64_t)prev == key) {
              ls = (int64_t)s;
              break;
            }
          }
          s = (s + 1) & lmask;
        }
      }
      const int64_t g = ls >= 0 ? -1 : table_slot(I, true, key);
      if (ls < 0 && g < 0) continue;  // HBM table overflow (flagged)
      for (int64_t j = 0; j < cnt; j++) {
        const int64_t br = q.jrowlist ? (int64_t)gld(q.jrowlist + st + j) : st;
        if (p.nu >= 0) {
          int64_t w[kMaxUsed];
          uint32_t nulls = 0;
#pragma unroll
          for (int k = 0; k < kMaxUsed; k++) {
            bool ok = true;
            w[k] = k < p.nu ? join_word(q, p.ucol[k], td.batch, r, br, &ok) : 0;
            if (!ok) nulls |= 1u << k;
          }
          const RegRow row{p, w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7], nulls};
          if (ls >= 0) apply_row_lds(p, s_tab, LS, ls, row, grow);
          else apply_row_hbm(p, I, g, row, grow);
        } else {
          const JoinRow row{q, td.batch, r, br};
          if (ls >= 0) apply_row_lds(p, s_tab, LS, ls, row, grow);
          else apply_row_hbm(p, I, g, row, grow);
        }
      }
    }
  }
  __syncthreads();
  lds_merge(p, I, s_tab, LS);
}

// Fix-up pass of the fused join -> aggregate (agg_fix_kernel's rule over joined
// pairs): one workgroup per probe tile walks the relation again, and a NaN pair
// whose probe row is the one its group's row lane kept stores its bits.  Pairs
// of one probe row share the ordinal; their values are equal whenever that
// matters (a build-side float min / max over repeated build keys is refused).
__global__ __launch_bounds__(256) void join_agg_fix_kernel(AggParams p, JoinAggParams q) {
  if (*(volatile int*)p.fix_flag == 0) return;
  const int64_t t = blockIdx.x;
  if (t >= q.T) return;
  const JoinTile td = q.tiles[t];
  const AggInst I = p.inst[0];
  const uint64_t jmask = q.jcap - 1;
  const uint8_t* kv = (const uint8_t*)q.bufs[(int64_t)td.batch * q.bstride + q.pkey * 3];
  const void* kd = q.bufs[(int64_t)td.batch * q.bstride + q.pkey * 3 + 2];
  for (int lr = threadIdx.x; lr < td.nrows; lr += blockDim.x) {
    const int64_t r = (int64_t)td.row_begin + lr;
    if (kv && !bit_get(kv, r)) continue;
    const int64_t key = read_i64(kd, q.key_type, r);
    int64_t st, cnt;
    if (key == kEmpty) {
      if (q.special_count == 0) continue;
      st = q.special_start;
      cnt = q.jrowlist ? q.special_count : 1;
    } else {
      uint64_t s = fmix64((uint64_t)key) & jmask;
      int64_t m = -1;
      for (uint64_t i = 0; i < q.jcap; i++) {
        const int64_t k = q.jkeys[s];
        if (k == key) {
          m = (int64_t)s;
          break;
        }
        if (k == kEmpty) break;
        s = (s + 1) & jmask;
      }
      if (m < 0) continue;
      st = q.jstart[m];
      cnt = q.jrowlist ? q.jcount[m] : 1;
    }
    const int64_t g = table_find(I, true, key);
    if (g < 0) continue;
    const int64_t grow = td.row_base + lr;
    for (int64_t j = 0; j < cnt; j++) {
      const int64_t br = q.jrowlist ? (int64_t)q.jrowlist[st + j] : st;
      for (int l = 0; l < p.nl; l++) {
        const Lane ln = p.lanes[l];
        if (ln.op != L_NANBITS) continue;
        const JoinIn& c = q.in[ln.col];
        int64_t raw;
        if (c.side == 0) {
          if (!c.bvalid[br]) continue;
          raw = raw_bits(c.bvalues, c.type, br);
        } else {
          const uint8_t* vb = (const uint8_t*)q.bufs[(int64_t)td.batch * q.bstride + c.idx * 3];
          if (vb && !bit_get(vb, r)) continue;
          raw = raw_bits(q.bufs[(int64_t)td.batch * q.bstride + c.idx * 3 + 2], c.type, r);
        }
        if (!raw_is_nan(raw, c.type)) continue;
        const int64_t have = ln.col2 < 0 ? *tab_first(I, g) : tab_lanes(I, g)[ln.col2];
        if (have == grow) tab_lanes(I, g)[l] = raw;
      }
    }
  }
}

// ---------------------------------------------------------------- partitioned
// Partitioned placement for medium cardinalities: the rows of an instance are
// scattered into 2^fan_log2 partitions by the high bits of the key hash (the LDS
// table indexes by the low bits), as fixed-size records carrying the row's
// instance-global index, so the partitions need no stable order.  Each
// partition's groups then fit an LDS table; workgroups aggregate chunks of
// records in LDS and merge into the instance's HBM table, which partitions of
// different keys never share.
//
// Record (8-byte words): key, first-row index | null mask << 48 (bit 15 = key
// null), then the used input columns widened to 8 bytes.
constexpr int kRecKeyNull = 15;
constexpr int kPartTile = 1024;       // rows per tile of the partition kernels

__device__ __forceinline__ int part_of(const AggInst& I, bool valid, int64_t key) {
  if (!valid || I.fan_log2 == 0) return 0;
  return (int)(fmix64((uint64_t)key) >> (64 - I.fan_log2));
}

// Rows per partition of every tile (LDS histogram), written partition-major:
// instance i's counts are a [F][tiles] block at cnt_off, so one exclusive scan
// over all blocks gives every (partition, tile) its first record, with the
// records of a partition contiguous and tile-ordered.  No global atomics.
__global__ __launch_bounds__(256) void agg_part_hist_kernel(AggParams p, uint32_t* cnt) {
  __shared__ unsigned int s_h[kMaxFanout];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const AggTile td = p.tiles[t];
  const AggInst I = p.inst[td.inst];
  const int F = 1 << I.fan_log2;
  const int64_t tl = p.tile_base + t - I.tile0;
  constexpr int RPT = kPartTile / 256;
  int64_t key[RPT];
  const uint32_t vm = load_keys<RPT, 256>(p, td.batch, td.row_begin, threadIdx.x, td.nrows, key);
  for (int b = threadIdx.x; b < F; b += blockDim.x) s_h[b] = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (k * 256 + (int)threadIdx.x < td.nrows) atomicAdd(&s_h[part_of(I, (vm >> k) & 1, key[k])], 1u);
  __syncthreads();
  for (int b = threadIdx.x; b < F; b += blockDim.x) gst(cnt + I.cnt_off + (int64_t)b * I.cnt_tiles + tl, s_h[b]);
}

// A work item of the partitioned update: records [a, e) of partition b.  An
// owned item holds the whole partition, so its workgroup is the only writer of
// the partition's groups; a partition of more than kPartItemRecords records is
// cut into items of that many, merged with atomics.
struct PartItem {
  int32_t b, owned;
  int64_t a, e;
};
constexpr int64_t kPartItemRecords = 32768;

// First record of global partition g = (instance pinst[g], partition b): the
// scanned count of its first tile; past the last partition, nrec.
__device__ __forceinline__ int64_t part_first(const uint64_t* prefix, const AggInst* inst, const int* pinst,
                                              int B, int64_t nrec, int g) {
  if (g >= B) return nrec;
  const AggInst& I = inst[pinst[g]];
  return (int64_t)prefix[I.cnt_off + (int64_t)(g - I.bucket_base) * I.cnt_tiles];
}

// The work items: item g < B is partition g's first (an empty partition's is
// empty); a partition of more than kPartItemRecords records appends its other
// items after the first B, counted at *extra (zero before the launch).  One
// thread per partition.
__global__ __launch_bounds__(256) void agg_part_items_kernel(const uint64_t* prefix, const AggInst* inst,
                                                             const int* pinst, int B, int64_t nrec, PartItem* items,
                                                             unsigned long long* extra) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B) return;
  const int64_t a0 = part_first(prefix, inst, pinst, B, nrec, g);
  const int64_t v = part_first(prefix, inst, pinst, B, nrec, g + 1) - a0;
  const int64_t k = (v + kPartItemRecords - 1) / kPartItemRecords;  // items of partition g
  items[g] = PartItem{g, k <= 1 ? 1 : 0, a0, a0 + min(v, kPartItemRecords)};
  if (k > 1) {
    const int64_t at = B + (int64_t)atomicAdd(extra, (unsigned long long)(k - 1));
    for (int64_t j = 1; j < k; j++) {
      const int64_t a = a0 + j * kPartItemRecords;
      items[at + j - 1] = PartItem{g, 0, a, min(a + kPartItemRecords, a0 + v)};
    }
  }
}

// Reset of a partition-structured table's partitions, by what the work items
// do with them: a partition of more than kPartItemRecords records (its items
// merge with atomics) gets every slot reset, an empty one only its keys and
// first rows (no item writes it; the collection reads first rows); an owned
// partition nothing (its item stores every slot's key and first row and every
// group's lanes).  One workgroup per partition.  lane_op(l): lane l's operation.
template <class LaneOp>
__device__ void part_slots_reset(const AggInst& I, int b, bool bare, int nl, LaneOp lane_op) {
  const int64_t base = (int64_t)b << I.sub_log2;
  const int64_t n = (int64_t)1 << I.sub_log2;
  for (int64_t s = threadIdx.x; s < n; s += blockDim.x) {
    I.base[base + s] = kEmpty;
    *tab_first(I, base + s) = INT64_MAX;
    if (bare) continue;
    int64_t* ln = tab_lanes(I, base + s);
    for (int l = 0; l < nl; l++) ln[l] = lane_init(lane_op(l));
  }
}

__global__ __launch_bounds__(256) void agg_part_reset_kernel(const uint64_t* prefix, const AggInst* inst,
                                                             const int* pinst, int B, int64_t nrec, int nl,
                                                             const int* lane_ops) {
  const int g = blockIdx.x;
  if (g >= B) return;
  const int64_t v = part_first(prefix, inst, pinst, B, nrec, g + 1) - part_first(prefix, inst, pinst, B, nrec, g);
  if (v > 0 && v <= kPartItemRecords) return;
  const AggInst I = inst[pinst[g]];
  part_slots_reset(I, g - I.bucket_base, v == 0, nl, [&](int l) { return lane_ops[l]; });
}

// Records of a tile: partition counts in LDS, one range reserved per non-empty
// partition of the tile, the tile's records laid out in LDS in partition order
// (word-major), then written out as one contiguous run per partition and word
// (consecutive threads, consecutive records).  Tiles of this kernel are at most
// kPartTile rows.  The key and the first kPreU used columns are loaded before
// the ranking; the other used columns after it.
constexpr int kPreU = 2;

// Dynamic LDS of the scatter for fan-outs up to fcap: the tile's records, then
// per partition its first record, rows and tile offset, then the partition of
// every sorted slot.
__host__ __device__ constexpr size_t scatter_lds(int W, int fcap) {
  return sizeof(int64_t) * (size_t)W * kPartTile + sizeof(uint64_t) * fcap + sizeof(uint32_t) * (2 * fcap + 1) +
         sizeof(uint16_t) * kPartTile;
}

__global__ __launch_bounds__(256) void agg_part_scatter_kernel(AggParams p, const uint64_t* prefix, int fcap,
                                                               int64_t* recs) {
  extern __shared__ int64_t s_rec[];  // [W][kPartTile], then the per-partition arrays
  __shared__ uint64_t s_scan[256 / kWave + 1];
  int64_t t = xcd_tile(blockIdx.x, gridDim.x, p.T);
  if (t < 0) return;
  const AggTile td = p.tiles[t];
  const AggInst I = p.inst[td.inst];
  const int F = 1 << I.fan_log2;
  const int W = 2 + p.nu;
  unsigned long long* s_base = (unsigned long long*)(s_rec + (size_t)W * kPartTile);
  unsigned int* s_h = (unsigned int*)(s_base + fcap);
  unsigned int* s_off = s_h + fcap;
  uint16_t* s_pt = (uint16_t*)(s_off + fcap + 1);
  const int64_t tl = p.tile_base + t - I.tile0;
  constexpr int RPT = kPartTile / 256;
  const int tid = threadIdx.x, n = td.nrows;
  const int64_t rb = td.row_begin;
  const void* const* bt = p.bufs + (int64_t)td.batch * p.bstride;
  // ---- loads that do not depend on the partition
  int64_t key[RPT];
  const uint32_t kv = load_keys<RPT, 256>(p, td.batch, rb, tid, n, key);
  int64_t pre[kPreU][RPT];
  uint32_t pvalid[kPreU];
#pragma unroll
  for (int u = 0; u < kPreU; u++) {
    pvalid[u] = 0;
#pragma unroll
    for (int k = 0; k < RPT; k++) pre[u][k] = 0;
    if (u < p.nu) {
      const int c = p.ucol[u];
      uint8_t vx[RPT];
      valid_bytes<RPT, 256>((const uint8_t*)bt[c * 3], rb, tid, n, vx);
      load_words<RPT, 256>(bt[c * 3 + 2], p.col_type[c], rb, tid, n, pre[u]);
      pvalid[u] = valid_mask<RPT, 256>(vx, rb, tid, n);
    }
  }
  for (int b = tid; b < F; b += blockDim.x) s_h[b] = 0;
  __syncthreads();
  // partition and in-partition rank of the owned rows (rank order is free: the
  // record carries the row index)
  int part[RPT], rank[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    part[k] = 0;
    rank[k] = 0;
    if (k * 256 + tid < n) {
      part[k] = part_of(I, (kv >> k) & 1, key[k]);
      rank[k] = (int)atomicAdd(&s_h[part[k]], 1u);
    }
  }
  __syncthreads();
  {
    // tile offsets of the partitions (256 at a time) and each partition's first
    // record for this tile
    uint64_t carry = 0;
    for (int b0 = 0; b0 < F; b0 += 256) {
      const int b = b0 + tid;
      const uint64_t c = b < F ? s_h[b] : 0;
      uint64_t total;
      const uint64_t ex = block_exclusive_scan<uint64_t>(c, s_scan, &total);
      if (b < F) {
        s_off[b] = (unsigned)(carry + ex);
        // only partitions with records in this tile are read below (at a
        // fan-out of 4096 a 1024-row tile holds about 230 of them)
        if (c) s_base[b] = (unsigned long long)prefix[I.cnt_off + (int64_t)b * I.cnt_tiles + tl];
      }
      carry += total;
    }
    if (tid == 0) s_off[F] = (unsigned)carry;
  }
  __syncthreads();
  int slot[RPT];
  uint32_t nulls[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    slot[k] = (int)s_off[part[k]] + rank[k];
    const bool kval = (kv >> k) & 1;
    nulls[k] = kval ? 0u : (1u << kRecKeyNull);
#pragma unroll
    for (int u = 0; u < kPreU; u++)
      if (u < p.nu && !((pvalid[u] >> k) & 1)) nulls[k] |= 1u << u;
    if (k * 256 + tid < n) {
      s_rec[slot[k]] = kval ? key[k] : 0;
#pragma unroll
      for (int u = 0; u < kPreU; u++)
        if (u < p.nu) s_rec[(2 + u) * kPartTile + slot[k]] = ((pvalid[u] >> k) & 1) ? pre[u][k] : 0;
      s_pt[slot[k]] = (uint16_t)part[k];
    }
  }
  for (int u = kPreU; u < p.nu; u++) {
    const int c = p.ucol[u];
    uint8_t vx[RPT];
    int64_t v[RPT];
    valid_bytes<RPT, 256>((const uint8_t*)bt[c * 3], rb, tid, n, vx);
    load_words<RPT, 256>(bt[c * 3 + 2], p.col_type[c], rb, tid, n, v);
    const uint32_t vm = valid_mask<RPT, 256>(vx, rb, tid, n);
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      if (!((vm >> k) & 1)) nulls[k] |= 1u << u;
      if (k * 256 + tid < n) s_rec[(2 + u) * kPartTile + slot[k]] = ((vm >> k) & 1) ? v[k] : 0;
    }
  }
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (k * 256 + tid < n)
      s_rec[kPartTile + slot[k]] = (int64_t)((uint64_t)(td.row_base + k * 256 + tid) | ((uint64_t)nulls[k] << 48));
  __syncthreads();
  // word-major runs: thread tid writes sorted slots tid + k * 256 of every word
  int64_t pos[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const int sl = k * 256 + tid;
    pos[k] = 0;
    if (sl < n) {
      const int b = s_pt[sl];
      pos[k] = (int64_t)(s_base[b] + (unsigned long long)(sl - (int)s_off[b]));
    }
  }
  for (int w = 0; w < W; w++) {
    int64_t* dst = recs + (int64_t)w * p.rstride;
#pragma unroll
    for (int k = 0; k < RPT; k++)
      if (k * 256 + tid < n) gst(dst + pos[k], s_rec[w * kPartTile + k * 256 + tid]);
  }
}

#ifndef GZ_PART_GROUPS
#define GZ_PART_GROUPS 512
#endif
#ifndef GZ_UPD_RPT
#define GZ_UPD_RPT 4
#endif
constexpr int kUpdRPT = GZ_UPD_RPT;      // records per thread per step of the update kernel

// Workgroups take the work items in turn: an item's records are aggregated in
// an LDS table, which an owned item (a whole partition) writes out with one
// slot claim per group and plain stores, and an item of a larger partition
// merges with atomics.
// UB: the workgroup size (512: two workgroups per CU, 256: four); the
// autotuner picks one.
template <int UB>
__global__ __launch_bounds__(UB, 2 * UB / 256) void agg_part_update_kernel(AggParams p, const int64_t* recs,
                                                                           const PartItem* items, int B,
                                                                           const unsigned long long* extra,
                                                                           const int* pinst) {
  extern __shared__ int64_t s_tab[];
  __shared__ int s_failed;        // a record of the item missed the LDS table
  const int64_t LS = p.lcap + 3;  // stride: lcap slots, 2 special, 1 spare
  const uint64_t lmask = (uint64_t)p.lcap - 1;
  const int64_t NR = p.rstride;
  const int64_t ni = B + (int64_t)*extra;
  for (int64_t j = blockIdx.x; j < ni; j += gridDim.x) {
    {
      const PartItem it = items[j];
      const int64_t a = it.a, e = it.e;
      if (a == e) continue;  // an empty partition: its keys were reset (agg_part_reset_kernel)
      const AggInst I = p.inst[pinst[it.b]];
      __syncthreads();
      if (threadIdx.x == 0) s_failed = 0;
      lds_reset(p, s_tab, LS);
      __syncthreads();
      // kUpdRPT records per thread per step, their key, row and first two
      // used words loaded together
      for (int64_t q0 = a + threadIdx.x; q0 < e; q0 += (int64_t)blockDim.x * kUpdRPT) {
        int64_t kw[kUpdRPT], w1w[kUpdRPT], u0[kUpdRPT], u1[kUpdRPT];
#pragma unroll
        for (int k = 0; k < kUpdRPT; k++) {
          const int64_t q = q0 + (int64_t)k * blockDim.x;
          const bool act = q < e;
          kw[k] = act ? gld(recs + q) : 0;
          w1w[k] = act ? gld(recs + NR + q) : 0;
          u0[k] = act && p.nu > 0 ? gld(recs + 2 * NR + q) : 0;
          u1[k] = act && p.nu > 1 ? gld(recs + 3 * NR + q) : 0;
        }
        // LDS slot of every record (-1: none, -2: the LDS probe failed)
        int ls[kUpdRPT];
#pragma unroll
        for (int k = 0; k < kUpdRPT; k++) {
          const int64_t q = q0 + (int64_t)k * blockDim.x;
          ls[k] = -1;
          if (q >= e) continue;
          const uint32_t nulls = (uint32_t)((uint64_t)w1w[k] >> 48);
          const int64_t key = kw[k];
          if ((nulls >> kRecKeyNull) & 1) ls[k] = p.lcap;
          else if (key == kEmpty) ls[k] = p.lcap + 1;
          else {
            ls[k] = -2;
            uint64_t s = fmix64((uint64_t)key) & lmask;
            for (int i = 0; i < 32; i++) {  // short probe: a full LDS table spills to HBM
              int64_t kk = s_tab[s];
              if (kk == key) { ls[k] = (int)s; break; }
              if (kk == kEmpty) {
                unsigned long long prev = atomicCAS((unsigned long long*)&s_tab[s], (unsigned long long)kEmpty,
                                                    (unsigned long long)key);
                if ((int64_t)prev == kEmpty || (int64_t)prev == key) { ls[k] = (int)s; break; }
              }
              s = (s + 1) & lmask;
            }
          }
        }
        auto grow_of = [&](int k) { return (int64_t)((uint64_t)w1w[k] & ((1ull << 48) - 1)); };
        auto row_of = [&](int k) {
          return RecRow{p, u0[k], u1[k], recs + q0 + (int64_t)k * blockDim.x, (uint32_t)((uint64_t)w1w[k] >> 48)};
        };
        if (p.nu <= 2) {
          int slot[kUpdRPT];
#pragma unroll
          for (int k = 0; k < kUpdRPT; k++) slot[k] = ls[k] >= 0 ? ls[k] : -1;
          lds_apply_records<kUpdRPT>(p, s_tab, LS, slot, -1, w1w, u0, u1);
        } else {
#pragma unroll
          for (int k = 0; k < kUpdRPT; k++)
            if (ls[k] >= 0) apply_row_lds(p, s_tab, LS, ls[k], row_of(k), grow_of(k));
        }
        // records whose LDS probe failed go to the instance's HBM table: at
        // once for a chunk of a shared partition (reset before the launch),
        // after the LDS pass for an owned one (its slots are not reset)
#pragma unroll
        for (int k = 0; k < kUpdRPT; k++) {
          if (ls[k] != -2) continue;
          s_failed = 1;
          if (it.owned) continue;
          const int64_t g = table_slot(I, true, kw[k]);
          if (g >= 0) apply_row_hbm(p, I, g, row_of(k), grow_of(k));
        }
      }
      __syncthreads();
      if (!it.owned) {
        lds_merge(p, I, s_tab, LS);
      } else if (!s_failed) {
        lds_write_image(p, I, it.b - I.bucket_base, s_tab, LS);
      } else {
        // the LDS table overflowed: reset the partition's slots, claim a slot
        // for every group of the LDS table and store it, then apply the
        // records whose keys the LDS table lacks with atomics (a key is either
        // in the LDS table for all its records or for none)
        part_slots_reset(I, it.b - I.bucket_base, false, p.nl, [&](int l) { return p.lanes[l].op; });
        __threadfence();
        __syncthreads();
        lds_writeout(p, I, s_tab, LS);
        __threadfence();
        __syncthreads();
        for (int64_t q = a + threadIdx.x; q < e; q += blockDim.x) {
          const int64_t key = gld(recs + q);
          const int64_t w1 = gld(recs + NR + q);
          const uint32_t nulls = (uint32_t)((uint64_t)w1 >> 48);
          if (((nulls >> kRecKeyNull) & 1) || key == kEmpty) continue;  // special slots never miss
          bool found = false;
          uint64_t sl = fmix64((uint64_t)key) & lmask;
          for (int i = 0; i < 32; i++) {
            const int64_t kk = s_tab[sl];
            if (kk == key) {
              found = true;
              break;
            }
            if (kk == kEmpty) break;
            sl = (sl + 1) & lmask;
          }
          if (found) continue;
          const int64_t u0 = p.nu > 0 ? gld(recs + 2 * NR + q) : 0;
          const int64_t u1 = p.nu > 1 ? gld(recs + 3 * NR + q) : 0;
          const int64_t g = table_slot(I, true, key);
          if (g >= 0)
            apply_row_hbm(p, I, g, RecRow{p, u0, u1, recs + q, nulls}, (int64_t)((uint64_t)w1 & ((1ull << 48) - 1)));
        }
      }
    }
  }
}

// HyperLogLog registers per instance (2^11 registers), persistent static chunks
// with LDS registers merged by atomicMax.
__global__ __launch_bounds__(256) void agg_hll_kernel(AggParams p, unsigned int* regs) {
  __shared__ unsigned int s_reg[kHllRegs];
  int64_t t0 = (int64_t)blockIdx.x * p.tiles_per_block;
  int64_t t1 = min(p.T, t0 + p.tiles_per_block);
  if (t0 >= t1) return;
  int cur = -1;
  for (int64_t t = t0; t < t1; t++) {
    const AggTile td = p.tiles[t];
    if (td.inst != cur) {
      __syncthreads();
      if (cur >= 0)
        for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x)
          if (s_reg[i]) atomicMax(&regs[(int64_t)cur * kHllRegs + i], s_reg[i]);
      __syncthreads();
      for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x) s_reg[i] = 0;
      __syncthreads();
      cur = td.inst;
    }
    constexpr int RPT = 8;  // rows per thread per step, loaded together
    for (int base = 0; base < td.nrows; base += 256 * RPT) {
      int64_t key[RPT];
      const uint32_t vm = load_keys<RPT, 256>(p, td.batch, td.row_begin, base + (int)threadIdx.x, td.nrows, key);
#pragma unroll
      for (int k = 0; k < RPT; k++) {
        if (!((vm >> k) & 1)) continue;
        uint64_t h = fmix64((uint64_t)key[k] ^ 0x9e3779b97f4a7c15ull);
        unsigned idx = (unsigned)(h >> (64 - kHllBits));
        uint64_t w = (h << kHllBits) | (1ull << (kHllBits - 1));
        unsigned rank = (unsigned)__clzll((long long)w) + 1;
        if (rank > s_reg[idx]) atomicMax(&s_reg[idx], rank);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kHllRegs; i += blockDim.x)
    if (s_reg[i]) atomicMax(&regs[(int64_t)cur * kHllRegs + i], s_reg[i]);
}

// One block per instance: the HyperLogLog estimate (with the linear-counting
// correction for small cardinalities).
__global__ __launch_bounds__(256) void agg_hll_estimate_kernel(const unsigned int* regs, double* est) {
  __shared__ double s_sum[256];
  __shared__ int s_zero[256];
  const unsigned int* r = regs + (int64_t)blockIdx.x * kHllRegs;
  double sum = 0;
  int zeros = 0;
  for (int i = threadIdx.x; i < kHllRegs; i += 256) {
    sum += ldexp(1.0, -(int)r[i]);
    zeros += r[i] == 0;
  }
  s_sum[threadIdx.x] = sum;
  s_zero[threadIdx.x] = zeros;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + o];
      s_zero[threadIdx.x] += s_zero[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = kHllRegs;
    double e = (0.7213 / (1.0 + 1.079 / m)) * m * m / s_sum[0];
    if (e <= 2.5 * m && s_zero[0] > 0) e = m * log(m / (double)s_zero[0]);
    est[blockIdx.x] = e;
  }
}

// Rehash of an instance into a larger table (keys are unique: plain stores).
__global__ void agg_rehash_kernel(AggInst from, AggInst to, int nl) {
  const int64_t FS = (int64_t)from.cap + 2;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < FS;
       s += (int64_t)gridDim.x * blockDim.x) {
    int64_t key = from.base[s];
    int64_t d;
    if (s >= (int64_t)from.cap) {
      if (*tab_first(from, s) == INT64_MAX) continue;
      d = (int64_t)to.cap + (s - (int64_t)from.cap);
    } else {
      if (key == kEmpty) continue;
      d = table_slot(to, true, key);
      if (d < 0) continue;
    }
    *tab_first(to, d) = *tab_first(from, s);
    const int64_t* fl = tab_lanes(from, s);
    int64_t* tl = tab_lanes(to, d);
    for (int l = 0; l < nl; l++) tl[l] = fl[l];
  }
}

// Occupied slots of every instance, compacted without atomics, as
// (instance << rbits | first_row, slot) pairs (rbits: bits of the largest
// instance's row count): blocks own chunks of kCollectChunk slots of one
// instance, numbered flat over the instances (instance i's chunks from
// cpre[i]); a first pass counts each chunk, a scan over the chunks gives every
// chunk its output position, a second pass writes (instance, first row) sort
// keys and slot numbers in slot order.  A slot is occupied iff a row reached it.
constexpr int kCollectChunk = 2048;  // 256 threads x 8 slots

// The instance of flat chunk g (cpre: [ninst + 1] chunk prefix), and its chunk
// number within the instance.
__device__ __forceinline__ int chunk_inst(const int64_t* cpre, int ninst, int64_t g, int64_t* x) {
  int lo = 0, hi = ninst - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (cpre[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  *x = g - cpre[lo];
  return lo;
}

__global__ __launch_bounds__(256) void agg_occupancy_kernel(const AggInst* inst, const int64_t* cpre, int ninst,
                                                            unsigned long long* cc) {
  __shared__ uint64_t s_scan[256 / kWave + 1];
  int64_t x;
  const int i = chunk_inst(cpre, ninst, blockIdx.x, &x);
  const AggInst I = inst[i];
  const int64_t S = (int64_t)I.cap + 2;
  const int64_t s0 = x * kCollectChunk + threadIdx.x * 8;
  uint64_t c = 0;
  if (x * kCollectChunk < S) {
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (s0 + k < S && I.base[S + s0 + k] != INT64_MAX) c++;
  }
  uint64_t total;
  block_exclusive_scan<uint64_t>(c, s_scan, &total);
  if (threadIdx.x == 0) cc[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void agg_collect_kernel(const AggInst* inst, const int64_t* cpre, int ninst,
                                                          const unsigned long long* ccpre, int rbits,
                                                          unsigned long long* keys, unsigned long long* vals) {
  __shared__ uint64_t s_scan[256 / kWave + 1];
  int64_t x;
  const int i = chunk_inst(cpre, ninst, blockIdx.x, &x);
  const AggInst I = inst[i];
  const int64_t S = (int64_t)I.cap + 2;
  if (x * kCollectChunk >= S) return;
  const int64_t s0 = x * kCollectChunk + threadIdx.x * 8;
  int64_t fr[8];
  uint64_t c = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    fr[k] = s0 + k < S ? I.base[S + s0 + k] : INT64_MAX;
    c += fr[k] != INT64_MAX;
  }
  uint64_t pos = ccpre[blockIdx.x] + block_exclusive_scan<uint64_t>(c, s_scan, nullptr);
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (fr[k] != INT64_MAX) {
      keys[pos] = ((unsigned long long)i << rbits) | (unsigned long long)fr[k];
      vals[pos] = (unsigned long long)(s0 + k);
      pos++;
    }
  }
}

// Groups per instance from the chunk scan.
// Occupied slots per instance, and at counts[ninst] the number of instances
// whose table overflowed.
__global__ void agg_inst_counts_kernel(const AggInst* inst, const unsigned long long* cc,
                                       const unsigned long long* ccpre, const int64_t* cpre, int ninst,
                                       unsigned long long* counts) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  const int64_t last = cpre[i + 1] - 1;  // every instance has a chunk (cap + 2 >= 2 slots)
  counts[i] = ccpre[last] + cc[last] - ccpre[cpre[i]];
  if (inst[i].overflow && *inst[i].overflow) atomicAdd(counts + ninst, 1ull);
}

// First-seen order for few groups: the sort keys are distinct, so a key's
// place is the number of smaller keys.  A workgroup ranks 64 keys (one per
// lane); its eight waves split all keys between them and go through their share
// 64 at a time: each lane loads one key, and the wave compares its own keys
// with all 64 via readlane (no memory access per comparison).
constexpr int64_t kRankSortMax = 16384;
constexpr int kRankWaves = 8;

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int k) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, k);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), k);
  return (unsigned long long)hi << 32 | lo;
}

__global__ __launch_bounds__(64 * kRankWaves) void agg_rank_sort_kernel(const unsigned long long* __restrict__ keys,
                                                                        const unsigned long long* __restrict__ vals,
                                                                        int64_t n, unsigned long long* skeys,
                                                                        unsigned long long* svals) {
  __shared__ unsigned s_c[kRankWaves][64];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), l = threadIdx.x & 63;
  const int64_t i = (int64_t)blockIdx.x * 64 + l;
  const unsigned long long mine = i < n ? keys[i] : ~0ull;
  const int64_t q0 = n * w / kRankWaves, q1 = n * (w + 1) / kRankWaves;
  unsigned c = 0;
  for (int64_t j0 = q0; j0 < q1; j0 += 64) {
    // keys past the share read as all ones: never below a key (keys are below 2^63)
    const unsigned long long kk = j0 + l < q1 ? keys[j0 + l] : ~0ull;
#pragma unroll
    for (int k = 0; k < 64; k++) c += readlane_u64(kk, k) < mine ? 1u : 0u;
  }
  s_c[w][l] = c;
  __syncthreads();
  if (w == 0 && i < n) {
    unsigned r = 0;
#pragma unroll
    for (int x = 0; x < kRankWaves; x++) r += s_c[x][l];
    skeys[r] = mine;
    svals[r] = vals[i];
  }
}

// Keys inserted and overflow flag of every instance, gathered for one copy:
// out[2 * i] = keys inserted, out[2 * i + 1] = overflow.
__global__ void agg_read_counters_kernel(const AggInst* inst, int ninst, unsigned int* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ninst) return;
  const AggInst I = inst[i];  // an instance without a table yet reads as empty
  out[2 * i] = I.counter ? *I.counter : 0u;
  out[2 * i + 1] = I.overflow ? (unsigned)*I.overflow : 0u;
}

struct ResultSpec {
  int kind;      // action kind
  int sub;       // 0 / 1 (sum_count yields two columns)
  int rtype;     // result type
  int lane_a, lane_b, lane_c, lane_d, lane_e, lane_f;  // lanes used
  int nan_check;
  int global;    // no group-by key
  // floating min / max: lane_a ordered value, lane_b non-null count, lane_c
  // first zero (L_FIRSTZ), lane_d the NaN row lane, lane_e its bits
  // (L_NANBITS), lane_f first non-null row; fmode says when the result is the
  // kept NaN: 0 when no non-NaN value was seen (min with NaN check, and every
  // min without a group-by key), 1 when any NaN was seen (max with NaN check),
  // 2 when the first non-null value was NaN (plain compares).
  // A float group key: lane_c first zero, lane_e the first row's NaN bits.
  int fmode;
};

constexpr int kMaxResults = 40;

struct FinishParams {
  const AggInst* inst;
  const unsigned long long* skeys;  // sorted keys
  const unsigned long long* svals;  // slots
  int64_t n;
  int key_type;
  int rbits;  // sort key: instance << rbits | first row
  int nl;     // lanes per slot
  int nres;
  ResultSpec res[kMaxResults];
  void* out_values[kMaxResults];
  uint8_t* out_valid[kMaxResults];
};

__device__ __forceinline__ double decode_ordered(uint64_t u) {
  uint64_t b = (u & 0x8000000000000000ull) ? (u & 0x7fffffffffffffffull) : ~u;
  return __longlong_as_double((long long)b);
}

__device__ __forceinline__ void store_typed(void* out, int64_t i, int t, int64_t iv, double dv, bool isf) {
  switch (t) {
    case GZ_TYPE_DOUBLE: ((double*)out)[i] = isf ? dv : (double)iv; break;
    case GZ_TYPE_FLOAT: ((float*)out)[i] = (float)(isf ? dv : (double)iv); break;
    case GZ_TYPE_INT8: case GZ_TYPE_UINT8: case GZ_TYPE_BOOL: ((int8_t*)out)[i] = (int8_t)iv; break;
    case GZ_TYPE_INT16: case GZ_TYPE_UINT16: ((int16_t*)out)[i] = (int16_t)iv; break;
    case GZ_TYPE_INT32: case GZ_TYPE_UINT32: case GZ_TYPE_DATE32: ((int32_t*)out)[i] = (int32_t)iv; break;
    default: ((int64_t*)out)[i] = iv; break;
  }
}

// One output group per thread, in first-seen order, so its slot is a random
// one: the group's key and all its lanes are loaded at once (the lanes into the
// thread's own column of LDS, s_lv[l * blockDim + thread]) before any result is
// computed, so each of its lines is fetched once, not once per result column
// after other groups' traffic has evicted it.
__global__ void agg_finish_kernel(FinishParams f) {
  extern __shared__ int64_t s_lv[];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int inst = (int)(f.skeys[i] >> f.rbits);
    int64_t slot = (int64_t)f.svals[i];
    const AggInst I = f.inst[inst];
    const int64_t* ls = tab_lanes(I, slot);
    const int64_t kv = I.base[slot];
    {
      int64_t v[kMaxLanes];
#pragma unroll
      for (int l = 0; l < kMaxLanes; l++) v[l] = l < f.nl ? ls[l] : 0;
#pragma unroll
      for (int l = 0; l < kMaxLanes; l++)
        if (l < f.nl) s_lv[l * blockDim.x + threadIdx.x] = v[l];
    }
    auto lane = [&](int l) -> int64_t { return s_lv[l * blockDim.x + threadIdx.x]; };
    for (int c = 0; c < f.nres; c++) {
      const ResultSpec& rs = f.res[c];
      bool ok = true;
      int64_t iv = 0;
      double dv = 0;
      bool isf = false;
      bool rawset = false;  // raw: the value's bits as stored
      int64_t raw = 0;
      switch (rs.kind) {
        case GZ_ACTION_GROUPBY:
          ok = slot != (int64_t)I.cap;  // the null-key group
          iv = kv;
          if (rs.rtype == GZ_TYPE_DOUBLE || rs.rtype == GZ_TYPE_FLOAT) {
            // the group's first row's bits: the table holds -0.0 as 0.0 and
            // one NaN for all
            rawset = true;
            if (iv == 0) {
              const bool neg = (lane(rs.lane_c) & 1) != 0;
              raw = rs.rtype == GZ_TYPE_DOUBLE ? (neg ? INT64_MIN : 0) : (neg ? 0x80000000ll : 0);
            } else if (iv == kCanonNaN) {
              raw = lane(rs.lane_e);
            } else {
              raw = rs.rtype == GZ_TYPE_DOUBLE ? iv : (int64_t)(uint32_t)__float_as_uint((float)__longlong_as_double(iv));
            }
          }
          break;
        case GZ_ACTION_COUNT:
        case GZ_ACTION_COUNT_LITERAL: iv = lane(rs.lane_a); break;
        case GZ_ACTION_SUM:
          ok = lane(rs.lane_b) != 0;
          if (rs.rtype == GZ_TYPE_DOUBLE) { dv = __longlong_as_double(lane(rs.lane_a)); isf = true; }
          else iv = lane(rs.lane_a);
          break;
        case GZ_ACTION_AVG: {
          int64_t n = lane(rs.lane_b);
          ok = rs.global ? true : n != 0;
          dv = __longlong_as_double(lane(rs.lane_a)) / (double)n;
          isf = true;
          break;
        }
        case GZ_ACTION_SUM_COUNT:
          ok = lane(rs.lane_b) != 0;
          if (rs.sub == 0) { dv = __longlong_as_double(lane(rs.lane_a)); isf = true; }
          else iv = lane(rs.lane_b);
          break;
        case GZ_ACTION_AVG_BY_COUNT: {
          int64_t n = lane(rs.lane_b);
          ok = n != 0;
          dv = ok ? __longlong_as_double(lane(rs.lane_a)) / (double)n : 0.0;
          isf = true;
          break;
        }
        case GZ_ACTION_MIN:
        case GZ_ACTION_MAX: {
          ok = lane(rs.lane_b) != 0;
          if (rs.rtype == GZ_TYPE_DOUBLE || rs.rtype == GZ_TYPE_FLOAT) {
            const uint64_t u = (uint64_t)lane(rs.lane_a);
            const bool has_value = rs.kind == GZ_ACTION_MIN ? u != ~0ull : u != 0ull;
            bool nan;
            if (rs.fmode == 0) nan = !has_value;
            else if (rs.fmode == 1) nan = lane(rs.lane_d) != -1;
            else nan = lane(rs.lane_f) == lane(rs.lane_d);
            rawset = true;
            if (nan) {
              raw = lane(rs.lane_e);
            } else {
              double d = has_value ? decode_ordered(u) : 0.0;
              if (d == 0.0) d = (lane(rs.lane_c) & 1) ? -0.0 : 0.0;  // the first zero's sign
              raw = rs.rtype == GZ_TYPE_DOUBLE ? __double_as_longlong(d)
                                               : (int64_t)(uint32_t)__float_as_uint((float)d);
            }
          } else {
            iv = lane(rs.lane_a);
          }
          break;
        }
      }
      f.out_valid[c][i] = ok ? 1 : 0;
      if (rawset) {
        if (rs.rtype == GZ_TYPE_DOUBLE) ((int64_t*)f.out_values[c])[i] = raw;
        else ((uint32_t*)f.out_values[c])[i] = (uint32_t)raw;
      } else {
        store_typed(f.out_values[c], i, rs.rtype, iv, dv, isf);
      }
    }
  }
}

}  // namespace gz

// =============================================================== host side
using namespace gz;

struct AggPlan {
  int ncols = 0;
  std::vector<int> types;
  int key_col = -1, key_type = GZ_TYPE_INT64;
  std::vector<gz_agg_action> actions;
  std::vector<Lane> lanes;
  std::vector<ResultSpec> results;
};

struct JoinAggState {
  gz_relation* rel = nullptr;
  std::vector<int> ptypes;  // probe batch column types
  int pkey = 0;
  std::vector<int> sides, idx;  // aggregate input column -> join result column
  int64_t probe_rows = 0;       // global probe row index of the next call's first row
};

struct gz_agg {
  MultiKey* mk = nullptr;        // grouped by several key columns
  JoinAggState* join = nullptr;  // a fused join -> aggregate
  gz_context* ctx = nullptr;
  std::vector<int32_t> in_types;  // input column types
  BatchStage stage;               // per-batch calls staged for the batched path (gz_agg_process)
  AggPlan plan;
  gz_agg_options opts;
  int ninst = 1;
  std::vector<AggInst> inst;       // host copies of the device descriptors
  std::vector<DBuf> tables, counters;
  std::vector<int64_t> rows_seen;  // per instance
  std::vector<int64_t> g_bound;    // upper bound of groups per instance
  std::vector<int> placement;      // 1 = LDS first level, 2 = HBM only
  DBuf d_inst, d_lane_ops, d_tiles, d_bufs, d_regs, d_est, d_blob, d_cnt, d_cpre;
  // finish
  bool finished = false;
  std::vector<int64_t> ngroups, group_off;
  DBuf skeys, svals, tkeys, tvals, cursor, ccpre, counts, sort_tmp;
  std::vector<DBuf> res_values, res_valid;
  std::vector<DBuf> res_bits;  // Arrow validity bitmaps per (instance, column), packed on request
  std::vector<uint8_t> bits_ready;
  int64_t total_groups = 0;
  bool inst_dirty = true;
  // planned form (gz_agg_set_prepare / run / capture / launch)
  struct Launch {
    bool ready = false;
    AggParams p_hbm, p_lds;
    unsigned grid_hbm = 0, blocks_hbm = 0, blocks_lds = 0;
    int lds = 0, lds_hbm = 0;
    int64_t total_slots = 0;
    DBuf d_hbm, d_lds, d_slot_prefix, d_bufs, d_fix, d_gb;
    AggParams p_fix;
    int64_t T_fix = 0;
    int n_gb = 0;
    std::vector<int64_t> rows;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } L;
  // value-keeping float min / max and float keys: the NaN flag of the fix-up
  // pass; integer min / max without a group-by key: the all-null batch rule
  DBuf d_fix, d_fix_tiles, d_gbatch;
  bool need_fix = false, need_allnull = false, has_glast = false;
  int lds_threshold = 256;  // groups up to which an instance gets the LDS first level
  int part_groups = GZ_PART_GROUPS;  // groups per partition on the partitioned placement (0 = off)
  // partitioned placement buffers of the plan
  DBuf d_part, pinst, recs, pcnt, pprefix, pscan_tmp, pitems, pnitems;
  int64_t ncnt = 0;         // [partition][tile] counts over the partitioned instances
  size_t pscan_bytes = 0;   // temporary storage of their scan
  int64_t part_small_tiles = 0;                 // tiles of the instances with fan-out <= kSmallFanout (first)
  int part_fcap_small = 0, part_fcap_large = 0;  // largest fan-out of either group
  int B = 0;                // partitions over all partitioned instances
  int64_t nrec = 0;
  AggParams p_part;
  unsigned grid_part = 0, blocks_part_upd = 0;
  int upd_block = 512;      // workgroup size of the partitioned update (512 or 256)
  int lds_part = 0;
};

static int add_lane(AggPlan& P, int op, int col, int col2) {
  for (size_t i = 0; i < P.lanes.size(); i++)
    if (P.lanes[i].op == op && P.lanes[i].col == col && P.lanes[i].col2 == col2) return (int)i;
  if ((int)P.lanes.size() >= kMaxLanes) return -1;
  P.lanes.push_back({op, col, col2});
  return (int)P.lanes.size() - 1;
}

static bool is_float_t(int t) { return t == GZ_TYPE_FLOAT || t == GZ_TYPE_DOUBLE; }
static bool is_int_t(int t) {
  return t == GZ_TYPE_BOOL || t == GZ_TYPE_INT8 || t == GZ_TYPE_UINT8 || t == GZ_TYPE_INT16 ||
         t == GZ_TYPE_UINT16 || t == GZ_TYPE_INT32 || t == GZ_TYPE_UINT32 || t == GZ_TYPE_INT64 ||
         t == GZ_TYPE_UINT64 || t == GZ_TYPE_DATE32 || t == GZ_TYPE_DATE64 || t == GZ_TYPE_TIMESTAMP;
}

static int build_plan(AggPlan& P, int ncols, const int32_t* types, int nact, const gz_agg_action* acts) {
  P.ncols = ncols;
  P.types.assign(types, types + ncols);
  if (nact <= 0 || nact > kMaxActions) return fail(GZ_INVALID, "bad action count");
  if (ncols > kAggMaxCols) return fail(GZ_NOT_IMPLEMENTED, "more than 32 input columns");
  P.actions.assign(acts, acts + nact);
  bool global = true;
  for (auto& a : P.actions)
    if (a.kind == GZ_ACTION_GROUPBY) global = false;
  for (auto& a : P.actions) {
    auto col_ok = [&](int c) { return c >= 0 && c < ncols; };
    ResultSpec r{};
    r.kind = a.kind;
    r.nan_check = a.nan_check;
    r.global = global;
    r.lane_a = r.lane_b = r.lane_c = r.lane_d = r.lane_e = r.lane_f = 0;
    r.fmode = 0;
    int t = col_ok(a.input0) ? types[a.input0] : -1;
    switch (a.kind) {
      case GZ_ACTION_GROUPBY:
        if (P.key_col >= 0) return fail(GZ_NOT_IMPLEMENTED, "multiple group-by keys");
        if (!col_ok(a.input0)) return fail(GZ_INVALID, "groupby input out of range");
        if (!is_int_t(t) && !is_float_t(t))
          return fail(GZ_NOT_IMPLEMENTED, "group-by key type " + std::to_string(t) +
                                              " (the HIP aggregate groups integer and floating keys)");
        P.key_col = a.input0;
        P.key_type = t;
        r.rtype = t;
        if (is_float_t(t)) {
          r.lane_c = add_lane(P, L_FIRSTZ, a.input0, -1);
          r.lane_e = add_lane(P, L_NANBITS, a.input0, -1);
        }
        P.results.push_back(r);
        break;
      case GZ_ACTION_COUNT:
        if (!col_ok(a.input0)) return fail(GZ_INVALID, "count input out of range");
        r.lane_a = col_ok(a.input1) ? add_lane(P, L_NN2, a.input0, a.input1) : add_lane(P, L_NN, a.input0, -1);
        r.rtype = GZ_TYPE_INT64;
        P.results.push_back(r);
        break;
      case GZ_ACTION_COUNT_LITERAL:
        r.lane_a = add_lane(P, L_CNT, -1, -1);
        r.rtype = GZ_TYPE_INT64;
        P.results.push_back(r);
        break;
      case GZ_ACTION_SUM:
        if (!col_ok(a.input0) || !(is_int_t(t) || is_float_t(t))) return fail(GZ_INVALID, "sum input");
        r.lane_a = is_float_t(t) ? add_lane(P, L_SUMF, a.input0, -1) : add_lane(P, L_SUMI, a.input0, -1);
        r.lane_b = add_lane(P, L_NN, a.input0, -1);
        r.rtype = is_float_t(t) ? GZ_TYPE_DOUBLE : GZ_TYPE_INT64;
        P.results.push_back(r);
        break;
      case GZ_ACTION_AVG:
      case GZ_ACTION_SUM_COUNT:
        if (!col_ok(a.input0) || !(is_int_t(t) || is_float_t(t))) return fail(GZ_INVALID, "avg input");
        r.lane_a = add_lane(P, L_SUMF, a.input0, -1);
        r.lane_b = add_lane(P, L_NN, a.input0, -1);
        r.rtype = GZ_TYPE_DOUBLE;
        P.results.push_back(r);
        if (a.kind == GZ_ACTION_SUM_COUNT) {
          r.sub = 1;
          r.rtype = GZ_TYPE_INT64;
          P.results.push_back(r);
        }
        break;
      case GZ_ACTION_AVG_BY_COUNT:
        if (!col_ok(a.input0) || !col_ok(a.input1)) return fail(GZ_INVALID, "avgByCount needs (sum, count) inputs");
        r.lane_a = add_lane(P, L_SUMF, a.input0, -1);
        if (global) {
          // AvgByCountAction::Evaluate (actions_impl.cc:3818-3838): arrow Sum of
          // each input, each skipping its own nulls; any integer count type
          if (!is_int_t(types[a.input1])) return fail(GZ_INVALID, "avgByCount needs an integer count input");
          r.lane_b = add_lane(P, L_SUMI, a.input1, -1);
        } else {
          // the grouped lambda (:3785-3795) adds the count of rows whose sum
          // input is non-null, read as int64 (undefined for narrower columns)
          if (types[a.input1] != GZ_TYPE_INT64) return fail(GZ_INVALID, "avgByCount needs (sum, int64 count) inputs");
          r.lane_b = add_lane(P, L_SUMCNT, a.input0, a.input1);
        }
        r.rtype = GZ_TYPE_DOUBLE;
        P.results.push_back(r);
        break;
      case GZ_ACTION_MIN:
      case GZ_ACTION_MAX: {
        bool mn = a.kind == GZ_ACTION_MIN;
        if (!col_ok(a.input0)) return fail(GZ_INVALID, "min/max input out of range");
        r.rtype = t;
        r.lane_b = add_lane(P, L_NN, a.input0, -1);
        if (is_float_t(t)) {
          // which NaN a result keeps (see ResultSpec::fmode)
          r.lane_a = add_lane(P, mn ? L_MINF : L_MAXF, a.input0, -1);
          r.lane_c = add_lane(P, L_FIRSTZ, a.input0, -1);
          if (mn && (a.nan_check || global)) {
            r.fmode = 0;
            r.lane_d = add_lane(P, L_FIRSTNAN, a.input0, -1);
          } else if (!mn && a.nan_check) {
            r.fmode = 1;
            r.lane_d = add_lane(P, global ? L_GLASTNAN : L_LASTNAN, a.input0, -1);
          } else {
            r.fmode = 2;
            r.lane_d = add_lane(P, L_FIRSTNAN, a.input0, -1);
            r.lane_f = add_lane(P, L_FIRSTNN, a.input0, -1);
          }
          r.lane_e = r.lane_d < 0 ? -1 : add_lane(P, L_NANBITS, a.input0, r.lane_d);
        } else if (t == GZ_TYPE_UINT64) {
          r.lane_a = add_lane(P, mn ? L_MINU : L_MAXU, a.input0, -1);
          if (global) r.lane_f = add_lane(P, L_FIRSTNN, a.input0, -1);  // for the all-null batch rule
        } else if (is_int_t(t)) {
          r.lane_a = add_lane(P, mn ? L_MINI : L_MAXI, a.input0, -1);
          if (global) r.lane_f = add_lane(P, L_FIRSTNN, a.input0, -1);
        } else {
          return fail(GZ_NOT_IMPLEMENTED, "min/max input type");
        }
        P.results.push_back(r);
        break;
      }
      default:
        return fail(GZ_NOT_IMPLEMENTED, "action kind " + std::to_string(a.kind) + " is not implemented");
    }
    if (r.lane_a < 0 || r.lane_b < 0 || r.lane_c < 0 || r.lane_d < 0 || r.lane_e < 0 || r.lane_f < 0)
      return fail(GZ_NOT_IMPLEMENTED, "too many aggregate states");
  }
  if ((int)P.results.size() > kMaxResults) return fail(GZ_NOT_IMPLEMENTED, "too many result columns");
  return GZ_OK;
}

static uint64_t pow2_at_least(uint64_t x) {
  uint64_t c = 16;
  while (c < x) c <<= 1;
  return c;
}

static AggParams base_params(gz_agg* h) {
  AggParams p;
  memset(&p, 0, sizeof(p));
  p.key_col = h->plan.key_col;
  p.key_type = h->plan.key_type;
  p.nl = (int)h->plan.lanes.size();
  for (int l = 0; l < p.nl; l++) p.lanes[l] = h->plan.lanes[l];
  for (int c = 0; c < h->plan.ncols; c++) p.col_type[c] = h->plan.types[c];
  p.lcap = 1024;
  p.lprobe = 32;
  p.fix_flag = h->need_fix ? (int*)h->d_fix.p : nullptr;
  // used input columns of a partition record
  for (int c = 0; c < kAggMaxCols; c++) p.uslot[c] = -1;
  p.nu = 0;
  for (int l = 0; l < p.nl; l++) {
    if (p.lanes[l].op == L_NANBITS) continue;  // read by the fix-up pass from the input; col2 names a lane
    for (int c : {p.lanes[l].col, p.lanes[l].col2}) {
      if (c < 0 || p.uslot[c] >= 0) continue;
      if (p.nu == kMaxUsed) {
        p.nu = -1;  // too many for a record: no partitioned placement
        return p;
      }
      p.uslot[c] = p.nu;
      p.ucol[p.nu++] = c;
    }
  }
  return p;
}

// LDS table: lcap slots, the null-key and kEmpty-key slots, and a spare slot.
static int lds_bytes(const gz_agg* h, int lcap) {
  return (int)((lcap + 3) * (2 + h->plan.lanes.size()) * 8);
}

// HBM-placed instances (more groups than an LDS table holds) still get a small
// LDS first level when the launch has tiles enough (eight per CU) to pay for
// the tables' reset and merge: persistent workgroups (four per CU) with a
// 512-key table and a 4-slot probe, so the head of a skewed key distribution
// aggregates in LDS and only its tail goes to the HBM table; with uniform keys
// the table fills at once and costs a few LDS reads per row.
static void hbm_first_level(const gz_agg* h, AggParams* q, unsigned* blocks_out, int* lds_out) {
  int64_t blocks = std::min<int64_t>(q->T, (int64_t)h->ctx->num_cus * 4);
  q->tiles_per_block = (int)((q->T + blocks - 1) / blocks);
  *blocks_out = (unsigned)((q->T + q->tiles_per_block - 1) / q->tiles_per_block);
  q->lcap = 512;
  q->lprobe = 4;
  int lds = lds_bytes(h, q->lcap);
  while (lds > 40 * 1024 && q->lcap > 64) {
    q->lcap >>= 1;
    lds = lds_bytes(h, q->lcap);
  }
  *lds_out = lds;
}

// (Re)allocate an instance's table with capacity cap (copying existing groups).
static int inst_alloc(gz_agg* h, int i, uint64_t cap) {
  gz_context* ctx = h->ctx;
  const int nl = (int)h->plan.lanes.size();
  DBuf nb;
  GZ_TRY(ctx_alloc(ctx, sizeof(int64_t) * (cap + 2) * (2 + nl), &nb));
  DBuf nc;
  GZ_TRY(ctx_alloc(ctx, 16, &nc));
  AggInst ni{};
  ni.base = (int64_t*)nb.p;
  ni.cap = cap;
  ni.counter = (unsigned int*)nc.p;
  ni.overflow = (int*)((char*)nc.p + 8);
  ni.nl = nl;
  // initialize this one instance
  DBuf one;
  GZ_TRY(upload(ctx, &one, &ni, sizeof(AggInst)));
  std::vector<int> ops(std::max(nl, 1));
  for (int l = 0; l < nl; l++) ops[l] = h->plan.lanes[l].op;
  GZ_TRY(upload(ctx, &h->d_lane_ops, ops.data(), sizeof(int) * ops.size()));
  unsigned gx = (unsigned)std::min<uint64_t>((cap + 2 + 255) / 256, 4096);
  hipLaunchKernelGGL(agg_init_kernel, dim3(gx, 1), dim3(256), 0, ctx->stream, (const AggInst*)one.p, 1, nl,
                     (const int*)h->d_lane_ops.p);
  if (h->tables[i].p) {
    AggInst old = h->inst[i];
    unsigned g2 = (unsigned)std::min<uint64_t>((old.cap + 2 + 255) / 256, 4096);
    hipLaunchKernelGGL(agg_rehash_kernel, dim3(g2), dim3(256), 0, ctx->stream, old, ni, nl);
    // the inserted-key counter of the new table counts the moved keys; the old
    // buffers go back to the pool in stream order
    ctx_free(ctx, &h->tables[i]);
    ctx_free(ctx, &h->counters[i]);
  }
  GZ_HIP(hipGetLastError());
  ctx_free(ctx, &one);
  h->tables[i] = nb;
  h->counters[i] = nc;
  h->inst[i] = ni;
  h->inst_dirty = true;
  return GZ_OK;
}

// Groups bound (keys inserted + the two special slots) and overflow of every
// instance: one gather kernel, one copy, one synchronization.
static int read_counts(gz_agg* h, std::vector<int64_t>& groups, int* overflow) {
  gz_context* ctx = h->ctx;
  *overflow = 0;
  groups.assign(h->ninst, 0);
  if (h->inst_dirty) {
    GZ_TRY(upload(ctx, &h->d_inst, h->inst.data(), sizeof(AggInst) * h->inst.size()));
    h->inst_dirty = false;
  }
  GZ_TRY(ctx_ensure(ctx, &h->d_cnt, sizeof(unsigned) * 2 * h->ninst));
  hipLaunchKernelGGL(agg_read_counters_kernel, dim3((unsigned)((h->ninst + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const AggInst*)h->d_inst.p, h->ninst, (unsigned*)h->d_cnt.p);
  std::vector<unsigned int> buf(2 * h->ninst);
  GZ_HIP(hipMemcpyAsync(buf.data(), h->d_cnt.p, sizeof(unsigned) * 2 * h->ninst, hipMemcpyDeviceToHost, ctx->stream));
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < h->ninst; i++) {
    groups[i] = buf[2 * i] + 2;
    if (buf[2 * i + 1]) *overflow = 1;
  }
  return GZ_OK;
}

static int agg_set_make(gz_context* ctx, int32_t ncols, const int32_t* types, int32_t nact,
                        const gz_agg_action* acts, const gz_agg_options* opts, int32_t ninst,
                        gz_agg** out) {
  if (!ctx || !out || (ncols > 0 && !types) || !acts) return fail(GZ_INVALID, "gz_agg_make: null argument");
  if (ninst <= 0 || ninst > 65535) return fail(GZ_INVALID, "instance count must be in [1, 65535]");
  int nkeys = 0;
  bool str_key = false;
  for (int i = 0; i < nact; i++) {
    if (acts[i].kind != GZ_ACTION_GROUPBY) continue;
    nkeys++;
    if (acts[i].input0 >= 0 && acts[i].input0 < ncols && types[acts[i].input0] == GZ_TYPE_STRING) str_key = true;
  }
  if (nkeys > 1 || str_key) {
    if (ninst > 1) return fail(GZ_NOT_IMPLEMENTED, "multi-column or utf8 group-by keys in an aggregate set");
    if (nact > kMaxActions) return fail(GZ_INVALID, "bad action count");
    auto* h = new gz_agg();
    h->ctx = ctx;
    h->in_types.assign(types, types + ncols);
    int rc = mk_make(ctx, ncols, types, nact, acts, opts, &h->mk);
    if (rc) {
      delete h;
      return rc;
    }
    *out = h;
    return GZ_OK;
  }
  auto* h = new gz_agg();
  h->ctx = ctx;
  h->in_types.assign(types, types + ncols);
  int rc = build_plan(h->plan, ncols, types, nact, acts);
  if (rc) {
    delete h;
    return rc;
  }
  for (auto& l : h->plan.lanes) {
    if (l.op == L_NANBITS) h->need_fix = true;
    if (l.op == L_GLASTNAN) h->has_glast = true;
    if (h->plan.key_col < 0 && (l.op == L_MINI || l.op == L_MAXI || l.op == L_MINU)) h->need_allnull = true;
  }
  if (h->need_fix) {
    rc = ctx_alloc(ctx, 256, &h->d_fix);
    if (rc) {
      delete h;
      return rc;
    }
    hipLaunchKernelGGL(fill_u32_kernel_agg, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)h->d_fix.p, (int64_t)1);
  }
  gz_agg_options_defaults(&h->opts);
  if (opts) h->opts = *opts;
  if (h->opts.load_factor <= 0 || h->opts.load_factor > 0.9) h->opts.load_factor = 0.5;
  h->ninst = ninst;
  h->inst.resize(ninst);
  h->tables.resize(ninst);
  h->counters.resize(ninst);
  h->rows_seen.assign(ninst, 0);
  h->g_bound.assign(ninst, 2);
  h->placement.assign(ninst, h->opts.placement >= 1 && h->opts.placement <= 3 ? h->opts.placement : 2);
  *out = h;
  return GZ_OK;
}

// Tile table + buffer table of a call.
static int agg_stage(gz_agg* h, int32_t nb, const gz_batch* batches, const int32_t* inst_of,
                     std::vector<AggTile>& tiles, std::vector<int64_t>& rows_per_inst, int tile_rows,
                     DBuf* bufs_dst = nullptr, std::vector<GBatch>* gbatches = nullptr, UploadBatch* ub = nullptr) {
  const int nc = h->plan.ncols;
  std::vector<const void*> bufs((size_t)nb * nc * 3, nullptr);
  rows_per_inst.assign(h->ninst, 0);
  std::vector<int64_t> base = h->rows_seen;
  for (int bi = 0; bi < nb; bi++) {
    const gz_batch& bt = batches[bi];
    int ii = inst_of ? inst_of[bi] : 0;
    if (ii < 0 || ii >= h->ninst) return fail(GZ_INDEX_ERROR, "instance out of range");
    if (bt.num_rows < 0 || bt.num_rows > INT32_MAX) return fail(GZ_INVALID, "bad num_rows");
    if (h->has_glast && (bt.num_rows >= (1 << 24) || base[ii] + bt.num_rows >= (1ll << 39)))
      return fail(GZ_NOT_IMPLEMENTED, "max with NaN check and no group-by key: batches of 2^24 rows or more");
    int need = 0;
    for (int c = 0; c < nc; c++) need += type_num_bufs(h->plan.types[c]);
    if (bt.num_bufs < need) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
    int k = 0;
    for (int c = 0; c < nc; c++) {
      int t = h->plan.types[c];
      const void** slot = &bufs[((size_t)bi * nc + c) * 3];
      slot[0] = (bt.buf_addrs[k] != 0 && bt.buf_sizes[k] > 0) ? (const void*)bt.buf_addrs[k] : nullptr;
      if (slot[0] && bt.buf_sizes[k] < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
      k++;
      if (type_is_binary(t)) slot[1] = (const void*)bt.buf_addrs[k++];
      slot[2] = (const void*)bt.buf_addrs[k];
      if (t != GZ_TYPE_NA && bt.num_rows > 0 && !slot[2]) return fail(GZ_INVALID, "missing values buffer");
      int64_t need_b = t == GZ_TYPE_BOOL ? (bt.num_rows + 7) / 8 : bt.num_rows * type_width(t);
      if (t != GZ_TYPE_NA && !type_is_binary(t) && bt.buf_sizes[k] < need_b)
        return fail(GZ_INVALID, "values buffer too small");
      k++;
    }
    for (int64_t r0 = 0; r0 < bt.num_rows; r0 += tile_rows) {
      AggTile td;
      td.batch = bi;
      td.row_begin = (int32_t)r0;
      td.nrows = (int32_t)std::min<int64_t>(tile_rows, bt.num_rows - r0);
      td.inst = ii;
      td.row_base = base[ii] + r0;
      tiles.push_back(td);
    }
    if (gbatches && bt.num_rows > 0) gbatches->push_back(GBatch{bi, ii, bt.num_rows, base[ii]});
    base[ii] += bt.num_rows;
    rows_per_inst[ii] += bt.num_rows;
  }
  DBuf* dst = bufs_dst ? bufs_dst : &h->d_bufs;
  const size_t nbytes = sizeof(void*) * std::max<size_t>(bufs.size(), 1);
  if (ub) return ub->add(h->ctx, dst, bufs.data(), nbytes);
  return upload(h->ctx, dst, bufs.data(), nbytes);
}

static int agg_prepare(gz_agg* h, int32_t nb, const gz_batch* batches, const int32_t* inst_of);
static int agg_run(gz_agg* h);

static int agg_process(gz_agg* h, int32_t nb, const gz_batch* batches, const int32_t* inst_of) {
  gz_context* ctx = h->ctx;
  if (h->finished) return fail(GZ_INVALID, "aggregate already finished");
  if (h->join) return fail(GZ_INVALID, "a join aggregate takes probe batches (gz_join_agg_process_batches)");
  GZ_HIP(hipSetDevice(ctx->device));
  // the first call of a fresh aggregate goes through the planned path (all
  // placements, the partitioned one included); later calls add to its tables
  bool fresh = !h->L.ready;
  for (int i = 0; i < h->ninst && fresh; i++)
    if (h->tables[i].p) fresh = false;
  if (fresh && nb > 0) {
    GZ_TRY(agg_prepare(h, nb, batches, inst_of));
    return agg_run(h);
  }
  // tables a plan left partition-structured become one hash table first (the
  // per-call kernels insert anywhere in it)
  for (int i = 0; i < h->ninst; i++)
    if (h->tables[i].p && h->inst[i].sub_log2) GZ_TRY(inst_alloc(h, i, h->inst[i].cap));
  // small calls (one batch at a time) get small tiles so a call still spreads
  // over many workgroups
  int64_t call_rows = 0;
  for (int32_t b = 0; b < nb; b++) call_rows += std::max<int64_t>(0, batches[b].num_rows);
  const int tile_rows = h->opts.tile_rows > 0 ? h->opts.tile_rows : (call_rows < (1 << 20) ? 512 : 4096);
  std::vector<AggTile> tiles;
  std::vector<int64_t> rows;
  std::vector<GBatch> gb;
  GZ_TRY(agg_stage(h, nb, batches, inst_of, tiles, rows, tile_rows, nullptr, &gb));
  if (tiles.empty()) return GZ_OK;
  GZ_TRY(upload(ctx, &h->d_tiles, tiles.data(), sizeof(AggTile) * tiles.size()));
  AggParams p = base_params(h);
  p.tiles = (const AggTile*)h->d_tiles.p;
  p.T = (int64_t)tiles.size();
  p.bufs = (const void* const*)h->d_bufs.p;
  p.bstride = h->plan.ncols * 3;
  const double lf = h->opts.load_factor;
  // ---- size tables: first use of an instance in a multi-row call estimates its
  // cardinality with HyperLogLog; later calls grow from an upper bound.
  bool need_est = false;
  for (int i = 0; i < h->ninst; i++)
    if (rows[i] > 0 && !h->tables[i].p && rows[i] > 4096 && h->opts.expected_groups <= 0) need_est = true;
  std::vector<double> est(h->ninst, 0.0);
  if (need_est) {
    GZ_TRY(ctx_ensure(ctx, &h->d_regs, sizeof(unsigned) * kHllRegs * h->ninst));
    GZ_TRY(ctx_ensure(ctx, &h->d_est, sizeof(double) * h->ninst));
    int64_t n32 = (int64_t)kHllRegs * h->ninst;
    hipLaunchKernelGGL(fill_u32_kernel_agg, dim3((unsigned)std::min<int64_t>((n32 + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, (uint32_t*)h->d_regs.p, n32);
    int64_t blocks = std::min<int64_t>(p.T, (int64_t)ctx->num_cus * 4);
    p.tiles_per_block = (int)((p.T + blocks - 1) / blocks);
    blocks = (p.T + p.tiles_per_block - 1) / p.tiles_per_block;
    {
      KernelTimer kt(ctx, "agg_hll");
      hipLaunchKernelGGL(agg_hll_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p,
                         (unsigned int*)h->d_regs.p);
      hipLaunchKernelGGL(agg_hll_estimate_kernel, dim3((unsigned)h->ninst), dim3(256), 0, ctx->stream,
                         (const unsigned int*)h->d_regs.p, (double*)h->d_est.p);
    }
    GZ_HIP(hipMemcpyAsync(est.data(), h->d_est.p, sizeof(double) * h->ninst, hipMemcpyDeviceToHost, ctx->stream));
    GZ_HIP(hipStreamSynchronize(ctx->stream));
  }
  bool sync_needed = false;
  for (int i = 0; i < h->ninst; i++)
    if (rows[i] > 0 && h->tables[i].p && (double)(h->g_bound[i] + rows[i]) > lf * h->inst[i].cap)
      sync_needed = true;
  std::vector<int64_t> actual;
  if (sync_needed) {
    int ovf = 0;
    GZ_TRY(read_counts(h, actual, &ovf));
    if (ovf) return fail(GZ_CAPACITY_ERROR, "aggregate hash table overflowed");
    for (int i = 0; i < h->ninst; i++)
      if (h->tables[i].p) h->g_bound[i] = actual[i];
  }
  bool est_used = false;
  for (int i = 0; i < h->ninst; i++) {
    if (rows[i] == 0) continue;
    if (!h->tables[i].p) {
      double g;
      bool estimated = false;
      if (h->opts.expected_groups > 0) {
        g = std::min<double>((double)h->opts.expected_groups, (double)rows[i]);
        estimated = true;
      } else if (need_est && rows[i] > 4096) {
        g = std::min<double>(est[i] * 1.3 + 64, (double)rows[i]);
        estimated = true;
      } else {
        g = (double)rows[i];
      }
      est_used = est_used || estimated;
      uint64_t cap = pow2_at_least((uint64_t)(g / lf) + 16);
      if (h->plan.key_col < 0) cap = 16;
      GZ_TRY(inst_alloc(h, i, cap));
      h->g_bound[i] = 2 + (int64_t)std::min<double>(g, (double)rows[i]);
      // placement: LDS first-level table when the groups fit in LDS
      if (h->opts.placement == 0)
        h->placement[i] = (g <= (double)h->lds_threshold || h->plan.key_col < 0) ? 1 : 2;
      if (!estimated) h->g_bound[i] = 2 + rows[i];
    } else if ((double)(h->g_bound[i] + rows[i]) > lf * h->inst[i].cap) {
      uint64_t cap = pow2_at_least((uint64_t)((double)(h->g_bound[i] + rows[i]) * 2 / lf));
      GZ_TRY(inst_alloc(h, i, cap));
      h->g_bound[i] += rows[i];
    } else {
      h->g_bound[i] += rows[i];
    }
  }
  if (h->inst_dirty) {
    GZ_TRY(upload(ctx, &h->d_inst, h->inst.data(), sizeof(AggInst) * h->inst.size()));
    h->inst_dirty = false;
  }
  p.inst = (const AggInst*)h->d_inst.p;
  // ---- update: split tiles by placement
  std::vector<AggTile> lds_tiles, hbm_tiles;
  for (auto& t : tiles) (h->placement[t.inst] == 1 ? lds_tiles : hbm_tiles).push_back(t);
  if (!hbm_tiles.empty()) {
    if (!lds_tiles.empty())
      GZ_TRY(upload(ctx, &h->d_tiles, hbm_tiles.data(), sizeof(AggTile) * hbm_tiles.size()));
    p.tiles = (const AggTile*)h->d_tiles.p;
    p.T = (int64_t)hbm_tiles.size();
    KernelTimer kt(ctx, "agg_update");
    if (p.T >= 8ll * ctx->num_cus) {
      AggParams ph = p;
      unsigned blocks = 0;
      int lds = 0;
      hbm_first_level(h, &ph, &blocks, &lds);
      hipLaunchKernelGGL(agg_update_lds_kernel, dim3(blocks), dim3(256), lds, ctx->stream, ph);
    } else {  // a small call: too few tiles per workgroup to pay for an LDS table's reset and merge
      const int64_t grid = (p.T + 7) / 8 * 8;
      hipLaunchKernelGGL(agg_update_kernel, dim3((unsigned)grid), dim3(256), 0, ctx->stream, p);
    }
  }
  if (!lds_tiles.empty()) {
    DBuf* dt = hbm_tiles.empty() ? &h->d_tiles : nullptr;
    DBuf second;
    if (!dt) {
      GZ_TRY(ctx_alloc(ctx, sizeof(AggTile) * lds_tiles.size(), &second));
      dt = &second;
    }
    GZ_TRY(upload(ctx, dt, lds_tiles.data(), sizeof(AggTile) * lds_tiles.size()));
    p.tiles = (const AggTile*)dt->p;
    p.T = (int64_t)lds_tiles.size();
    int64_t blocks = std::min<int64_t>(p.T, (int64_t)ctx->num_cus * 2);
    p.tiles_per_block = (int)((p.T + blocks - 1) / blocks);
    blocks = (p.T + p.tiles_per_block - 1) / p.tiles_per_block;
    p.lcap = 1024;
    int lds = lds_bytes(h, p.lcap);
    while (lds > 96 * 1024 && p.lcap > 64) {
      p.lcap >>= 1;
      lds = lds_bytes(h, p.lcap);
    }
    KernelTimer kt(ctx, "agg_update_lds");
    hipLaunchKernelGGL(agg_update_lds_kernel, dim3((unsigned)blocks), dim3(256), lds, ctx->stream, p);
    if (second.p) ctx_free(ctx, &second);
  }
  if (h->need_fix) {  // NaN bits of value-keeping results (returns at once without a NaN)
    GZ_TRY(upload(ctx, &h->d_fix_tiles, tiles.data(), sizeof(AggTile) * tiles.size()));
    AggParams pf = p;
    pf.tiles = (const AggTile*)h->d_fix_tiles.p;
    pf.T = (int64_t)tiles.size();
    hipLaunchKernelGGL(agg_fix_kernel, dim3((unsigned)pf.T), dim3(256), 0, ctx->stream, pf);
  }
  if (h->need_allnull && !gb.empty()) {
    GZ_TRY(upload(ctx, &h->d_gbatch, gb.data(), sizeof(GBatch) * gb.size()));
    hipLaunchKernelGGL(agg_allnull_kernel, dim3((unsigned)gb.size()), dim3(256), 0, ctx->stream, p,
                       (const GBatch*)h->d_gbatch.p, (int)gb.size());
  }
  GZ_HIP(hipGetLastError());
  for (int i = 0; i < h->ninst; i++) h->rows_seen[i] += rows[i];
  if (est_used) {
    // an estimate may be wrong: verify this call did not overflow
    std::vector<int64_t> g;
    int ovf = 0;
    GZ_TRY(read_counts(h, g, &ovf));
    if (ovf) return fail(GZ_CAPACITY_ERROR, "aggregate hash table overflowed (cardinality estimate)");
    for (int i = 0; i < h->ninst; i++)
      if (h->tables[i].p) h->g_bound[i] = g[i];
  }
  return GZ_OK;
}

static int agg_finish(gz_agg* h) {
  if (h->finished) return GZ_OK;
  std::fill(h->bits_ready.begin(), h->bits_ready.end(), 0);
  gz_context* ctx = h->ctx;
  GZ_HIP(hipSetDevice(ctx->device));
  // instances never fed get an empty table so the collect kernel can run
  for (int i = 0; i < h->ninst; i++)
    if (!h->tables[i].p) GZ_TRY(inst_alloc(h, i, 16));
  if (h->inst_dirty) {
    GZ_TRY(upload(ctx, &h->d_inst, h->inst.data(), sizeof(AggInst) * h->inst.size()));
    h->inst_dirty = false;
  }
  int64_t maxg = 0;
  std::vector<int64_t> cpre(h->ninst + 1, 0);  // chunks of kCollectChunk slots per instance
  int64_t maxrows = 1;
  for (int i = 0; i < h->ninst; i++) {
    maxg += (int64_t)h->inst[i].cap + 2;
    cpre[i + 1] = cpre[i] + (int64_t)((h->inst[i].cap + 2 + kCollectChunk - 1) / kCollectChunk);
    maxrows = std::max(maxrows, h->rows_seen[i]);
  }
  // first rows of instance i are below rows_seen[i]
  int rbits = 1;
  while ((1ll << rbits) < maxrows) rbits++;
  GZ_TRY(ctx_ensure(ctx, &h->skeys, sizeof(uint64_t) * maxg));
  GZ_TRY(ctx_ensure(ctx, &h->svals, sizeof(uint64_t) * maxg));
  GZ_TRY(ctx_ensure(ctx, &h->tkeys, sizeof(uint64_t) * maxg));
  GZ_TRY(ctx_ensure(ctx, &h->tvals, sizeof(uint64_t) * maxg));
  // compaction of the occupied slots: chunk counts, scan, ordered writes
  const int64_t NC = cpre[h->ninst];
  GZ_TRY(upload(ctx, &h->d_cpre, cpre.data(), sizeof(int64_t) * cpre.size()));
  GZ_TRY(ctx_ensure(ctx, &h->cursor, sizeof(uint64_t) * NC));  // chunk counts
  GZ_TRY(ctx_ensure(ctx, &h->ccpre, sizeof(uint64_t) * NC));   // their exclusive scan
  GZ_TRY(ctx_ensure(ctx, &h->counts, sizeof(uint64_t) * (h->ninst + 1)));
  {
    KernelTimer kt(ctx, "agg_collect");
    hipLaunchKernelGGL(agg_occupancy_kernel, dim3((unsigned)NC), dim3(256), 0, ctx->stream,
                       (const AggInst*)h->d_inst.p, (const int64_t*)h->d_cpre.p, h->ninst,
                       (unsigned long long*)h->cursor.p);
    size_t tmp = 0;
    GZ_HIP(rocprim::exclusive_scan(nullptr, tmp, (const unsigned long long*)h->cursor.p,
                                   (unsigned long long*)h->ccpre.p, 0ull, (size_t)NC,
                                   rocprim::plus<unsigned long long>(), ctx->stream));
    GZ_TRY(ctx_ensure(ctx, &h->sort_tmp, tmp));
    GZ_HIP(rocprim::exclusive_scan(h->sort_tmp.p, tmp, (const unsigned long long*)h->cursor.p,
                                   (unsigned long long*)h->ccpre.p, 0ull, (size_t)NC,
                                   rocprim::plus<unsigned long long>(), ctx->stream));
    hipLaunchKernelGGL(agg_collect_kernel, dim3((unsigned)NC), dim3(256), 0, ctx->stream,
                       (const AggInst*)h->d_inst.p, (const int64_t*)h->d_cpre.p, h->ninst,
                       (const unsigned long long*)h->ccpre.p, rbits, (unsigned long long*)h->tkeys.p,
                       (unsigned long long*)h->tvals.p);
    hipLaunchKernelGGL(fill_u32_kernel_agg, dim3(1), dim3(64), 0, ctx->stream, (uint32_t*)h->counts.p + 2 * h->ninst,
                       (int64_t)2);
    hipLaunchKernelGGL(agg_inst_counts_kernel, dim3((unsigned)((h->ninst + 255) / 256)), dim3(256), 0, ctx->stream,
                       (const AggInst*)h->d_inst.p, (const unsigned long long*)h->cursor.p,
                       (const unsigned long long*)h->ccpre.p, (const int64_t*)h->d_cpre.p, h->ninst,
                       (unsigned long long*)h->counts.p);
  }
  std::vector<unsigned long long> cnt(h->ninst + 1, 0);
  GZ_HIP(hipMemcpyAsync(cnt.data(), h->counts.p, sizeof(uint64_t) * (h->ninst + 1), hipMemcpyDeviceToHost,
                        ctx->stream));
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  // tables sized from an estimate or a hint may have overflowed: rows were lost
  if (cnt[h->ninst]) return fail(GZ_CAPACITY_ERROR, "aggregate hash table overflowed (cardinality estimate); "
                                                    "prepare again with expected_groups");
  cnt[h->ninst] = 0;
  for (int i = 0; i < h->ninst; i++) cnt[h->ninst] += cnt[i];
  const int64_t n = (int64_t)cnt[h->ninst];
  h->ngroups.assign(h->ninst, 0);
  h->group_off.assign(h->ninst + 1, 0);
  for (int i = 0; i < h->ninst; i++) {
    h->ngroups[i] = (int64_t)cnt[i];
    // the global aggregate has one output row once any row was processed
    if (h->plan.key_col < 0) h->ngroups[i] = h->rows_seen[i] > 0 ? 1 : 0;
    h->group_off[i + 1] = h->group_off[i] + (int64_t)cnt[i];
  }
  h->total_groups = n;
  // sort by (instance, first row): first-seen order inside each instance
  if (n > 0 && n <= kRankSortMax) {
    KernelTimer kt(ctx, "agg_sort");
    hipLaunchKernelGGL(agg_rank_sort_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64 * kRankWaves), 0, ctx->stream,
                       (const unsigned long long*)h->tkeys.p, (const unsigned long long*)h->tvals.p, n,
                       (unsigned long long*)h->skeys.p, (unsigned long long*)h->svals.p);
  } else if (n > 0) {
    int end_bit = rbits;
    while ((1ll << (end_bit - rbits)) < h->ninst) end_bit++;
    size_t tmp = 0;
    GZ_HIP(rocprim::radix_sort_pairs(nullptr, tmp, (unsigned long long*)h->tkeys.p,
                                     (unsigned long long*)h->skeys.p, (unsigned long long*)h->tvals.p,
                                     (unsigned long long*)h->svals.p, (size_t)n, 0, end_bit, ctx->stream));
    GZ_TRY(ctx_ensure(ctx, &h->sort_tmp, tmp));
    KernelTimer kt(ctx, "agg_sort");
    GZ_HIP(rocprim::radix_sort_pairs(h->sort_tmp.p, tmp, (unsigned long long*)h->tkeys.p,
                                     (unsigned long long*)h->skeys.p, (unsigned long long*)h->tvals.p,
                                     (unsigned long long*)h->svals.p, (size_t)n, 0, end_bit, ctx->stream));
  }
  const int nres = (int)h->plan.results.size();
  for (auto& b : h->res_values) ctx_free(ctx, &b);
  for (auto& b : h->res_valid) ctx_free(ctx, &b);
  h->res_values.assign(nres, DBuf());
  h->res_valid.assign(nres, DBuf());
  FinishParams f;
  memset(&f, 0, sizeof(f));
  f.inst = (const AggInst*)h->d_inst.p;
  f.skeys = (const unsigned long long*)h->skeys.p;
  f.svals = (const unsigned long long*)h->svals.p;
  f.n = n;
  f.key_type = h->plan.key_type;
  f.rbits = rbits;
  f.nl = (int)h->plan.lanes.size();
  f.nres = nres;
  for (int c = 0; c < nres; c++) {
    f.res[c] = h->plan.results[c];
    int w = std::max(1, type_width(f.res[c].rtype));
    GZ_TRY(ctx_alloc(ctx, (size_t)std::max<int64_t>(n, 1) * w, &h->res_values[c]));
    GZ_TRY(ctx_alloc(ctx, (size_t)std::max<int64_t>(n, 1), &h->res_valid[c]));
    f.out_values[c] = h->res_values[c].p;
    f.out_valid[c] = (uint8_t*)h->res_valid[c].p;
  }
  if (n > 0) {
    KernelTimer kt(ctx, "agg_finish");
    unsigned gx = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(agg_finish_kernel, dim3(gx), dim3(256), sizeof(int64_t) * 256 * std::max(f.nl, 1),
                       ctx->stream, f);
  }
  GZ_HIP(hipGetLastError());
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  h->finished = true;
  return GZ_OK;
}

// ---- planned form: prepare (host decisions, allocation, uploads), then a
// kernel-only run that can be captured into a HIP graph and replayed.
static void drop_graph(gz_agg* h) {
  if (h->L.exec) hipGraphExecDestroy(h->L.exec);
  if (h->L.graph) hipGraphDestroy(h->L.graph);
  h->L.exec = nullptr;
  h->L.graph = nullptr;
}

// A plan's table for instance i: no initialization and no rehash (every run of
// the plan resets its tables first).
static int plan_table_alloc(gz_agg* h, int i, uint64_t cap) {
  gz_context* ctx = h->ctx;
  const int nl = (int)h->plan.lanes.size();
  ctx_free(ctx, &h->tables[i]);
  ctx_free(ctx, &h->counters[i]);
  GZ_TRY(ctx_alloc(ctx, sizeof(int64_t) * (cap + 2) * (2 + nl), &h->tables[i]));
  GZ_TRY(ctx_alloc(ctx, 16, &h->counters[i]));
  AggInst ni{};
  ni.base = (int64_t*)h->tables[i].p;
  ni.cap = cap;
  ni.counter = (unsigned int*)h->counters[i].p;
  ni.overflow = (int*)((char*)h->counters[i].p + 8);
  ni.nl = nl;
  h->inst[i] = ni;
  h->inst_dirty = true;
  return GZ_OK;
}

static int agg_prepare(gz_agg* h, int32_t nb, const gz_batch* batches, const int32_t* inst_of) {
  gz_context* ctx = h->ctx;
  GZ_HIP(hipSetDevice(ctx->device));
  drop_graph(h);
  h->L.ready = false;
  const int tile_rows = h->opts.tile_rows > 0 ? h->opts.tile_rows : 4096;
  std::vector<AggTile> tiles;
  std::vector<int64_t> rows;
  // a prepared run aggregates its batches from scratch: instance rows start at 0
  std::fill(h->rows_seen.begin(), h->rows_seen.end(), 0);
  std::vector<GBatch> gb;
  // the plan's host arrays go up in two transfers: the batch table (and the
  // tiles the estimate reads) before the estimate, everything else after it
  UploadBatch ub;
  GZ_TRY(agg_stage(h, nb, batches, inst_of, tiles, rows, tile_rows, &h->L.d_bufs, &gb, &ub));
  AggParams p = base_params(h);
  p.bufs = (const void* const*)h->L.d_bufs.p;  // the plan's own batch table
  p.bstride = h->plan.ncols * 3;
  const double lf = h->opts.load_factor;
  // cardinality per instance: the hint, else HyperLogLog over its rows
  std::vector<double> est(h->ninst, 0.0);
  bool need_est = false;
  for (int i = 0; i < h->ninst; i++)
    if (rows[i] > 4096 && h->opts.expected_groups <= 0) need_est = true;
  if (need_est && !tiles.empty()) {
    GZ_TRY(ub.add(ctx, &h->d_tiles, tiles.data(), sizeof(AggTile) * tiles.size()));
    GZ_TRY(ub.flush(ctx, &h->d_blob));
    p.tiles = (const AggTile*)h->d_tiles.p;
    p.T = (int64_t)tiles.size();
    GZ_TRY(ctx_ensure(ctx, &h->d_regs, sizeof(unsigned) * kHllRegs * h->ninst));
    GZ_TRY(ctx_ensure(ctx, &h->d_est, sizeof(double) * h->ninst));
    int64_t n32 = (int64_t)kHllRegs * h->ninst;
    hipLaunchKernelGGL(fill_u32_kernel_agg, dim3((unsigned)std::min<int64_t>((n32 + 255) / 256, 4096)), dim3(256), 0,
                       ctx->stream, (uint32_t*)h->d_regs.p, n32);
    int64_t blocks = std::min<int64_t>(p.T, (int64_t)ctx->num_cus * 4);
    p.tiles_per_block = (int)((p.T + blocks - 1) / blocks);
    blocks = (p.T + p.tiles_per_block - 1) / p.tiles_per_block;
    hipLaunchKernelGGL(agg_hll_kernel, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, p,
                       (unsigned int*)h->d_regs.p);
    hipLaunchKernelGGL(agg_hll_estimate_kernel, dim3((unsigned)h->ninst), dim3(256), 0, ctx->stream,
                       (const unsigned int*)h->d_regs.p, (double*)h->d_est.p);
    GZ_HIP(hipMemcpyAsync(est.data(), h->d_est.p, sizeof(double) * h->ninst, hipMemcpyDeviceToHost, ctx->stream));
  }
  // while the estimate runs: every instance's tiles cut to kPartTile rows, the
  // tiles of one instance contiguous (the partition scan relies on it)
  std::vector<AggTile> split;
  split.reserve(tiles.size() * ((tile_rows + kPartTile - 1) / kPartTile));
  for (auto& t : tiles)
    for (int32_t r0 = 0; r0 < t.nrows; r0 += kPartTile) {
      AggTile u = t;
      u.row_begin = t.row_begin + r0;
      u.nrows = std::min<int32_t>(kPartTile, t.nrows - r0);
      u.row_base = t.row_base + r0;
      split.push_back(u);
    }
  if (!std::is_sorted(split.begin(), split.end(), [](const AggTile& a, const AggTile& b) { return a.inst < b.inst; }))
    std::stable_sort(split.begin(), split.end(), [](const AggTile& a, const AggTile& b) { return a.inst < b.inst; });
  if (need_est && !tiles.empty()) GZ_HIP(hipStreamSynchronize(ctx->stream));
  // LDS table of the partitioned update: lcap slots (a power of two), so that
  // two 512-thread (or four 256-thread) workgroups share a CU
  int part_lcap = 1024;
  while (part_lcap > 64 && (lds_bytes(h, part_lcap) > 96 * 1024 ||
                            (h->upd_block == 256 && lds_bytes(h, part_lcap) > 40 * 1024)))
    part_lcap >>= 1;
  int part_sub = 0;
  while ((1 << part_sub) < part_lcap) part_sub++;
  // tables (reused when large enough: the run resets them) and placement.  A
  // partitioned instance's table is partition-structured: its partitions'
  // slot ranges have the update's LDS table layout (lcap slots each).
  std::vector<int64_t> prefix(h->ninst + 1, 0);
  std::vector<int64_t> ptiles(h->ninst, 0);  // partition tiles per instance
  for (auto& t : split) ptiles[t.inst]++;
  for (int i = 0; i < h->ninst; i++) {
    double g;
    if (rows[i] == 0) g = 1;
    else if (h->opts.expected_groups > 0) g = std::min<double>((double)h->opts.expected_groups, (double)rows[i]);
    else if (rows[i] > 4096) g = std::min<double>(est[i] * 1.3 + 64, (double)rows[i]);
    else g = (double)rows[i];
    // the partitioned placement: up to kMaxFanout partitions, and a bounded
    // [partition][tile] count block
    int lg = 1;
    while (lg < 12 && (double)(1 << lg) * h->part_groups < g) lg++;
    const bool part_ok = p.nu >= 0 && h->plan.key_col >= 0 && h->part_groups > 0 &&
                         g <= (double)kMaxFanout * h->part_groups && ptiles[i] > 0 &&
                         ((int64_t)ptiles[i] << lg) <= (1ll << 26);
    if (h->opts.placement == 0)
      h->placement[i] = (g <= (double)h->lds_threshold || h->plan.key_col < 0) ? 1 : part_ok ? 3 : 2;
    else
      h->placement[i] = h->opts.placement == 3 && !part_ok ? 2 : h->opts.placement;
    if (h->placement[i] != 3) lg = 0;
    const uint64_t cap = h->plan.key_col < 0 ? 16
                         : lg ? (uint64_t)part_lcap << lg
                              : pow2_at_least((uint64_t)(g / lf) + 16);
    if (!h->tables[i].p || h->inst[i].cap != cap) GZ_TRY(plan_table_alloc(h, i, cap));
    h->g_bound[i] = 2 + (int64_t)g;
    h->inst[i].fan_log2 = lg;
    h->inst[i].sub_log2 = lg ? part_sub : 0;
    h->inst[i].bucket_base = 0;
    prefix[i + 1] = prefix[i] + (h->inst[i].sub_log2 ? 0 : (int64_t)h->inst[i].cap) + 2;
  }
  // partitioned instances, small fan-outs first (their tiles form a scatter
  // launch with a small LDS footprint): partition numbering, their tiles in
  // that order, and each one's [partition][tile] count block
  std::vector<int> porder;
  for (int i = 0; i < h->ninst; i++)
    if (h->placement[i] == 3) porder.push_back(i);
  std::stable_sort(porder.begin(), porder.end(), [&](int x, int y) {
    return (h->inst[x].fan_log2 > 8) < (h->inst[y].fan_log2 > 8);
  });
  std::vector<int> rank(h->ninst, -1);
  for (size_t k = 0; k < porder.size(); k++) rank[porder[k]] = (int)k;
  std::vector<AggTile> part_tiles;
  for (auto& t : split)
    if (h->placement[t.inst] == 3) part_tiles.push_back(t);
  std::stable_sort(part_tiles.begin(), part_tiles.end(),
                   [&](const AggTile& x, const AggTile& y) { return rank[x.inst] < rank[y.inst]; });
  int B = 0;
  int64_t ncnt = 0, tpos = 0;
  std::vector<int> pinst;
  h->part_small_tiles = 0;
  h->part_fcap_small = h->part_fcap_large = 0;
  for (int i : porder) {
    AggInst& I = h->inst[i];
    const int F = 1 << I.fan_log2;
    I.bucket_base = B;
    for (int b = 0; b < F; b++) pinst.push_back(i);
    B += F;
    I.tile0 = (int32_t)tpos;
    I.cnt_tiles = (int32_t)ptiles[i];
    I.cnt_off = ncnt;
    ncnt += (int64_t)F * ptiles[i];
    tpos += ptiles[i];
    if (F <= kSmallFanout) {
      h->part_small_tiles += ptiles[i];
      h->part_fcap_small = std::max(h->part_fcap_small, F);
    } else {
      h->part_fcap_large = std::max(h->part_fcap_large, F);
    }
  }
  h->B = B;
  h->ncnt = ncnt;
  GZ_TRY(ub.add(ctx, &h->d_inst, h->inst.data(), sizeof(AggInst) * h->inst.size()));
  h->inst_dirty = false;
  GZ_TRY(ub.add(ctx, &h->L.d_slot_prefix, prefix.data(), sizeof(int64_t) * prefix.size()));
  h->L.total_slots = prefix[h->ninst];
  std::vector<int> ops(std::max<size_t>(h->plan.lanes.size(), 1));
  for (size_t l = 0; l < h->plan.lanes.size(); l++) ops[l] = h->plan.lanes[l].op;
  GZ_TRY(ub.add(ctx, &h->d_lane_ops, ops.data(), sizeof(int) * ops.size()));
  p.inst = (const AggInst*)h->d_inst.p;
  // launch plan
  std::vector<AggTile> lds_tiles, hbm_tiles;
  for (auto& t : tiles) {
    if (h->placement[t.inst] == 3) continue;
    (h->placement[t.inst] == 1 ? lds_tiles : hbm_tiles).push_back(t);
  }
  h->grid_part = 0;
  h->nrec = 0;
  if (!part_tiles.empty()) {
    for (auto& t : part_tiles) h->nrec += t.nrows;
    GZ_TRY(ub.add(ctx, &h->d_part, part_tiles.data(), sizeof(AggTile) * part_tiles.size()));
    GZ_TRY(ub.add(ctx, &h->pinst, pinst.data(), sizeof(int) * pinst.size()));
    GZ_TRY(ctx_ensure(ctx, &h->pcnt, sizeof(uint32_t) * ncnt));
    GZ_TRY(ctx_ensure(ctx, &h->pprefix, sizeof(uint64_t) * ncnt));
    size_t tmp = 0;
    GZ_HIP(rocprim::exclusive_scan(nullptr, tmp, (const uint32_t*)h->pcnt.p, (uint64_t*)h->pprefix.p, (uint64_t)0,
                                   (size_t)ncnt, rocprim::plus<uint64_t>(), ctx->stream));
    GZ_TRY(ctx_ensure(ctx, &h->pscan_tmp, std::max<size_t>(tmp, 16)));
    h->pscan_bytes = tmp;
    GZ_TRY(ctx_ensure(ctx, &h->recs, sizeof(int64_t) * (2 + p.nu) * h->nrec));
    AggParams& q = h->p_part;
    q = p;
    q.rstride = h->nrec;
    q.tiles = (const AggTile*)h->d_part.p;
    q.T = (int64_t)part_tiles.size();
    h->grid_part = (unsigned)((q.T + 7) / 8 * 8);
    q.lcap = part_lcap;  // the structured tables' partition size
    h->lds_part = lds_bytes(h, q.lcap);
    // work items: one per partition plus one per further kPartItemRecords records
    const int64_t max_items = B + (h->nrec + kPartItemRecords - 1) / kPartItemRecords + 1;
    GZ_TRY(ctx_ensure(ctx, &h->pitems, sizeof(PartItem) * max_items));
    GZ_TRY(ctx_ensure(ctx, &h->pnitems, sizeof(int64_t)));
    h->blocks_part_upd = (unsigned)std::max<int64_t>(
        1, std::min<int64_t>(max_items, (int64_t)(1024 / h->upd_block) * ctx->num_cus));
  }
  h->L.p_hbm = p;
  h->L.p_lds = p;
  h->L.grid_hbm = 0;
  h->L.blocks_hbm = 0;
  h->L.blocks_lds = 0;
  if (!hbm_tiles.empty()) {
    GZ_TRY(ub.add(ctx, &h->L.d_hbm, hbm_tiles.data(), sizeof(AggTile) * hbm_tiles.size()));
    h->L.p_hbm.tiles = (const AggTile*)h->L.d_hbm.p;
    h->L.p_hbm.T = (int64_t)hbm_tiles.size();
    if (h->L.p_hbm.T >= 8ll * ctx->num_cus) hbm_first_level(h, &h->L.p_hbm, &h->L.blocks_hbm, &h->L.lds_hbm);
    else h->L.grid_hbm = (unsigned)((h->L.p_hbm.T + 7) / 8 * 8);
  }
  if (!lds_tiles.empty()) {
    GZ_TRY(ub.add(ctx, &h->L.d_lds, lds_tiles.data(), sizeof(AggTile) * lds_tiles.size()));
    AggParams& q = h->L.p_lds;
    q.tiles = (const AggTile*)h->L.d_lds.p;
    q.T = (int64_t)lds_tiles.size();
    int64_t blocks = std::min<int64_t>(q.T, (int64_t)ctx->num_cus * 2);
    q.tiles_per_block = (int)((q.T + blocks - 1) / blocks);
    h->L.blocks_lds = (unsigned)((q.T + q.tiles_per_block - 1) / q.tiles_per_block);
    q.lcap = 1024;
    int lds = lds_bytes(h, q.lcap);
    while (lds > 96 * 1024 && q.lcap > 64) {
      q.lcap >>= 1;
      lds = lds_bytes(h, q.lcap);
    }
    h->L.lds = lds;
  }
  h->L.T_fix = 0;
  h->L.n_gb = 0;
  if (h->need_fix && !tiles.empty()) {
    GZ_TRY(ub.add(ctx, &h->L.d_fix, tiles.data(), sizeof(AggTile) * tiles.size()));
    h->L.p_fix = p;
    h->L.p_fix.tiles = (const AggTile*)h->L.d_fix.p;
    h->L.p_fix.T = h->L.T_fix = (int64_t)tiles.size();
  }
  if (h->need_allnull && !gb.empty()) {
    GZ_TRY(ub.add(ctx, &h->L.d_gb, gb.data(), sizeof(GBatch) * gb.size()));
    h->L.n_gb = (int)gb.size();
  }
  GZ_TRY(ub.flush(ctx, &h->d_blob));
  h->L.rows = rows;
  h->L.ready = true;
  GZ_HIP(hipGetLastError());
  return GZ_OK;
}

// Kernels only: no allocation, no host synchronization, no memset nodes.
static int agg_run_kernels(gz_agg* h) {
  gz_context* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  const int nl = (int)h->plan.lanes.size();
  unsigned gi = (unsigned)std::min<int64_t>((h->L.total_slots + 255) / 256, 16384);
  hipLaunchKernelGGL(agg_init_all_kernel, dim3(gi), dim3(256), 0, st, (const AggInst*)h->d_inst.p, h->ninst,
                     (const int64_t*)h->L.d_slot_prefix.p, h->L.total_slots, nl, (const int*)h->d_lane_ops.p);
  if (h->L.blocks_hbm)
    hipLaunchKernelGGL(agg_update_lds_kernel, dim3(h->L.blocks_hbm), dim3(256), h->L.lds_hbm, st, h->L.p_hbm);
  if (h->L.grid_hbm)
    hipLaunchKernelGGL(agg_update_kernel, dim3(h->L.grid_hbm), dim3(256), 0, st, h->L.p_hbm);
  if (h->L.blocks_lds)
    hipLaunchKernelGGL(agg_update_lds_kernel, dim3(h->L.blocks_lds), dim3(256), h->L.lds, st, h->L.p_lds);
  if (h->grid_part) {
    const int64_t TP = h->p_part.T, TA = h->part_small_tiles;
    const int W = 2 + h->p_part.nu;
    hipLaunchKernelGGL(agg_part_hist_kernel, dim3(h->grid_part), dim3(256), 0, st, h->p_part, (uint32_t*)h->pcnt.p);
    GZ_HIP(rocprim::exclusive_scan(h->pscan_tmp.p, h->pscan_bytes, (const uint32_t*)h->pcnt.p,
                                   (uint64_t*)h->pprefix.p, (uint64_t)0, (size_t)h->ncnt, rocprim::plus<uint64_t>(),
                                   st));
    hipLaunchKernelGGL(fill_u32_kernel_agg, dim3(1), dim3(64), 0, st, (uint32_t*)h->pnitems.p, (int64_t)2);
    hipLaunchKernelGGL(agg_part_items_kernel, dim3((unsigned)((h->B + 255) / 256)), dim3(256), 0, st,
                       (const uint64_t*)h->pprefix.p, (const AggInst*)h->d_inst.p, (const int*)h->pinst.p, h->B,
                       h->nrec, (PartItem*)h->pitems.p, (unsigned long long*)h->pnitems.p);
    hipLaunchKernelGGL(agg_part_reset_kernel, dim3((unsigned)h->B), dim3(256), 0, st, (const uint64_t*)h->pprefix.p,
                       (const AggInst*)h->d_inst.p, (const int*)h->pinst.p, h->B, h->nrec, nl,
                       (const int*)h->d_lane_ops.p);
    // scatter: the small fan-outs' tiles, then the large ones' (per-partition LDS arrays sized for each)
    for (int grp = 0; grp < 2; grp++) {
      AggParams q = h->p_part;
      const int64_t t0 = grp ? TA : 0, t1 = grp ? TP : TA;
      if (t1 <= t0) continue;
      q.tiles = h->p_part.tiles + t0;
      q.T = t1 - t0;
      q.tile_base = t0;
      const int fcap = grp ? h->part_fcap_large : h->part_fcap_small;
      hipLaunchKernelGGL(agg_part_scatter_kernel, dim3((unsigned)((q.T + 7) / 8 * 8)), dim3(256),
                         scatter_lds(W, fcap), st, q, (const uint64_t*)h->pprefix.p, fcap, (int64_t*)h->recs.p);
    }
    if (h->upd_block == 256)
      hipLaunchKernelGGL(agg_part_update_kernel<256>, dim3(h->blocks_part_upd), dim3(256), h->lds_part, st,
                         h->p_part, (const int64_t*)h->recs.p, (const PartItem*)h->pitems.p, h->B,
                         (const unsigned long long*)h->pnitems.p, (const int*)h->pinst.p);
    else
      hipLaunchKernelGGL(agg_part_update_kernel<512>, dim3(h->blocks_part_upd), dim3(512), h->lds_part, st,
                         h->p_part, (const int64_t*)h->recs.p, (const PartItem*)h->pitems.p, h->B,
                         (const unsigned long long*)h->pnitems.p, (const int*)h->pinst.p);
  }
  if (h->L.T_fix)
    hipLaunchKernelGGL(agg_fix_kernel, dim3((unsigned)h->L.T_fix), dim3(256), 0, st, h->L.p_fix);
  if (h->L.n_gb)
    hipLaunchKernelGGL(agg_allnull_kernel, dim3((unsigned)h->L.n_gb), dim3(256), 0, st, h->L.p_hbm,
                       (const GBatch*)h->L.d_gb.p, h->L.n_gb);
  GZ_HIP(hipGetLastError());
  return GZ_OK;
}

static void mark_run(gz_agg* h) {
  for (int i = 0; i < h->ninst; i++) h->rows_seen[i] = h->L.rows[i];
  h->finished = false;
}

static int agg_run(gz_agg* h) {
  if (!h->L.ready) return fail(GZ_INVALID, "gz_agg_set_prepare first");
  GZ_HIP(hipSetDevice(h->ctx->device));
  {
    KernelTimer kt(h->ctx, "agg_run");
    GZ_TRY(agg_run_kernels(h));
  }
  mark_run(h);
  return GZ_OK;
}

static int agg_capture(gz_agg* h) {
  if (!h->L.ready) return fail(GZ_INVALID, "gz_agg_set_prepare first");
  GZ_HIP(hipSetDevice(h->ctx->device));
  drop_graph(h);
  hipStream_t st = h->ctx->stream;
  GZ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = agg_run_kernels(h);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(GZ_HIP_ERROR, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  h->L.graph = g;
  e = hipGraphInstantiate(&h->L.exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(h);
    return fail(GZ_HIP_ERROR, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return GZ_OK;
}

static int agg_launch(gz_agg* h) {
  if (!h->L.exec) return agg_run(h);
  GZ_HIP(hipSetDevice(h->ctx->device));
  {
    KernelTimer kt(h->ctx, "agg_run");
    GZ_HIP(hipGraphLaunch(h->L.exec, h->ctx->stream));
  }
  mark_run(h);
  return GZ_OK;
}

// Autotuner: a coordinate descent over the plan's knobs, each candidate
// prepared and timed over two runs (after a warm one) on the given batches:
// starting from the default plan, every dimension in turn (update tile rows,
// LDS placement threshold, groups per partition, i.e. the partitioned fan-out,
// table load factor, workgroup size of the partitioned update) tries its other
// values with the others at their best so far and keeps the fastest.  A
// candidate whose table overflowed is not eligible.  *chosen: the number of the
// winning candidate in the order tried (0: the default plan).
static int agg_autotune(gz_agg* h, int32_t nb, const gz_batch* batches, const int32_t* inst_of, int32_t* chosen) {
  struct Knobs {
    int tile_rows, threshold, part, block;
    double lf;
  };
  gz_context* ctx = h->ctx;
  const gz_agg_options saved = h->opts;
  const int saved_th = h->lds_threshold, saved_part = h->part_groups, saved_block = h->upd_block;
  const bool free_placement = saved.placement == 0;
  auto apply = [&](const Knobs& k) {
    h->opts.tile_rows = k.tile_rows;
    h->opts.load_factor = k.lf;
    h->upd_block = k.block;
    if (free_placement) {
      h->lds_threshold = k.threshold;
      h->part_groups = k.part;
    }
  };
  hipEvent_t e0, e1;
  GZ_HIP(hipEventCreate(&e0));
  GZ_HIP(hipEventCreate(&e1));
  int rc = GZ_OK, tried = 0, best_id = -1;
  float best_ms = 0;
  Knobs best{h->opts.tile_rows > 0 ? h->opts.tile_rows : 4096, saved_th, saved_part, saved_block,
             saved.load_factor > 0 ? saved.load_factor : 0.5};
  auto time_plan = [&](const Knobs& k, float* ms, bool* ok) -> int {
    apply(k);
    GZ_TRY(agg_prepare(h, nb, batches, inst_of));
    GZ_TRY(agg_run_kernels(h));  // warm
    GZ_HIP(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < 2; r++) GZ_TRY(agg_run_kernels(h));
    GZ_HIP(hipEventRecord(e1, ctx->stream));
    GZ_HIP(hipEventSynchronize(e1));
    std::vector<int64_t> g;
    int ovf = 0;
    GZ_TRY(read_counts(h, g, &ovf));
    GZ_HIP(hipEventElapsedTime(ms, e0, e1));
    *ok = !ovf;
    return GZ_OK;
  };
  auto consider = [&](const Knobs& k) {
    if (rc) return;
    float ms = 0;
    bool ok = false;
    rc = time_plan(k, &ms, &ok);
    const int id = tried++;
    if (rc == GZ_OK && ok && (best_id < 0 || ms < best_ms)) {
      best_id = id;
      best_ms = ms;
      best = k;
    }
  };
  consider(best);  // the default plan
  for (int dim = 0; dim < 5 && rc == GZ_OK; dim++) {
    if (!free_placement && (dim == 1 || dim == 2)) continue;  // placement fixed by the caller
    const Knobs base = best;
    std::vector<Knobs> alts;
    auto with = [&](auto set) {
      Knobs k = base;
      set(k);
      alts.push_back(k);
    };
    switch (dim) {
      case 0: for (int v : {2048, 4096, 8192}) if (v != base.tile_rows) with([&](Knobs& k) { k.tile_rows = v; }); break;
      case 1: for (int v : {0, 256, 1024}) if (v != base.threshold) with([&](Knobs& k) { k.threshold = v; }); break;
      case 2: for (int v : {0, 256, 512, 768, 1024}) if (v != base.part) with([&](Knobs& k) { k.part = v; }); break;
      case 3: for (double v : {0.35, 0.5, 0.7}) if (v != base.lf) with([&](Knobs& k) { k.lf = v; }); break;
      case 4: for (int v : {256, 512}) if (v != base.block) with([&](Knobs& k) { k.block = v; }); break;
    }
    for (auto& k : alts) consider(k);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (rc == GZ_OK && best_id < 0) rc = fail(GZ_CAPACITY_ERROR, "no candidate plan fits");
  if (rc) {
    h->opts = saved;
    h->lds_threshold = saved_th;
    h->part_groups = saved_part;
    h->upd_block = saved_block;
    return rc;
  }
  apply(best);
  GZ_TRY(agg_prepare(h, nb, batches, inst_of));
  if (chosen) *chosen = best_id;
  return GZ_OK;
}

extern "C" {

void gz_agg_options_defaults(gz_agg_options* o) {
  if (!o) return;
  o->placement = 0;
  o->tile_rows = 0;
  o->load_factor = 0.5;
  o->expected_groups = 0;
}

int gz_agg_make(gz_context* ctx, int32_t num_cols, const int32_t* col_types, int32_t num_actions,
                const gz_agg_action* actions, const gz_agg_options* opts, gz_agg** out) {
  return agg_set_make(ctx, num_cols, col_types, num_actions, actions, opts, 1, out);
}

int gz_join_agg_make(gz_relation* rel, gz_context* ctx, int32_t num_probe_cols, const int32_t* probe_types,
                     int32_t probe_key_col, int32_t num_inputs, const int32_t* input_sides,
                     const int32_t* input_indices, int32_t num_actions, const gz_agg_action* actions,
                     const gz_agg_options* opts, gz_agg** out) {
  if (!rel || !out || !probe_types || !input_sides || !input_indices || !actions)
    return fail(GZ_INVALID, "gz_join_agg_make: null argument");
  GZ_TRY(relation_synchronize(rel));
  RelationView v;
  GZ_TRY(relation_view(rel, &v));
  if (!ctx) ctx = v.ctx;
  if (num_probe_cols <= 0 || num_probe_cols > kAggMaxCols || probe_key_col < 0 || probe_key_col >= num_probe_cols)
    return fail(GZ_INVALID, "bad probe columns");
  if (probe_types[probe_key_col] != v.types[v.key_col])
    return fail(GZ_TYPE_ERROR, "probe and build key types differ");
  if (num_inputs <= 0 || num_inputs > kAggMaxCols) return fail(GZ_INVALID, "bad input column count");
  std::vector<int32_t> types(num_inputs);
  for (int c = 0; c < num_inputs; c++) {
    const int sd = input_sides[c], ix = input_indices[c];
    if (sd == 0 && (ix < 0 || ix >= v.ncols)) return fail(GZ_INVALID, "build column out of range");
    if (sd == 1 && (ix < 0 || ix >= num_probe_cols)) return fail(GZ_INVALID, "probe column out of range");
    if (sd != 0 && sd != 1) return fail(GZ_INVALID, "input side must be 0 (build) or 1 (probe)");
    types[c] = sd == 0 ? v.types[ix] : probe_types[ix];
    if (type_is_binary(types[c])) return fail(GZ_NOT_IMPLEMENTED, "aggregate inputs must be fixed-width");
  }
  gz_agg* h = nullptr;
  GZ_TRY(agg_set_make(ctx, num_inputs, types.data(), num_actions, actions, opts, 1, &h));
  auto bail = [&](int rc) {
    gz_agg_release(h);
    return rc;
  };
  const int kc = h->plan.key_col;
  if (kc < 0) return bail(fail(GZ_NOT_IMPLEMENTED, "a join aggregate groups by the join key"));
  if (is_float_t(h->plan.key_type)) return bail(fail(GZ_NOT_IMPLEMENTED, "a join aggregate over a float join key"));
  const bool key_ok = (input_sides[kc] == 1 && input_indices[kc] == probe_key_col) ||
                      (input_sides[kc] == 0 && input_indices[kc] == v.key_col);
  if (!key_ok) return bail(fail(GZ_NOT_IMPLEMENTED, "a join aggregate groups by the join key"));
  // the pairs of one probe row share its ordinal, so a value-keeping float
  // min / max must not see different values of one probe row: refuse a build
  // column when build keys repeat
  if (v.rowlist)
    for (auto& l : h->plan.lanes)
      if ((l.op == L_FIRSTNN || l.op == L_FIRSTNAN || l.op == L_FIRSTZ || l.op == L_LASTNAN ||
           l.op == L_NANBITS) && l.col >= 0 && input_sides[l.col] == 0 && l.col != kc)
        return bail(fail(GZ_NOT_IMPLEMENTED, "float min/max of a build column over repeated build keys"));
  h->join = new JoinAggState();
  h->join->rel = rel;
  h->join->ptypes.assign(probe_types, probe_types + num_probe_cols);
  h->join->pkey = probe_key_col;
  h->join->sides.assign(input_sides, input_sides + num_inputs);
  h->join->idx.assign(input_indices, input_indices + num_inputs);
  *out = h;
  return GZ_OK;
}

int gz_join_agg_process_batches(gz_agg* h, int32_t nb, const gz_batch* batches) {
  if (!h || !h->join) return fail(GZ_INVALID, "not a join aggregate");
  if (nb < 0 || (nb > 0 && !batches)) return fail(GZ_INVALID, "bad batch list");
  if (h->finished) return fail(GZ_INVALID, "aggregate already finished");
  JoinAggState& J = *h->join;
  gz_context* ctx = h->ctx;
  GZ_HIP(hipSetDevice(ctx->device));
  RelationView v;
  GZ_TRY(relation_view(J.rel, &v));
  const int nc = (int)J.ptypes.size();
  constexpr int kTile = 4096;
  std::vector<JoinTile> tiles;
  std::vector<const void*> bufs((size_t)std::max(nb, 1) * nc * 3, nullptr);
  int64_t call_rows = 0;
  for (int bi = 0; bi < nb; bi++) {
    const gz_batch& bt = batches[bi];
    if (bt.num_rows < 0 || bt.num_rows > INT32_MAX) return fail(GZ_INVALID, "bad num_rows");
    int need = 0;
    for (int c = 0; c < nc; c++) need += type_num_bufs(J.ptypes[c]);
    if (bt.num_bufs < need) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
    int k = 0;
    for (int c = 0; c < nc; c++) {
      const int t = J.ptypes[c];
      const void** slot = &bufs[((size_t)bi * nc + c) * 3];
      slot[0] = (bt.buf_addrs[k] && bt.buf_sizes[k] > 0) ? (const void*)bt.buf_addrs[k] : nullptr;
      if (slot[0] && bt.buf_sizes[k] < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
      k++;
      if (type_is_binary(t)) slot[1] = (const void*)bt.buf_addrs[k++];
      const int64_t need_b = t == GZ_TYPE_BOOL ? (bt.num_rows + 7) / 8 : bt.num_rows * type_width(t);
      if (t != GZ_TYPE_NA && !type_is_binary(t) && bt.buf_sizes[k] < need_b)
        return fail(GZ_INVALID, "values buffer too small");
      slot[2] = (const void*)bt.buf_addrs[k++];
      if (t != GZ_TYPE_NA && bt.num_rows > 0 && !slot[2]) return fail(GZ_INVALID, "missing values buffer");
    }
    for (int64_t r0 = 0; r0 < bt.num_rows; r0 += kTile) {
      JoinTile t{};
      t.batch = bi;
      t.row_begin = (int32_t)r0;
      t.nrows = (int32_t)std::min<int64_t>(kTile, bt.num_rows - r0);
      t.row_base = J.probe_rows + call_rows + r0;
      tiles.push_back(t);
    }
    call_rows += bt.num_rows;
  }
  if (call_rows + J.probe_rows >= (1ll << 44)) return fail(GZ_CAPACITY_ERROR, "more than 2^44 probe rows");
  J.probe_rows += call_rows;
  h->rows_seen[0] += call_rows;
  if (tiles.empty()) return GZ_OK;
  // groups: at most the relation's distinct keys, at most the probe rows so far
  const double lf = h->opts.load_factor;
  const int64_t bound = 2 + std::min<int64_t>(v.nkeys, h->g_bound[0] + call_rows);
  if (!h->tables[0].p) {
    GZ_TRY(inst_alloc(h, 0, pow2_at_least((uint64_t)(bound / lf) + 16)));
  } else if ((double)bound > lf * h->inst[0].cap) {
    GZ_TRY(inst_alloc(h, 0, pow2_at_least((uint64_t)(bound * 2 / lf))));
  }
  h->g_bound[0] = bound;
  if (h->inst_dirty) {
    GZ_TRY(upload(ctx, &h->d_inst, h->inst.data(), sizeof(AggInst) * h->inst.size()));
    h->inst_dirty = false;
  }
  GZ_TRY(upload(ctx, &h->d_tiles, tiles.data(), sizeof(JoinTile) * tiles.size()));
  GZ_TRY(upload(ctx, &h->d_bufs, bufs.data(), sizeof(void*) * bufs.size()));
  AggParams p = base_params(h);
  p.inst = (const AggInst*)h->d_inst.p;
  JoinAggParams q;
  memset(&q, 0, sizeof(q));
  q.tiles = (const JoinTile*)h->d_tiles.p;
  q.T = (int64_t)tiles.size();
  q.bufs = (const void* const*)h->d_bufs.p;
  q.bstride = nc * 3;
  q.pkey = J.pkey;
  q.key_type = J.ptypes[J.pkey];
  q.jkeys = v.keys;
  q.jstart = v.start;
  q.jcount = v.count;
  q.jrowlist = v.rowlist;
  q.jcap = v.cap;
  q.special_start = v.special_start;
  q.special_count = v.special_count;
  for (int c = 0; c < h->plan.ncols; c++) {
    JoinIn& in = q.in[c];
    in.side = J.sides[c];
    in.idx = J.idx[c];
    in.type = h->plan.types[c];
    if (in.side == 0) {
      in.bvalid = v.valid[in.idx];
      in.bvalues = (const uint8_t*)v.values[in.idx];
    }
  }
  // four workgroups per CU: the LDS table holds the head of a skewed key
  // distribution (512 keys), the rest spills to HBM
  int64_t blocks = std::min<int64_t>(q.T, (int64_t)ctx->num_cus * 4);
  q.tiles_per_block = (int)((q.T + blocks - 1) / blocks);
  blocks = (q.T + q.tiles_per_block - 1) / q.tiles_per_block;
  p.lcap = 512;
  p.lprobe = 8;
  int lds = lds_bytes(h, p.lcap);
  while (lds > 40 * 1024 && p.lcap > 64) {
    p.lcap >>= 1;
    lds = lds_bytes(h, p.lcap);
  }
  {
    KernelTimer kt(ctx, "join_agg_update");
    hipLaunchKernelGGL(join_agg_update_kernel, dim3((unsigned)blocks), dim3(256), lds, ctx->stream, p, q);
  }
  if (p.fix_flag)
    hipLaunchKernelGGL(join_agg_fix_kernel, dim3((unsigned)q.T), dim3(256), 0, ctx->stream, p, q);
  GZ_HIP(hipGetLastError());
  return GZ_OK;
}

// Runs the staged per-batch calls through the batched path.
static int agg_flush(gz_agg* h) {
  if (h->stage.empty()) return GZ_OK;
  std::vector<gz_batch> v;
  stage_view(h->stage, &v);
  const int rc = h->mk ? mk_process(h->mk, (int32_t)v.size(), v.data()) : agg_process(h, (int32_t)v.size(), v.data(), nullptr);
  stage_clear(&h->stage);  // the arena is reused by later copies, queued after this run's kernels
  return rc;
}

int gz_agg_process(gz_agg* h, int64_t num_rows, const int64_t* buf_addrs, const int64_t* buf_sizes,
                   int32_t num_bufs) {
  if (!h) return fail(GZ_INVALID, "null aggregate");
  if (num_rows == 0) return GZ_OK;  // ProcessAndCacheOne returns early on empty input
  if (!buf_addrs || !buf_sizes) return fail(GZ_INVALID, "buf_addrs can't be null");
  if (h->join) return fail(GZ_INVALID, "a join aggregate takes probe batches (gz_join_agg_process_batches)");
  gz_batch b{num_rows, buf_addrs, buf_sizes, num_bufs};
  if (h->ninst == 1 && !h->finished && !(h->mk && mk_finished(h->mk))) {
    // staged for the batched path: results are observed only at finish
    GZ_HIP(hipSetDevice(h->ctx->device));
    GZ_TRY(stage_check((int32_t)h->in_types.size(), h->in_types.data(), b));
    bool staged = false;
    GZ_TRY(stage_add(h->ctx, &h->stage, b, &staged));
    if (!staged && !h->stage.empty()) {  // the arena is full: run what it holds, then stage again
      GZ_TRY(agg_flush(h));
      GZ_TRY(stage_add(h->ctx, &h->stage, b, &staged));
    }
    if (staged) return GZ_OK;
  }
  GZ_TRY(agg_flush(h));
  if (h->mk) return mk_process(h->mk, 1, &b);
  return agg_process(h, 1, &b, nullptr);
}

int gz_agg_process_batches(gz_agg* h, int32_t num_batches, const gz_batch* batches) {
  if (!h) return fail(GZ_INVALID, "null aggregate");
  if (num_batches < 0 || (num_batches > 0 && !batches)) return fail(GZ_INVALID, "bad batch list");
  GZ_TRY(agg_flush(h));
  if (h->mk) return mk_process(h->mk, num_batches, batches);
  return agg_process(h, num_batches, batches, nullptr);
}

int gz_agg_num_groups(gz_agg* h, int64_t* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  GZ_TRY(agg_flush(h));
  if (h->mk) return mk_num_groups(h->mk, out);
  if (h->finished) {
    *out = h->ngroups[0];
    return GZ_OK;
  }
  if (h->plan.key_col < 0) {
    *out = h->rows_seen[0] > 0 ? 1 : 0;
    return GZ_OK;
  }
  if (!h->tables[0].p) {
    *out = 0;
    return GZ_OK;
  }
  GZ_TRY(agg_finish(h));
  *out = h->ngroups[0];
  return GZ_OK;
}

int gz_agg_num_result_columns(gz_agg* h) {
  if (h && h->mk) return mk_num_result_columns(h->mk);
  return h ? (int)h->plan.results.size() : 0;
}

int gz_agg_result_type(gz_agg* h, int32_t col) {
  if (h && h->mk) return mk_result_type(h->mk, col);
  if (!h || col < 0 || col >= (int)h->plan.results.size()) return -1;
  return h->plan.results[col].rtype;
}

int gz_agg_finish(gz_agg* h) {
  if (!h) return fail(GZ_INVALID, "null aggregate");
  GZ_TRY(agg_flush(h));
  if (h->mk) return mk_finish(h->mk);
  return agg_finish(h);
}

int gz_agg_result_column(gz_agg* h, int32_t col, void** values, void** valid) {
  if (!h) return fail(GZ_INVALID, "null aggregate");
  if (h->mk) return mk_result_column(h->mk, col, values, valid);
  if (!h->finished) return fail(GZ_INVALID, "call gz_agg_finish first");
  if (col < 0 || col >= (int)h->plan.results.size()) return fail(GZ_INDEX_ERROR, "result column out of range");
  int w = std::max(1, type_width(h->plan.results[col].rtype));
  if (values) *values = (char*)h->res_values[col].p + h->group_off[0] * w;
  if (valid) *valid = (char*)h->res_valid[col].p + h->group_off[0];
  return GZ_OK;
}

int gz_agg_result_binary(gz_agg* h, int32_t col, void** offsets, void** data, void** valid) {
  if (!h) return fail(GZ_INVALID, "null aggregate");
  if (!h->mk) return fail(GZ_TYPE_ERROR, "result column is not utf8");
  return mk_result_binary(h->mk, col, offsets, data, valid);
}

int gz_agg_release(gz_agg* h) {
  if (!h) return GZ_OK;
  stage_release(h->ctx, &h->stage);
  if (h->mk) {
    for (auto& b : h->res_bits) ctx_free(h->ctx, &b);
    mk_release(h->mk);
    delete h;
    return GZ_OK;
  }
  gz_context* ctx = h->ctx;
  for (auto& b : h->tables) ctx_free(ctx, &b);
  for (auto& b : h->counters) ctx_free(ctx, &b);
  for (auto& b : h->res_values) ctx_free(ctx, &b);
  for (auto& b : h->res_valid) ctx_free(ctx, &b);
  for (DBuf* b : {&h->d_cpre, &h->d_inst, &h->d_lane_ops, &h->d_tiles, &h->d_bufs, &h->d_regs, &h->d_est, &h->skeys,
                  &h->svals, &h->tkeys, &h->tvals, &h->cursor, &h->ccpre, &h->counts, &h->sort_tmp, &h->L.d_hbm,
                  &h->L.d_lds, &h->L.d_slot_prefix, &h->L.d_bufs, &h->d_part, &h->pinst,
                  &h->recs, &h->pcnt, &h->pprefix, &h->pscan_tmp, &h->d_fix,
                  &h->d_fix_tiles, &h->d_gbatch, &h->L.d_fix, &h->L.d_gb, &h->d_blob, &h->d_cnt, &h->pitems, &h->pnitems})
    ctx_free(ctx, b);
  for (auto& b : h->res_bits) ctx_free(ctx, &b);
  if (h->L.exec) hipGraphExecDestroy(h->L.exec);
  if (h->L.graph) hipGraphDestroy(h->L.graph);
  delete h->join;
  delete h;
  return GZ_OK;
}

// ---- aggregate sets (many independent GROUP BYs in one launch sequence)
int gz_agg_set_make(gz_context* ctx, int32_t num_cols, const int32_t* col_types, int32_t num_actions,
                    const gz_agg_action* actions, const gz_agg_options* opts, int32_t num_instances,
                    gz_agg** out) {
  return agg_set_make(ctx, num_cols, col_types, num_actions, actions, opts, num_instances, out);
}

int gz_agg_set_process(gz_agg* h, int32_t num_batches, const gz_batch* batches, const int32_t* instance_of) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  if (num_batches < 0 || (num_batches > 0 && (!batches || !instance_of))) return fail(GZ_INVALID, "bad batch list");
  GZ_TRY(agg_flush(h));
  return agg_process(h, num_batches, batches, instance_of);
}

int gz_agg_set_num_groups(gz_agg* h, int32_t inst, int64_t* out) {
  if (!h || !out) return fail(GZ_INVALID, "null argument");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  if (!h->finished) return fail(GZ_INVALID, "call gz_agg_finish first");
  if (inst < 0 || inst >= h->ninst) return fail(GZ_INDEX_ERROR, "instance out of range");
  *out = h->ngroups[inst];
  return GZ_OK;
}

int gz_agg_set_result_column(gz_agg* h, int32_t inst, int32_t col, void** values, void** valid) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  if (!h->finished) return fail(GZ_INVALID, "call gz_agg_finish first");
  if (inst < 0 || inst >= h->ninst) return fail(GZ_INDEX_ERROR, "instance out of range");
  if (col < 0 || col >= (int)h->plan.results.size()) return fail(GZ_INDEX_ERROR, "result column out of range");
  int w = std::max(1, type_width(h->plan.results[col].rtype));
  if (values) *values = (char*)h->res_values[col].p + h->group_off[inst] * w;
  if (valid) *valid = (char*)h->res_valid[col].p + h->group_off[inst];
  return GZ_OK;
}

int gz_agg_set_result_bitmap(gz_agg* h, int32_t inst, int32_t col, void** bitmap) {
  if (!h || !bitmap) return fail(GZ_INVALID, "null argument");
  void* valid = nullptr;
  int64_t n = 0;
  int nres = 0;
  if (h->mk) {
    if (inst != 0) return fail(GZ_INDEX_ERROR, "instance out of range");
    GZ_TRY(mk_result_column(h->mk, col, nullptr, &valid));
    GZ_TRY(mk_num_groups(h->mk, &n));
    nres = gz_agg_num_result_columns(h);
  } else {
    GZ_TRY(gz_agg_set_result_column(h, inst, col, nullptr, &valid));
    n = h->ngroups[inst];
    nres = (int)h->plan.results.size();
  }
  const size_t k = (size_t)inst * nres + col;
  if (h->res_bits.size() < (size_t)h->ninst * nres) {
    h->res_bits.resize((size_t)h->ninst * nres);
    h->bits_ready.resize((size_t)h->ninst * nres, 0);
  }
  if (!h->bits_ready[k]) {
    GZ_TRY(ctx_ensure(h->ctx, &h->res_bits[k], (size_t)std::max<int64_t>((n + 63) / 64 * 8, 8)));
    GZ_TRY(pack_validity(h->ctx, (const uint8_t*)valid, n, (uint8_t*)h->res_bits[k].p));
    h->bits_ready[k] = 1;
  }
  *bitmap = h->res_bits[k].p;
  return GZ_OK;
}

int gz_agg_result_bitmap(gz_agg* h, int32_t col, void** bitmap) { return gz_agg_set_result_bitmap(h, 0, col, bitmap); }

int gz_agg_set_placement(gz_agg* h, int32_t inst, int32_t* out) {
  if (!h || !out || inst < 0 || inst >= h->ninst) return fail(GZ_INVALID, "bad argument");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  *out = h->placement[inst];
  return GZ_OK;
}

int gz_agg_set_prepare(gz_agg* h, int32_t num_batches, const gz_batch* batches, const int32_t* instance_of) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  if (num_batches < 0 || (num_batches > 0 && (!batches || !instance_of))) return fail(GZ_INVALID, "bad batch list");
  GZ_TRY(agg_flush(h));
  return agg_prepare(h, num_batches, batches, instance_of);
}

int gz_agg_set_run(gz_agg* h) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  return agg_run(h);
}

int gz_agg_set_capture(gz_agg* h) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  return agg_capture(h);
}

int gz_agg_set_launch(gz_agg* h) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  return agg_launch(h);
}

int gz_agg_set_autotune(gz_agg* h, int32_t num_batches, const gz_batch* batches, const int32_t* instance_of,
                        int32_t* chosen) {
  if (!h) return fail(GZ_INVALID, "null aggregate set");
  if (h->mk) return fail(GZ_NOT_IMPLEMENTED, "planned and set forms group by one key column");
  if (num_batches < 0 || (num_batches > 0 && (!batches || !instance_of))) return fail(GZ_INVALID, "bad batch list");
  GZ_TRY(agg_flush(h));
  return agg_autotune(h, num_batches, batches, instance_of, chosen);
}

}  // extern "C"
