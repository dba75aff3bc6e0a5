// Group-by over several key columns (the hash aggregate's multi-key form).
//
// The reference builds an UnsafeRow of the key columns for every row (a null
// bitset, then the raw bytes of each non-null key) and looks the bytes up in a
// LargeStringHashMap that hands out ids in first-seen order
// (ext/hash_aggregate_kernel.cc:970-1000, third_party/row_wise_memory/
// unsafe_row.h:45-149, precompile/unsafe_array.h:58-108).  Two rows share a
// group iff the same keys are null and the other keys have identical bits.
//
// Here a device key dictionary turns every row's key tuple into a dense int64
// code with that equality, and the single-key aggregate groups by the code:
//   key_find_kernel    one thread per row: load the key tuple (raw bits per
//                      key, null mask), hash it, probe a 64-bit slot table
//                      (fingerprint | reference); an empty slot is claimed by
//                      CAS with the row's own reference, a fingerprint match is
//                      compared against the referenced tuple (an earlier code's
//                      stored tuple, or the input row that claimed the slot in
//                      this launch).  No thread ever waits for another.
//   key_assign_kernel  the row that claimed a slot takes the next code, stores
//                      its tuple and turns the slot into that code
//   key_code_kernel    every row reads its slot's code
// The aggregate's own first-row order then gives the reference's group order,
// and the finish gathers each key column from the stored tuples by code.
//
// A utf8 key (alone, or the last of several keys: callers order the keys with
// kd_key_order) enters the tuple as a 64-bit hash of its bytes; equal hashes
// are confirmed on the bytes, and every code keeps its string in a device byte
// heap.  Alone it groups like the reference's LargeStringHashMap (the null key
// its own group); in an UnsafeRow with one utf8 key, equal row bytes mean equal
// fields wherever the utf8 key sits, so tuple equality is the row's byte
// equality.
#include <algorithm>
#include <cstring>
#include <rocprim/rocprim.hpp>
#include <vector>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

constexpr int kMaxKeys = 8;
constexpr uint32_t kPermanent = 0x80000000u;  // slot reference is a code
constexpr int64_t kClaimed = 1ll << 62;        // per-row slot: this row claimed it

struct KeyBatchPtrs {
  const uint8_t* valid[kMaxKeys];
  const uint8_t* vals[kMaxKeys];  // the utf8 key: its data buffer
  const int32_t* soffs;           // the utf8 key's offsets
};

// 256 rows of one batch: the work of one workgroup.  A row's reference in the
// slot table is tile * 256 + lane.
struct KeyTile {
  int32_t batch, row0, nrows, pad;
  int64_t g0;  // call row of row0
};

struct KeyParams {
  const KeyBatchPtrs* kb;  // per batch of the call
  const KeyTile* tiles;
  int K;
  int width[kMaxKeys];  // bytes per value: 1, 2, 4 or 8
  int sj;               // index of the utf8 key (the last one), -1 if none
  uint32_t norm;        // float/double keys normalised before hashing (bit j: key j)
  int mode;             // 0 insert every tuple, 1 insert (a null key: code -1), 2 look up (miss or null: -2)
  int64_t rows;         // rows of the call
  unsigned long long* table;
  uint64_t mask;  // table capacity - 1
  unsigned long long* kvals;  // [code * K + j] raw key bits (0 under a null)
  uint32_t* kmask;            // [code] null mask
  unsigned long long* counter;  // [0] codes, [1] heap bytes used
  int64_t* rowslot;  // per call row: its slot, then its code
  uint8_t* heap;     // strings of the codes (utf8 key)
  int64_t* koff;     // [code] heap offset of its string
  int32_t* klen;     // [code] its length
};

// A utf8 key's bytes as seen by a comparison.
struct StrRef {
  const uint8_t* p;
  int32_t n;
};

__device__ __forceinline__ uint64_t bytes_hash64(const uint8_t* p, int32_t n) {
  uint64_t h = 0x27d4eb2f165667c5ull ^ (uint64_t)n;
  uint64_t w = 0;
  int k = 0;
  for (int32_t i = 0; i < n; i++) {
    w |= (uint64_t)p[i] << (8 * k);
    if (++k == 8) {
      h = (h ^ w) * 0x9e3779b97f4a7c15ull;
      h ^= h >> 29;
      w = 0;
      k = 0;
    }
  }
  h = (h ^ w ^ ((uint64_t)k << 56)) * 0xbf58476d1ce4e5b9ull;
  return h ^ (h >> 31);
}

// Key tuple of call row g: raw bits per key (zero-extended; the utf8 key's
// hash), null mask, and the utf8 key's bytes.
__device__ __forceinline__ uint32_t load_key(const KeyParams& p, int b, int64_t r, uint64_t* raw, StrRef* str) {
  const KeyBatchPtrs& B = p.kb[b];
  uint32_t m = 0;
  str->p = nullptr;
  str->n = 0;
#pragma unroll
  for (int j = 0; j < kMaxKeys; j++) {
    raw[j] = 0;
    if (j < p.K) {
      const uint8_t* vb = B.valid[j];
      const bool ok = !vb || ((vb[r >> 3] >> (r & 7)) & 1);
      if (ok && j == p.sj) {
        const int32_t o0 = B.soffs[r];
        str->p = B.vals[j] + o0;
        str->n = B.soffs[r + 1] - o0;
        raw[j] = bytes_hash64(str->p, str->n);
      } else if (ok) {
        const uint8_t* v = B.vals[j];
        switch (p.width[j]) {
          case 1: raw[j] = v[r]; break;
          case 2: raw[j] = ((const uint16_t*)v)[r]; break;
          case 4: raw[j] = ((const uint32_t*)v)[r]; break;
          default: raw[j] = ((const uint64_t*)v)[r]; break;
        }
        if ((p.norm >> j) & 1) raw[j] = norm_key_bits(raw[j], p.width[j]);
      } else {
        m |= 1u << j;
      }
    }
  }
  return m;
}

__device__ __forceinline__ uint64_t key_hash(const uint64_t* raw, uint32_t m, int K) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ m;
#pragma unroll
  for (int j = 0; j < kMaxKeys; j++) {
    if (j < K) {
      h = (h ^ raw[j]) * 0xff51afd7ed558ccdull;
      h ^= h >> 32;
    }
  }
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  return h;
}

__device__ __forceinline__ bool key_equal(const KeyParams& p, uint32_t ref, const uint64_t* raw, uint32_t m,
                                          const StrRef& str) {
  uint64_t o[kMaxKeys];
  uint32_t om;
  StrRef os;
  if (ref & kPermanent) {
    const uint64_t c = ref & ~kPermanent;
    om = p.kmask[c];
#pragma unroll
    for (int j = 0; j < kMaxKeys; j++) o[j] = j < p.K ? p.kvals[c * p.K + j] : 0;
    os.p = p.sj >= 0 ? p.heap + p.koff[c] : nullptr;
    os.n = p.sj >= 0 ? p.klen[c] : 0;
  } else {
    const KeyTile T = p.tiles[ref >> 8];
    om = load_key(p, T.batch, (int64_t)T.row0 + (ref & 255), o, &os);
  }
  bool eq = om == m;
#pragma unroll
  for (int j = 0; j < kMaxKeys; j++)
    if (j < p.K) eq = eq && o[j] == raw[j];
  if (eq && p.sj >= 0) {
    eq = os.n == str.n;
    for (int32_t i = 0; eq && i < str.n; i++) eq = os.p[i] == str.p[i];
  }
  return eq;
}

__global__ __launch_bounds__(256) void key_find_kernel(KeyParams p) {
  const KeyTile T = p.tiles[blockIdx.x];
  if ((int)threadIdx.x >= T.nrows) return;
  const int64_t g = T.g0 + threadIdx.x;
  const uint32_t ref = blockIdx.x * 256u + threadIdx.x;
  uint64_t raw[kMaxKeys];
  StrRef str;
  const uint32_t m = load_key(p, T.batch, (int64_t)T.row0 + threadIdx.x, raw, &str);
  if (p.mode != 0 && m != 0) {  // join keys: a row with a null key never matches
    p.rowslot[g] = p.mode == 1 ? -1 : -2;
    return;
  }
  const uint64_t h = key_hash(raw, m, p.K);
  const uint64_t fp = (h >> 32) | 1u;  // never zero: an occupied slot is never 0
  const unsigned long long mine = (fp << 32) | (uint64_t)ref;
  uint64_t s = h & p.mask;
  unsigned long long w;
  // the table is at most half full, so the walk ends
  while (true) {
    w = __hip_atomic_load(&p.table[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (w == 0 && p.mode == 2) {  // look-up: an empty slot ends the walk
      p.rowslot[g] = -2;
      return;
    }
    if (w == 0) {
      const unsigned long long old = atomicCAS(&p.table[s], 0ull, mine);
      if (old == 0) {
        p.rowslot[g] = (int64_t)s | kClaimed;
        return;
      }
      w = old;
    }
    if ((w >> 32) == fp && key_equal(p, (uint32_t)w, raw, m, str)) break;
    s = (s + 1) & p.mask;
  }
  // a look-up only meets finished codes: the code itself, no code pass
  p.rowslot[g] = p.mode == 2 ? (int64_t)(w & 0x7fffffffull) : (int64_t)s;
}

__global__ __launch_bounds__(256) void key_assign_kernel(KeyParams p) {
  // codes and heap bytes are reserved once per workgroup (LDS counts, one
  // device atomic each), not once per claiming row on a single word
  __shared__ unsigned int s_n;
  __shared__ unsigned long long s_bytes, s_code0, s_heap0;
  const KeyTile T = p.tiles[blockIdx.x];
  const int64_t g = T.g0 + threadIdx.x;
  bool claimer = false;
  int64_t s = 0;
  if ((int)threadIdx.x < T.nrows) {
    const int64_t rs = p.rowslot[g];
    // only the rows that claimed a slot (marked by the walk) take codes
    claimer = rs >= 0 && (rs & kClaimed);
    s = rs & ~kClaimed;
  }
  if (threadIdx.x == 0) {
    s_n = 0;
    s_bytes = 0;
  }
  __syncthreads();
  uint64_t raw[kMaxKeys];
  StrRef str{nullptr, 0};
  uint32_t m = 0;
  unsigned my = 0;
  unsigned long long my_off = 0;
  if (claimer) {
    m = load_key(p, T.batch, (int64_t)T.row0 + threadIdx.x, raw, &str);
    my = atomicAdd(&s_n, 1u);
    if (p.sj >= 0 && str.n > 0) my_off = atomicAdd(&s_bytes, (unsigned long long)str.n);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    s_code0 = s_n ? atomicAdd(p.counter, (unsigned long long)s_n) : 0ull;
    s_heap0 = s_bytes ? atomicAdd(p.counter + 1, s_bytes) : 0ull;
  }
  __syncthreads();
  if (!claimer) return;
  const unsigned long long w = p.table[s];
  const unsigned long long c = s_code0 + my;
#pragma unroll
  for (int j = 0; j < kMaxKeys; j++)
    if (j < p.K) p.kvals[c * p.K + j] = raw[j];
  p.kmask[c] = m;
  if (p.sj >= 0) {
    const unsigned long long at = str.n > 0 ? s_heap0 + my_off : 0ull;
    for (int32_t i = 0; i < str.n; i++) p.heap[at + i] = str.p[i];
    p.koff[c] = (int64_t)at;
    p.klen[c] = str.n;
  }
  p.table[s] = (w & 0xffffffff00000000ull) | kPermanent | (uint32_t)c;
}

__global__ __launch_bounds__(256) void key_code_kernel(const unsigned long long* table, int64_t* rowslot, int64_t rows) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= rows) return;
  const int64_t s = rowslot[g];
  if (s >= 0) rowslot[g] = (int64_t)(table[s & ~kClaimed] & 0x7fffffffull);
}

// Re-inserts codes [0, n) into a larger empty table.
__global__ __launch_bounds__(256) void key_rehash_kernel(unsigned long long* table, uint64_t mask,
                                                         const unsigned long long* kvals, const uint32_t* kmask,
                                                         int K, int64_t n) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  uint64_t raw[kMaxKeys];
#pragma unroll
  for (int j = 0; j < kMaxKeys; j++) raw[j] = j < K ? kvals[c * K + j] : 0;
  const uint64_t h = key_hash(raw, kmask[c], K);
  const unsigned long long word = (((h >> 32) | 1u) << 32) | kPermanent | (uint32_t)c;
  uint64_t s = h & mask;
  while (atomicCAS(&table[s], 0ull, word) != 0ull) s = (s + 1) & mask;
}

struct KeyGatherParams {
  const int64_t* codes;  // per group
  int64_t n;
  int K;
  int sj;  // the utf8 key (validity only here), -1 if none
  int width[kMaxKeys];
  const unsigned long long* kvals;
  const uint32_t* kmask;
  void* out[kMaxKeys];
  uint8_t* valid[kMaxKeys];
};

__global__ __launch_bounds__(256) void key_gather_kernel(KeyGatherParams f) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < f.n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = f.codes[i];
    const uint32_t m = f.kmask[c];
    for (int j = 0; j < f.K; j++) {
      f.valid[j][i] = ((m >> j) & 1) ? 0 : 1;
      if (j == f.sj) continue;
      const uint64_t v = f.kvals[c * f.K + j];
      switch (f.width[j]) {
        case 1: ((uint8_t*)f.out[j])[i] = (uint8_t)v; break;
        case 2: ((uint16_t*)f.out[j])[i] = (uint16_t)v; break;
        case 4: ((uint32_t*)f.out[j])[i] = (uint32_t)v; break;
        default: ((uint64_t*)f.out[j])[i] = v; break;
      }
    }
  }
}

// utf8 key result: lengths per group (n + 1 entries, the last 0), scanned into
// offsets, then each group's bytes copied from the heap.
__global__ __launch_bounds__(256) void key_str_len_kernel(const int64_t* codes, int64_t n, const int32_t* klen,
                                                          int32_t* lens) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lens[i] = klen[codes[i]];
  else if (i == n) lens[n] = 0;
}

__global__ __launch_bounds__(256) void key_str_copy_kernel(const int64_t* codes, int64_t n, const int64_t* koff,
                                                           const int32_t* offs, const uint8_t* heap, uint8_t* data) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t* src = heap + koff[codes[i]];
  const int32_t o = offs[i], len = offs[i + 1] - o;
  for (int32_t b = 0; b < len; b++) data[o + b] = src[b];
}

__global__ void fill_u64_kernel_keys(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ------------------------------------------------------------------ host side

// The key dictionary: slot table, per-code key tuples, utf8 key heap.
struct KeyDict {
  gz_context* ctx = nullptr;
  int K = 0;
  std::vector<int> key_types;
  int sj = -1;  // the utf8 key (last), -1 if none
  uint32_t norm = 0;  // normalised float/double keys
  DBuf table, kvals, kmask, counter;  // counter: [0] codes, [1] heap bytes used
  uint64_t tcap = 0;
  int64_t kcap = 0, ncodes = 0;
  DBuf heap, koff, klen;
  int64_t heap_cap = 0, heap_used = 0;
};

struct MultiKey {
  gz_context* ctx = nullptr;
  gz_agg* inner = nullptr;  // single-key aggregate over (columns..., code)
  int ncols = 0;
  std::vector<int> types;
  std::vector<int> key_cols;
  std::vector<int> res_map;  // outer result -> inner result (>= 0) or key -(j + 1)
  int code_res = 0;          // inner result holding the code
  KeyDict* kd = nullptr;
  KeyScratch sc;
  // finish
  bool finished = false;
  int64_t ngroups = 0;
  std::vector<DBuf> out_vals, out_valid;  // utf8 key: out_vals = offsets
  DBuf str_data, str_tmp;
};

static bool key_type_ok(int t) {
  switch (t) {
    case GZ_TYPE_INT8: case GZ_TYPE_UINT8: case GZ_TYPE_INT16: case GZ_TYPE_UINT16: case GZ_TYPE_INT32:
    case GZ_TYPE_UINT32: case GZ_TYPE_INT64: case GZ_TYPE_UINT64: case GZ_TYPE_FLOAT: case GZ_TYPE_DOUBLE:
    case GZ_TYPE_DATE32: case GZ_TYPE_DATE64: case GZ_TYPE_TIMESTAMP:
      return true;
    default:
      return false;
  }
}

std::vector<int> kd_key_order(int K, const int32_t* key_types) {
  std::vector<int> perm;
  int nstr = 0, sj = -1;
  for (int j = 0; j < K; j++)
    if (key_types[j] == GZ_TYPE_STRING) {
      nstr++;
      sj = j;
    }
  for (int j = 0; j < K; j++)
    if (nstr != 1 || j != sj) perm.push_back(j);
  if (nstr == 1) perm.push_back(sj);
  return perm;
}

int kd_make(gz_context* ctx, int32_t K, const int32_t* key_types, KeyDict** out) {
  if (K < 1 || K > kMaxKeys) return fail(GZ_NOT_IMPLEMENTED, "1 to 8 key columns");
  auto* d = new KeyDict();
  d->ctx = ctx;
  d->K = K;
  for (int j = 0; j < K; j++) {
    const int t = key_types[j];
    if (d->sj >= 0) {
      delete d;
      return fail(GZ_NOT_IMPLEMENTED, "a key column after a utf8 key");
    }
    if (t == GZ_TYPE_STRING) d->sj = j;
    else if (!key_type_ok(t)) {
      delete d;
      return fail(GZ_NOT_IMPLEMENTED, "key type " + std::to_string(t) +
                                          " (fixed-width numeric keys and one utf8 key, last)");
    }
    d->key_types.push_back(t);
  }
  int rc = ctx_alloc(ctx, 2 * sizeof(unsigned long long), &d->counter);
  if (!rc && hipMemsetAsync(d->counter.p, 0, 2 * sizeof(unsigned long long), ctx->stream) != hipSuccess)
    rc = fail(GZ_HIP_ERROR, "clearing the key counter");
  if (rc) {
    ctx_free(ctx, &d->counter);
    delete d;
    return rc;
  }
  *out = d;
  return GZ_OK;
}

int kd_set_normalize(KeyDict* d, uint32_t mask) {
  if (d->ncodes > 0) return fail(GZ_INVALID, "key normalisation is set before the first insert");
  for (int j = 0; j < 32; j++) {
    if (!((mask >> j) & 1)) continue;
    if (j >= d->K || (d->key_types[j] != GZ_TYPE_FLOAT && d->key_types[j] != GZ_TYPE_DOUBLE))
      return fail(GZ_TYPE_ERROR, "only float and double keys are normalised (key " + std::to_string(j) + ")");
  }
  d->norm = mask;
  return GZ_OK;
}

void kd_release(KeyDict* d) {
  if (!d) return;
  for (DBuf* b : {&d->table, &d->kvals, &d->kmask, &d->counter, &d->heap, &d->koff, &d->klen}) ctx_free(d->ctx, b);
  delete d;
}

void kd_scratch_free(gz_context* ctx, KeyScratch* sc) {
  for (DBuf* b : {&sc->rowslot, &sc->d_kb, &sc->d_prefix}) ctx_free(ctx, b);
}

int mk_make(gz_context* ctx, int32_t ncols, const int32_t* types, int32_t nact, const gz_agg_action* acts,
            const gz_agg_options* opts, MultiKey** out) {
  auto* m = new MultiKey();
  m->ctx = ctx;
  m->ncols = ncols;
  m->types.assign(types, types + ncols);
  std::vector<gz_agg_action> inner;
  int inner_pos = 0;
  auto bad = [&](int code, const std::string& msg) {
    delete m;
    return fail(code, msg);
  };
  for (int i = 0; i < nact; i++) {
    const gz_agg_action& a = acts[i];
    if (a.kind == GZ_ACTION_GROUPBY) {
      if (a.input0 < 0 || a.input0 >= ncols) return bad(GZ_INVALID, "groupby input out of range");
      if ((int)m->key_cols.size() == kMaxKeys) return bad(GZ_NOT_IMPLEMENTED, "more than 8 group-by keys");
      m->res_map.push_back(-((int)m->key_cols.size() + 1));
      if (m->key_cols.empty()) {
        inner.push_back({GZ_ACTION_GROUPBY, ncols, -1, 0});
        m->code_res = inner_pos++;
      }
      m->key_cols.push_back(a.input0);
    } else {
      inner.push_back(a);
      const int nres = a.kind == GZ_ACTION_SUM_COUNT ? 2 : 1;
      for (int r = 0; r < nres; r++) m->res_map.push_back(inner_pos++);
    }
  }
  // the dictionary's key order: one utf8 key goes last; result key columns
  // keep the plan's order
  std::vector<int32_t> ktypes;
  for (int c : m->key_cols) ktypes.push_back(types[c]);
  const std::vector<int> perm = kd_key_order((int)ktypes.size(), ktypes.data());
  std::vector<int> pos(perm.size());  // caller key j -> dictionary position
  std::vector<int> kc;
  ktypes.clear();
  for (size_t i = 0; i < perm.size(); i++) {
    pos[perm[i]] = (int)i;
    kc.push_back(m->key_cols[perm[i]]);
    ktypes.push_back(types[kc.back()]);
  }
  m->key_cols = kc;
  for (int& r : m->res_map)
    if (r < 0) r = -(pos[-r - 1] + 1);
  int rc = kd_make(ctx, (int32_t)ktypes.size(), ktypes.data(), &m->kd);
  if (rc) {
    delete m;
    return rc;
  }
  std::vector<int32_t> itypes(types, types + ncols);
  itypes.push_back(GZ_TYPE_INT64);
  rc = gz_agg_make(ctx, ncols + 1, itypes.data(), (int32_t)inner.size(), inner.data(), opts, &m->inner);
  if (rc) {
    kd_release(m->kd);
    delete m;
    return rc;
  }
  *out = m;
  return GZ_OK;
}

static int key_width(int t) { return t == GZ_TYPE_STRING ? 0 : type_width(t); }

// Grows a device buffer to `bytes`, keeping its first `keep` bytes.
static int grow_keep(gz_context* ctx, DBuf* b, size_t bytes, size_t keep) {
  DBuf nb;
  GZ_TRY(ctx_alloc(ctx, bytes, &nb));
  if (keep > 0) GZ_HIP(hipMemcpyAsync(nb.p, b->p, keep, hipMemcpyDeviceToDevice, ctx->stream));
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  ctx_free(ctx, b);
  *b = nb;
  return GZ_OK;
}

// Grows the code store to hold `need` codes, the utf8 heap to `heap_need`
// bytes, and the table to at least twice `need` slots, re-inserting the codes
// already assigned.
static int kd_reserve(KeyDict* d, int64_t need, int64_t heap_need) {
  gz_context* ctx = d->ctx;
  if (need > d->kcap) {
    const int64_t cap = std::max<int64_t>({need, 2 * d->kcap, 1024});
    const size_t n = (size_t)d->ncodes;
    GZ_TRY(grow_keep(ctx, &d->kvals, sizeof(unsigned long long) * (size_t)cap * d->K,
                     sizeof(unsigned long long) * n * d->K));
    GZ_TRY(grow_keep(ctx, &d->kmask, sizeof(uint32_t) * (size_t)cap, sizeof(uint32_t) * n));
    if (d->sj >= 0) {
      GZ_TRY(grow_keep(ctx, &d->koff, sizeof(int64_t) * (size_t)cap, sizeof(int64_t) * n));
      GZ_TRY(grow_keep(ctx, &d->klen, sizeof(int32_t) * (size_t)cap, sizeof(int32_t) * n));
    }
    d->kcap = cap;
  }
  if (d->sj >= 0 && (heap_need > d->heap_cap || !d->heap.p)) {
    const int64_t cap = std::max<int64_t>({heap_need, 2 * d->heap_cap, 4096});
    GZ_TRY(grow_keep(ctx, &d->heap, (size_t)cap, (size_t)d->heap_used));
    d->heap_cap = cap;
  }
  uint64_t tcap = 1024;
  while (tcap < 2 * (uint64_t)need) tcap <<= 1;
  if (tcap > d->tcap) {
    DBuf nt;
    GZ_TRY(ctx_alloc(ctx, sizeof(unsigned long long) * tcap, &nt));
    const unsigned fill_blocks = (unsigned)std::min<uint64_t>((tcap + 255) / 256, 65536);
    hipLaunchKernelGGL(fill_u64_kernel_keys, dim3(fill_blocks), dim3(256), 0, ctx->stream,
                       (unsigned long long*)nt.p, (int64_t)tcap);
    if (d->ncodes > 0) {
      KernelTimer kt(ctx, "key_rehash");
      hipLaunchKernelGGL(key_rehash_kernel, dim3((unsigned)((d->ncodes + 255) / 256)), dim3(256), 0, ctx->stream,
                         (unsigned long long*)nt.p, tcap - 1, (const unsigned long long*)d->kvals.p,
                         (const uint32_t*)d->kmask.p, d->K, d->ncodes);
    }
    GZ_HIP(hipGetLastError());
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, &d->table);
    d->table = nt;
    d->tcap = tcap;
  }
  return GZ_OK;
}

int kd_encode(KeyDict* d, gz_context* ctx, KeyScratch* sc, int32_t ncols, const int32_t* types,
              const int32_t* key_cols, int32_t nb, const gz_batch* batches, int mode, const int64_t** codes,
              int64_t* rows_out) {
  *rows_out = 0;
  if (mode == 2 && !d->table.p) {
    // nothing was inserted: every probe row misses
    int64_t rows = 0;
    for (int b = 0; b < nb; b++) rows += std::max<int64_t>(batches[b].num_rows, 0);
    if (rows == 0) return GZ_OK;
    GZ_TRY(ctx_ensure(ctx, &sc->rowslot, sizeof(int64_t) * (size_t)rows));
    std::vector<int64_t> miss((size_t)rows, -2);
    GZ_HIP(hipMemcpyAsync(sc->rowslot.p, miss.data(), sizeof(int64_t) * (size_t)rows, hipMemcpyHostToDevice,
                          ctx->stream));
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    *codes = (const int64_t*)sc->rowslot.p;
    *rows_out = rows;
    return GZ_OK;
  }
  int need_bufs = 0;
  std::vector<int> col_buf(ncols);
  for (int c = 0; c < ncols; c++) {
    col_buf[c] = need_bufs;
    need_bufs += type_num_bufs(types[c]);
  }
  std::vector<KeyBatchPtrs> kb(std::max(nb, 1));
  std::vector<int64_t> prefix(nb + 1, 0);
  int64_t heap_need = d->heap_used;
  // a utf8 key's bytes per batch: both ends of its offsets, read once for all
  // batches, bound the copies into the heap and must fit the declared data size
  std::vector<int32_t> ends;
  if (d->sj >= 0 && mode != 2) {
    ends.assign((size_t)nb * 2, 0);
    const int k = col_buf[key_cols[d->sj]];
    for (int b = 0; b < nb; b++) {
      const gz_batch& bt = batches[b];
      if (bt.num_rows <= 0 || !bt.buf_addrs || !bt.buf_sizes || bt.num_bufs <= k + 1 || !bt.buf_addrs[k + 1] ||
          bt.buf_sizes[k + 1] < (bt.num_rows + 1) * 4)
        continue;  // rejected below
      const int32_t* off = (const int32_t*)bt.buf_addrs[k + 1];
      GZ_HIP(hipMemcpyAsync(&ends[2 * b], off, 4, hipMemcpyDeviceToHost, ctx->stream));
      GZ_HIP(hipMemcpyAsync(&ends[2 * b + 1], off + bt.num_rows, 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    GZ_HIP(hipStreamSynchronize(ctx->stream));
  }
  for (int b = 0; b < nb; b++) {
    const gz_batch& bt = batches[b];
    if (bt.num_rows < 0 || bt.num_rows > INT32_MAX) return fail(GZ_INVALID, "bad num_rows");
    if (bt.num_rows > 0 && (!bt.buf_addrs || !bt.buf_sizes)) return fail(GZ_INVALID, "buf_addrs can't be null");
    if (bt.num_rows > 0 && bt.num_bufs < need_bufs) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
    memset(&kb[b], 0, sizeof(KeyBatchPtrs));
    if (bt.num_rows > 0) {
      for (int j = 0; j < d->K; j++) {
        const int k = col_buf[key_cols[j]];
        const bool has_valid = bt.buf_addrs[k] != 0 && bt.buf_sizes[k] > 0;
        if (has_valid && bt.buf_sizes[k] < (bt.num_rows + 7) / 8) return fail(GZ_INVALID, "validity buffer too small");
        kb[b].valid[j] = has_valid ? (const uint8_t*)bt.buf_addrs[k] : nullptr;
        if (j == d->sj) {
          if (!bt.buf_addrs[k + 1] || bt.buf_sizes[k + 1] < (bt.num_rows + 1) * 4)
            return fail(GZ_INVALID, "offsets buffer too small");
          if (bt.buf_sizes[k + 2] > 0 && !bt.buf_addrs[k + 2]) return fail(GZ_INVALID, "missing values buffer");
          kb[b].soffs = (const int32_t*)bt.buf_addrs[k + 1];
          kb[b].vals[j] = (const uint8_t*)bt.buf_addrs[k + 2];
          if (!ends.empty()) {
            const int64_t lo = ends[2 * b], hi = ends[2 * b + 1];
            if (lo < 0 || hi < lo || hi > bt.buf_sizes[k + 2])
              return fail(GZ_INVALID, "utf8 key offsets run past the declared data buffer (" + std::to_string(lo) +
                                          ".." + std::to_string(hi) + " of " + std::to_string(bt.buf_sizes[k + 2]) +
                                          " bytes)");
            heap_need += hi - lo;
          }
          continue;
        }
        if (!bt.buf_addrs[k + 1] || bt.buf_sizes[k + 1] < bt.num_rows * key_width(d->key_types[j]))
          return fail(GZ_INVALID, "values buffer too small");
        kb[b].vals[j] = (const uint8_t*)bt.buf_addrs[k + 1];
      }
    }
    prefix[b + 1] = prefix[b] + bt.num_rows;
  }
  const int64_t rows = prefix[nb];
  if (rows == 0) return GZ_OK;
  if (rows > INT32_MAX) return fail(GZ_NOT_IMPLEMENTED, "2^31 or more rows in one multi-key call");
  if (mode != 2) {
    if (d->ncodes + rows > INT32_MAX) return fail(GZ_CAPACITY_ERROR, "2^31 or more distinct keys");
    GZ_TRY(kd_reserve(d, d->ncodes + rows, heap_need));
  }
  GZ_TRY(upload(ctx, &sc->d_kb, kb.data(), sizeof(KeyBatchPtrs) * kb.size()));
  std::vector<KeyTile> tiles;
  tiles.reserve((size_t)(rows / 256 + nb + 1));
  for (int b = 0; b < nb; b++)
    for (int64_t r0 = 0; r0 < batches[b].num_rows; r0 += 256)
      tiles.push_back(KeyTile{b, (int32_t)r0, (int32_t)std::min<int64_t>(256, batches[b].num_rows - r0), 0,
                              prefix[b] + r0});
  if (tiles.size() >= (1u << 23)) return fail(GZ_NOT_IMPLEMENTED, "2^23 or more row tiles in one multi-key call");
  GZ_TRY(upload(ctx, &sc->d_prefix, tiles.data(), sizeof(KeyTile) * tiles.size()));
  const unsigned ntiles = (unsigned)tiles.size();
  GZ_TRY(ctx_ensure(ctx, &sc->rowslot, sizeof(int64_t) * (size_t)rows));
  KeyParams p;
  memset(&p, 0, sizeof(p));
  p.kb = (const KeyBatchPtrs*)sc->d_kb.p;
  p.tiles = (const KeyTile*)sc->d_prefix.p;
  p.K = d->K;
  for (int j = 0; j < d->K; j++) p.width[j] = key_width(d->key_types[j]);
  p.sj = d->sj;
  p.norm = d->norm;
  p.mode = mode;
  p.rows = rows;
  p.table = (unsigned long long*)d->table.p;
  p.mask = d->tcap - 1;
  p.kvals = (unsigned long long*)d->kvals.p;
  p.kmask = (uint32_t*)d->kmask.p;
  p.counter = (unsigned long long*)d->counter.p;
  p.rowslot = (int64_t*)sc->rowslot.p;
  p.heap = (uint8_t*)d->heap.p;
  p.koff = (int64_t*)d->koff.p;
  p.klen = (int32_t*)d->klen.p;
  const unsigned grid = (unsigned)((rows + 255) / 256);
  {
    KernelTimer kt(ctx, "key_find");
    hipLaunchKernelGGL(key_find_kernel, dim3(ntiles), dim3(256), 0, ctx->stream, p);
  }
  {
    KernelTimer kt(ctx, "key_assign");
    if (mode != 2) {
      hipLaunchKernelGGL(key_assign_kernel, dim3(ntiles), dim3(256), 0, ctx->stream, p);
      hipLaunchKernelGGL(key_code_kernel, dim3(grid), dim3(256), 0, ctx->stream,
                         (const unsigned long long*)d->table.p, (int64_t*)sc->rowslot.p, rows);
    }
  }
  GZ_HIP(hipGetLastError());
  if (mode != 2) {
    unsigned long long nc[2] = {0, 0};
    GZ_HIP(hipMemcpyAsync(nc, d->counter.p, sizeof(nc), hipMemcpyDeviceToHost, ctx->stream));
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    d->ncodes = (int64_t)nc[0];
    d->heap_used = (int64_t)nc[1];
  }
  *codes = (const int64_t*)sc->rowslot.p;
  *rows_out = rows;
  return GZ_OK;
}

// Each batch's buffer list with an int64 column of codes (no validity) appended.
void with_code_column(int32_t ncols, const int32_t* types, int32_t nb, const gz_batch* batches, const int64_t* codes,
                      std::vector<int64_t>& addrs, std::vector<int64_t>& sizes, std::vector<gz_batch>& out,
                      int valid_col) {
  int need_bufs = 0, vbuf = -1;
  for (int c = 0; c < ncols; c++) {
    if (c == valid_col) vbuf = need_bufs;
    need_bufs += type_num_bufs(types[c]);
  }
  addrs.assign((size_t)nb * (need_bufs + 2), 0);
  sizes.assign(addrs.size(), 0);
  out.resize(nb);
  int64_t off = 0;
  for (int b = 0; b < nb; b++) {
    const gz_batch& bt = batches[b];
    int64_t* a = &addrs[(size_t)b * (need_bufs + 2)];
    int64_t* s = &sizes[(size_t)b * (need_bufs + 2)];
    for (int k = 0; k < need_bufs; k++) {
      a[k] = bt.num_rows > 0 ? bt.buf_addrs[k] : 0;
      s[k] = bt.num_rows > 0 ? bt.buf_sizes[k] : 0;
    }
    if (vbuf >= 0 && bt.num_rows > 0) {
      a[need_bufs] = bt.buf_addrs[vbuf];
      s[need_bufs] = bt.buf_sizes[vbuf];
    }
    a[need_bufs + 1] = bt.num_rows > 0 ? (int64_t)(codes + off) : 0;
    s[need_bufs + 1] = bt.num_rows * (int64_t)sizeof(int64_t);
    out[b] = gz_batch{bt.num_rows, a, s, need_bufs + 2};
    off += std::max<int64_t>(bt.num_rows, 0);
  }
}

int mk_process(MultiKey* m, int32_t nb, const gz_batch* batches) {
  gz_context* ctx = m->ctx;
  if (m->finished) return fail(GZ_INVALID, "aggregate already finished");
  GZ_HIP(hipSetDevice(ctx->device));
  const int64_t* codes = nullptr;
  int64_t rows = 0;
  GZ_TRY(kd_encode(m->kd, ctx, &m->sc, m->ncols, m->types.data(), m->key_cols.data(), nb, batches, 0, &codes,
                   &rows));
  if (rows == 0) return GZ_OK;  // ProcessAndCacheOne returns early on empty input
  std::vector<int64_t> addrs, sizes;
  std::vector<gz_batch> inner;
  with_code_column(m->ncols, m->types.data(), nb, batches, codes, addrs, sizes, inner);
  return gz_agg_process_batches(m->inner, nb, inner.data());
}

int mk_num_result_columns(MultiKey* m) { return (int)m->res_map.size(); }

int mk_result_type(MultiKey* m, int col) {
  if (col < 0 || col >= (int)m->res_map.size()) return -1;
  const int r = m->res_map[col];
  return r < 0 ? m->kd->key_types[-r - 1] : gz_agg_result_type(m->inner, r);
}

bool mk_finished(const MultiKey* m) { return m->finished; }

int mk_finish(MultiKey* m) {
  KeyDict* d = m->kd;
  if (m->finished) return GZ_OK;
  gz_context* ctx = m->ctx;
  GZ_HIP(hipSetDevice(ctx->device));
  GZ_TRY(gz_agg_finish(m->inner));
  GZ_TRY(gz_agg_num_groups(m->inner, &m->ngroups));
  void* codes = nullptr;
  GZ_TRY(gz_agg_result_column(m->inner, m->code_res, &codes, nullptr));
  for (auto& b : m->out_vals) ctx_free(ctx, &b);
  for (auto& b : m->out_valid) ctx_free(ctx, &b);
  m->out_vals.assign(d->K, DBuf());
  m->out_valid.assign(d->K, DBuf());
  KeyGatherParams f;
  memset(&f, 0, sizeof(f));
  f.codes = (const int64_t*)codes;
  f.n = m->ngroups;
  f.K = d->K;
  f.kvals = (const unsigned long long*)d->kvals.p;
  f.kmask = (const uint32_t*)d->kmask.p;
  f.sj = d->sj;
  for (int j = 0; j < d->K; j++) {
    f.width[j] = key_width(d->key_types[j]);
    // the utf8 key's values are its int32 offsets (groups + 1)
    const size_t vbytes = j == d->sj ? sizeof(int32_t) * (size_t)(m->ngroups + 1)
                                     : (size_t)std::max<int64_t>(m->ngroups, 1) * f.width[j];
    GZ_TRY(ctx_alloc(ctx, vbytes, &m->out_vals[j]));
    GZ_TRY(ctx_alloc(ctx, (size_t)std::max<int64_t>(m->ngroups, 1), &m->out_valid[j]));
    f.out[j] = m->out_vals[j].p;
    f.valid[j] = (uint8_t*)m->out_valid[j].p;
  }
  if (m->ngroups > 0) {
    KernelTimer kt(ctx, "key_gather");
    const unsigned gx = (unsigned)std::min<int64_t>((m->ngroups + 255) / 256, 8192);
    hipLaunchKernelGGL(key_gather_kernel, dim3(gx), dim3(256), 0, ctx->stream, f);
  }
  if (d->sj >= 0) {
    if (d->heap_used > INT32_MAX) return fail(GZ_CAPACITY_ERROR, "utf8 keys of 2^31 bytes or more");
    ctx_free(ctx, &m->str_data);
    GZ_TRY(ctx_alloc(ctx, (size_t)std::max<int64_t>(d->heap_used, 1), &m->str_data));
    DBuf lens;
    GZ_TRY(ctx_alloc(ctx, sizeof(int32_t) * (size_t)(m->ngroups + 1), &lens));
    int32_t* offs = (int32_t*)m->out_vals[d->sj].p;
    const unsigned g1 = (unsigned)((m->ngroups + 1 + 255) / 256);
    hipLaunchKernelGGL(key_str_len_kernel, dim3(g1), dim3(256), 0, ctx->stream, (const int64_t*)codes, m->ngroups,
                       (const int32_t*)d->klen.p, (int32_t*)lens.p);
    size_t tmp = 0;
    GZ_HIP(rocprim::exclusive_scan(nullptr, tmp, (const int32_t*)lens.p, offs, 0, (size_t)(m->ngroups + 1),
                                   rocprim::plus<int32_t>(), ctx->stream));
    GZ_TRY(ctx_ensure(ctx, &m->str_tmp, std::max<size_t>(tmp, 1)));
    GZ_HIP(rocprim::exclusive_scan(m->str_tmp.p, tmp, (const int32_t*)lens.p, offs, 0, (size_t)(m->ngroups + 1),
                                   rocprim::plus<int32_t>(), ctx->stream));
    if (m->ngroups > 0)
      hipLaunchKernelGGL(key_str_copy_kernel, dim3((unsigned)((m->ngroups + 255) / 256)), dim3(256), 0, ctx->stream,
                         (const int64_t*)codes, m->ngroups, (const int64_t*)d->koff.p, (const int32_t*)offs,
                         (const uint8_t*)d->heap.p, (uint8_t*)m->str_data.p);
    GZ_HIP(hipGetLastError());
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, &lens);
  }
  GZ_HIP(hipGetLastError());
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  m->finished = true;
  return GZ_OK;
}

int mk_num_groups(MultiKey* m, int64_t* out) {
  GZ_TRY(mk_finish(m));
  *out = m->ngroups;
  return GZ_OK;
}

int mk_result_column(MultiKey* m, int col, void** values, void** valid) {
  if (!m->finished) return fail(GZ_INVALID, "call gz_agg_finish first");
  if (col < 0 || col >= (int)m->res_map.size()) return fail(GZ_INDEX_ERROR, "result column out of range");
  const int r = m->res_map[col];
  if (r >= 0) return gz_agg_result_column(m->inner, r, values, valid);
  if (values) *values = m->out_vals[-r - 1].p;
  if (valid) *valid = m->out_valid[-r - 1].p;
  return GZ_OK;
}

int mk_result_binary(MultiKey* m, int col, void** offsets, void** data, void** valid) {
  if (!m->finished) return fail(GZ_INVALID, "call gz_agg_finish first");
  if (col < 0 || col >= (int)m->res_map.size()) return fail(GZ_INDEX_ERROR, "result column out of range");
  const int r = m->res_map[col];
  if (r >= 0 || -r - 1 != m->kd->sj) return fail(GZ_TYPE_ERROR, "result column is not utf8");
  if (offsets) *offsets = m->out_vals[m->kd->sj].p;
  if (data) *data = m->str_data.p;
  if (valid) *valid = m->out_valid[m->kd->sj].p;
  return GZ_OK;
}

void mk_release(MultiKey* m) {
  if (!m) return;
  gz_context* ctx = m->ctx;
  gz_agg_release(m->inner);
  kd_release(m->kd);
  kd_scratch_free(ctx, &m->sc);
  for (auto& b : m->out_vals) ctx_free(ctx, &b);
  for (auto& b : m->out_valid) ctx_free(ctx, &b);
  for (DBuf* b : {&m->str_data, &m->str_tmp}) ctx_free(ctx, b);
  delete m;
}

}  // namespace gz
