/*
 * agg.c -- CPU restatement of the reference hash aggregate
 * (ext/hash_aggregate_kernel.cc:778-928 with the accumulators of
 * ext/actions_impl.cc).  TEST INFRASTRUCTURE ONLY (checker + CPU baseline).
 *
 * Group ids are dense and assigned in first-seen order, as SparseHashMap<T>
 * assigns them (third_party/sparsehash/sparse_hash_map.h:39-47); the null key
 * gets its own id when first seen (:49-58).  The map below is an
 * open-addressing table; only the id order is observable, and it does not
 * depend on the map implementation.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gz_oracle.h"

enum { T_NA = 0, T_BOOL = 1, T_STRING = 13, T_BINARY = 14, T_UINT8 = 2, T_INT8 = 3, T_UINT16 = 4, T_INT16 = 5,
       T_UINT32 = 6, T_INT32 = 7, T_UINT64 = 8, T_INT64 = 9, T_FLOAT = 11, T_DOUBLE = 12,
       T_DATE32 = 16, T_DATE64 = 17, T_TIMESTAMP = 18 };
enum { A_GROUPBY = 0, A_COUNT = 1, A_COUNT_LITERAL = 2, A_SUM = 3, A_AVG = 4, A_MIN = 5,
       A_MAX = 6, A_SUM_COUNT = 7, A_AVG_BY_COUNT = 8 };

static int is_float(int32_t t) { return t == T_FLOAT || t == T_DOUBLE; }
static int is_unsigned(int32_t t) { return t == T_UINT8 || t == T_UINT16 || t == T_UINT32 || t == T_UINT64; }
static int width_of(int32_t t) {
  switch (t) {
    case T_UINT8: case T_INT8: return 1;
    case T_UINT16: case T_INT16: return 2;
    case T_UINT32: case T_INT32: case T_FLOAT: case T_DATE32: return 4;
    default: return 8;
  }
}

static inline int bit_get(const uint8_t* bm, int64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }

/* integer value widened to int64 (unsigned types zero-extended) */
static int64_t get_i64(const void* v, int32_t t, int64_t r) {
  switch (t) {
    case T_INT8: return ((const int8_t*)v)[r];
    case T_UINT8: return ((const uint8_t*)v)[r];
    case T_INT16: return ((const int16_t*)v)[r];
    case T_UINT16: return ((const uint16_t*)v)[r];
    case T_INT32: case T_DATE32: return ((const int32_t*)v)[r];
    case T_UINT32: return ((const uint32_t*)v)[r];
    default: return ((const int64_t*)v)[r];
  }
}
static double get_f64(const void* v, int32_t t, int64_t r) {
  if (t == T_FLOAT) return ((const float*)v)[r];
  if (t == T_DOUBLE) return ((const double*)v)[r];
  if (t == T_UINT64) return (double)((const uint64_t*)v)[r];
  return (double)get_i64(v, t, r);
}

/* a float or double value's bits as stored (a float's 32 bits zero-extended) */
static int64_t raw_bits(const void* v, int32_t t, int64_t r) {
  if (t == T_FLOAT) return (int64_t)((const uint32_t*)v)[r];
  return ((const int64_t*)v)[r];
}

typedef struct {
  int32_t kind, in0, in1, nan_check;
  int32_t in_type;
  /* per-group state, grown with the group count (GrowByFactor) */
  double* dsum;    /* sum / avg / sum_count sum (floating) */
  int64_t* isum;   /* integer sum, count, countLiteral */
  int64_t* cnt;    /* avg / sum_count / avgByCount counts */
  int64_t* imm;    /* integer min/max */
  double* fmm;     /* floating min/max (value, for comparisons) */
  int64_t* fraw;   /* floating min/max: the cached value's bits as stored (a
                    * float's 32 bits zero-extended), so NaN payloads and the
                    * sign of a zero come out as the reference keeps them */
  uint8_t* valid;  /* cache_validity_ */
  uint8_t* nullf;  /* UniqueAction null_flag_ */
  int64_t* key;    /* UniqueAction cache_ (bit pattern) */
  int64_t* soff;   /* UniqueAction cache_ of a utf8 key: bytes in sarena */
  int32_t* slen;
  uint8_t* sarena;
  int64_t sused, scap_bytes;
} action_t;

struct gzo_agg {
  int ncols;
  int32_t* types;
  int nactions;
  action_t* acts;
  int key_col;          /* -1 = no group by */
  /* first-seen map */
  int64_t* mkeys;
  int32_t* mids;
  uint8_t* mused;
  int64_t mcap, msize;
  int32_t null_id;      /* -1 until the null key is seen */
  int64_t ngroups, gcap;
  int global_seen;
  /* several group-by keys: the UnsafeRow key bytes of every group, in a
   * byte-string first-seen map (LargeStringHashMap) */
  int nkeys;
  int32_t kcols[GZO_MAX_KEYS];
  uint8_t* arena;
  int64_t arena_len, arena_cap;
  int64_t* soff;
  int32_t* slen;
  int32_t* sid;
  uint64_t* shash;
  int64_t scap, ssize;
  uint8_t* rowbuf;  /* the current row's UnsafeRow bytes */
  int64_t rowcap;
};

static int is_binary(int32_t t) { return t == T_STRING || t == T_BINARY; }

static uint64_t mix64(uint64_t x) {
  x ^= x >> 30; x *= 0xbf58476d1ce4e5b9ULL;
  x ^= x >> 27; x *= 0x94d049bb133111ebULL;
  x ^= x >> 31;
  return x;
}

static void map_grow(gzo_agg* a) {
  int64_t nc = a->mcap ? a->mcap * 2 : 1024;
  int64_t* nk = (int64_t*)malloc(sizeof(int64_t) * nc);
  int32_t* ni = (int32_t*)malloc(sizeof(int32_t) * nc);
  uint8_t* nu = (uint8_t*)calloc((size_t)nc, 1);
  for (int64_t i = 0; i < a->mcap; i++) {
    if (!a->mused[i]) continue;
    uint64_t h = mix64((uint64_t)a->mkeys[i]) & (uint64_t)(nc - 1);
    while (nu[h]) h = (h + 1) & (uint64_t)(nc - 1);
    nu[h] = 1; nk[h] = a->mkeys[i]; ni[h] = a->mids[i];
  }
  free(a->mkeys); free(a->mids); free(a->mused);
  a->mkeys = nk; a->mids = ni; a->mused = nu; a->mcap = nc;
}

static void grow_groups(gzo_agg* a, int64_t need) {
  if (need <= a->gcap) return;
  int64_t nc = a->gcap ? a->gcap : 128; /* GrowByFactor: at least 128 */
  while (nc < need) nc *= 2;
  for (int i = 0; i < a->nactions; i++) {
    action_t* x = &a->acts[i];
#define GROW(f, T) do { x->f = (T*)realloc(x->f, sizeof(T) * nc); memset(x->f + a->gcap, 0, sizeof(T) * (nc - a->gcap)); } while (0)
    GROW(dsum, double); GROW(isum, int64_t); GROW(cnt, int64_t); GROW(imm, int64_t);
    GROW(fmm, double); GROW(fraw, int64_t); GROW(valid, uint8_t); GROW(nullf, uint8_t); GROW(key, int64_t);
    GROW(soff, int64_t); GROW(slen, int32_t);
#undef GROW
  }
  a->gcap = nc;
}

/* SparseHashMap::GetOrInsert: the id of a key, a new id (= size) if unseen */
static int32_t get_or_insert(gzo_agg* a, int64_t k) {
  if ((a->msize + 1) * 2 > a->mcap) map_grow(a);
  uint64_t h = mix64((uint64_t)k) & (uint64_t)(a->mcap - 1);
  while (a->mused[h]) {
    if (a->mkeys[h] == k) return a->mids[h];
    h = (h + 1) & (uint64_t)(a->mcap - 1);
  }
  a->mused[h] = 1; a->mkeys[h] = k; a->mids[h] = (int32_t)a->ngroups;
  a->msize++;
  return (int32_t)a->ngroups++;
}

static uint64_t bytes_hash(const uint8_t* p, int32_t n) {
  uint64_t h = 0xcbf29ce484222325ULL; /* FNV-1a, then a final mix */
  for (int32_t i = 0; i < n; i++) h = (h ^ p[i]) * 0x100000001b3ULL;
  return mix64(h);
}

static void bytes_map_grow(gzo_agg* a) {
  int64_t nc = a->scap ? a->scap * 2 : 1024;
  int64_t* no = (int64_t*)malloc(sizeof(int64_t) * nc);
  int32_t* nl = (int32_t*)malloc(sizeof(int32_t) * nc);
  int32_t* ni = (int32_t*)malloc(sizeof(int32_t) * nc);
  uint64_t* nh = (uint64_t*)malloc(sizeof(uint64_t) * nc);
  for (int64_t i = 0; i < nc; i++) ni[i] = -1;
  for (int64_t i = 0; i < a->scap; i++) {
    if (a->sid[i] < 0) continue;
    uint64_t h = a->shash[i] & (uint64_t)(nc - 1);
    while (ni[h] >= 0) h = (h + 1) & (uint64_t)(nc - 1);
    no[h] = a->soff[i]; nl[h] = a->slen[i]; ni[h] = a->sid[i]; nh[h] = a->shash[i];
  }
  free(a->soff); free(a->slen); free(a->sid); free(a->shash);
  a->soff = no; a->slen = nl; a->sid = ni; a->shash = nh; a->scap = nc;
}

/* LargeStringHashMap::GetOrInsert over the key's bytes: the id of an equal
 * byte string, else a new id (= size) */
static int32_t bytes_get_or_insert(gzo_agg* a, const uint8_t* p, int32_t n) {
  if ((a->ssize + 1) * 2 > a->scap) bytes_map_grow(a);
  uint64_t hv = bytes_hash(p, n);
  uint64_t h = hv & (uint64_t)(a->scap - 1);
  while (a->sid[h] >= 0) {
    if (a->shash[h] == hv && a->slen[h] == n && memcmp(a->arena + a->soff[h], p, (size_t)n) == 0)
      return a->sid[h];
    h = (h + 1) & (uint64_t)(a->scap - 1);
  }
  if (a->arena_len + n > a->arena_cap) {
    a->arena_cap = (a->arena_cap + n) * 2;
    a->arena = (uint8_t*)realloc(a->arena, (size_t)a->arena_cap);
  }
  memcpy(a->arena + a->arena_len, p, (size_t)n);
  a->soff[h] = a->arena_len; a->slen[h] = n; a->sid[h] = (int32_t)a->ngroups; a->shash[h] = hv;
  a->arena_len += n;
  a->ssize++;
  return (int32_t)a->ngroups++;
}

/* The key of row r as an UnsafeRow (unsafe_row.h:45-149, unsafe_array.h:58-108):
 * a null bitset and an "empty" bitset of nkeys/8+1 bytes each, then, for every
 * non-null key in order, its value's sizeof(T) bytes, or a string's bytes (no
 * length; an empty string sets its "empty" bit and appends nothing); a null key
 * sets its bit and appends nothing.  Returns the row's byte length; the bytes
 * are in a->rowbuf. */
static int32_t unsafe_row_key(gzo_agg* a, const gzo_batch* b, int64_t r) {
  int32_t bits = a->nkeys / 8 + 1;
  int64_t need = 2 * bits;
  for (int k = 0; k < a->nkeys; k++) {
    int c = a->kcols[k];
    if (is_binary(a->types[c])) {
      const int32_t* off = (const int32_t*)b->bufs[3 * c + 1];
      need += off[r + 1] - off[r];
    } else {
      need += width_of(a->types[c]);
    }
  }
  if (need > a->rowcap) {
    a->rowcap = need * 2;
    a->rowbuf = (uint8_t*)realloc(a->rowbuf, (size_t)a->rowcap);
  }
  uint8_t* row = a->rowbuf;
  memset(row, 0, (size_t)(2 * bits));
  int32_t cur = 2 * bits;
  for (int k = 0; k < a->nkeys; k++) {
    int c = a->kcols[k];
    const uint8_t* vb = (const uint8_t*)b->bufs[3 * c];
    const uint8_t* v = (const uint8_t*)b->bufs[3 * c + 2];
    if (vb && !bit_get(vb, r)) {
      row[k >> 3] |= (uint8_t)(1u << (k & 7));
      continue;
    }
    if (is_binary(a->types[c])) {
      const int32_t* off = (const int32_t*)b->bufs[3 * c + 1];
      int32_t len = off[r + 1] - off[r];
      if (len == 0) row[bits + (k >> 3)] |= (uint8_t)(1u << (k & 7));
      memcpy(row + cur, v + off[r], (size_t)len);
      cur += len;
      continue;
    }
    int w = width_of(a->types[c]);
    memcpy(row + cur, v + r * w, (size_t)w);
    cur += w;
  }
  return cur;
}

gzo_agg* gzo_agg_make(int ncols, const int32_t* types, int nactions, const int32_t* kinds,
                      const int32_t* in0, const int32_t* in1, const int32_t* nan_check) {
  gzo_agg* a = (gzo_agg*)calloc(1, sizeof(gzo_agg));
  a->ncols = ncols;
  a->types = (int32_t*)malloc(sizeof(int32_t) * ncols);
  memcpy(a->types, types, sizeof(int32_t) * ncols);
  a->nactions = nactions;
  a->acts = (action_t*)calloc(nactions, sizeof(action_t));
  a->key_col = -1;
  a->null_id = -1;
  for (int i = 0; i < nactions; i++) {
    a->acts[i].kind = kinds[i];
    a->acts[i].in0 = in0[i];
    a->acts[i].in1 = in1[i];
    a->acts[i].nan_check = nan_check[i];
    a->acts[i].in_type = in0[i] >= 0 ? types[in0[i]] : T_INT64;
    if (kinds[i] == A_GROUPBY) {
      a->key_col = in0[i];
      if (a->nkeys < GZO_MAX_KEYS) a->kcols[a->nkeys] = in0[i];
      a->nkeys++;
    }
  }
  if (a->nkeys > GZO_MAX_KEYS) { gzo_agg_free(a); return NULL; }
  for (int i = 0; i < nactions; i++) {
    /* avgByCount's count column: int64 when grouped (the lambda reads int64),
     * any integer type for the global aggregate (arrow Sum) */
    if (kinds[i] != A_AVG_BY_COUNT) continue;
    const int32_t ct = in1[i] >= 0 && in1[i] < ncols ? types[in1[i]] : -1;
    const int ok = a->key_col >= 0 ? ct == T_INT64 : (ct >= 2 && ct <= 9);
    if (!ok) { gzo_agg_free(a); return NULL; }
  }
  return a;
}

/* One row of one action: the on_valid lambdas of actions_impl.cc. */
static void update(action_t* x, const gzo_batch* b, int64_t r, int32_t g) {
  const uint8_t* vb = x->in0 >= 0 ? (const uint8_t*)b->bufs[3 * x->in0] : NULL;
  const void* v = x->in0 >= 0 ? b->bufs[3 * x->in0 + 2] : NULL;
  int is_null = vb && !bit_get(vb, r);
  int32_t t = x->in_type;
  switch (x->kind) {
    case A_GROUPBY: /* UniqueAction (:159-177): first value or null flag */
      if (!x->valid[g]) {
        x->valid[g] = 1;
        if (is_null) x->nullf[g] = 1;
        else if (is_binary(t)) {
          const int32_t* off = (const int32_t*)b->bufs[3 * x->in0 + 1];
          int32_t len = off[r + 1] - off[r];
          if (x->sused + len > x->scap_bytes) {
            x->scap_bytes = (x->scap_bytes + len) * 2 + 64;
            x->sarena = (uint8_t*)realloc(x->sarena, (size_t)x->scap_bytes);
          }
          memcpy(x->sarena + x->sused, (const uint8_t*)v + off[r], (size_t)len);
          x->soff[g] = x->sused;
          x->slen[g] = len;
          x->sused += len;
        } else if (t == T_DOUBLE) memcpy(&x->key[g], (const double*)v + r, 8);
        else if (t == T_FLOAT) { float f = ((const float*)v)[r]; memcpy(&x->key[g], &f, 4); }
        else x->key[g] = get_i64(v, t, r);
      }
      break;
    case A_COUNT: { /* CountAction (:329-351): rows whose inputs are all non-null */
      int nn = !is_null;
      if (x->in1 >= 0) {
        const uint8_t* vb1 = (const uint8_t*)b->bufs[3 * x->in1];
        if (vb1 && !bit_get(vb1, r)) nn = 0;
      }
      x->isum[g] += nn;
      break;
    }
    case A_COUNT_LITERAL: x->isum[g] += 1; break; /* (:664-667) */
    case A_SUM: /* SumAction (:2021-2042) */
      if (!is_null) {
        x->valid[g] = 1;
        if (is_float(t)) x->dsum[g] += get_f64(v, t, r);
        else x->isum[g] += get_i64(v, t, r);
      }
      break;
    case A_AVG: /* AvgAction (:2690-2710) */
    case A_SUM_COUNT: /* SumCountAction (:3050-3075): sum in double */
      if (!is_null) {
        x->valid[g] = 1;
        x->dsum[g] += get_f64(v, t, r);
        x->cnt[g] += 1;
      }
      break;
    case A_AVG_BY_COUNT: { /* AvgByCountAction (:3785-3795): an int64 count column */
      if (!is_null) {
        const void* cv = b->bufs[3 * x->in1 + 2];
        x->dsum[g] += get_f64(v, t, r);
        x->cnt[g] += ((const int64_t*)cv)[r];
      }
      break;
    }
    case A_MIN:
    case A_MAX:
      if (is_null) break;
      if (is_float(t)) {
        double d = get_f64(v, t, r);
        int64_t raw = raw_bits(v, t, r);
        if (x->kind == A_MIN) {
          /* MinAction (:905-936): LessNaNSafe when NaN_check */
          if (!x->valid[g]) { x->fmm[g] = d; x->fraw[g] = raw; x->valid[g] = 1; }
          int less;
          if (x->nan_check) {
            int ln = isnan(d), rn = isnan(x->fmm[g]);
            less = (ln && rn) ? 0 : ln ? 0 : rn ? 1 : d < x->fmm[g];
          } else less = d < x->fmm[g];
          if (less) { x->fmm[g] = d; x->fraw[g] = raw; }
        } else {
          /* MaxAction (:1504-1539): every NaN overwrites when NaN_check */
          if (x->nan_check && isnan(d)) { x->fmm[g] = d; x->fraw[g] = raw; x->valid[g] = 1; }
          else {
            if (!x->valid[g]) { x->fmm[g] = d; x->fraw[g] = raw; x->valid[g] = 1; }
            if (d > x->fmm[g]) { x->fmm[g] = d; x->fraw[g] = raw; }
          }
        }
      } else if (t == T_UINT64) {
        uint64_t u = ((const uint64_t*)v)[r];
        if (!x->valid[g]) { x->imm[g] = (int64_t)u; x->valid[g] = 1; }
        if (x->kind == A_MIN ? u < (uint64_t)x->imm[g] : u > (uint64_t)x->imm[g]) x->imm[g] = (int64_t)u;
      } else {
        int64_t i = get_i64(v, t, r);
        if (!x->valid[g]) { x->imm[g] = i; x->valid[g] = 1; }
        if (x->kind == A_MIN ? i < x->imm[g] : i > x->imm[g]) x->imm[g] = i;
      }
      break;
  }
}

/* Min/max of a batch with no group-by key: the action's Evaluate(in)
 * (hash_aggregate_kernel.cc:818-839), which differs from the grouped lambdas.
 * Floats (MinAction::GetMinResult, actions_impl.cc:976-993; MaxAction::
 * GetMaxResult, :1581-1603): min always compares with LessNaNSafe, whatever
 * NaN_check says; max with NaN_check takes a batch's first NaN and stops
 * reading that batch.  Integers (:995-1011, :1605-1621): arrow MinMax of the
 * batch, merged with `>` / `<`; an all-null batch yields a null scalar whose
 * value is 0, which is merged all the same once an earlier batch has set the
 * state. */
static void global_min_max(action_t* x, const gzo_batch* b) {
  const uint8_t* vb = (const uint8_t*)b->bufs[3 * x->in0];
  const void* v = b->bufs[3 * x->in0 + 2];
  const int32_t t = x->in_type;
  const int mn = x->kind == A_MIN;
  const int64_t n = b->num_rows;
  if (is_float(t)) {
    for (int64_t r = 0; r < n; r++) {
      if (vb && !bit_get(vb, r)) continue;
      double d = get_f64(v, t, r);
      int64_t raw = raw_bits(v, t, r);
      if (!mn && x->nan_check && isnan(d)) { x->fmm[0] = d; x->fraw[0] = raw; x->valid[0] = 1; break; }
      if (!x->valid[0]) { x->fmm[0] = d; x->fraw[0] = raw; x->valid[0] = 1; }
      int better;
      if (mn) {
        int ln = isnan(d), rn = isnan(x->fmm[0]);
        better = (ln && rn) ? 0 : ln ? 0 : rn ? 1 : d < x->fmm[0];
      } else better = d > x->fmm[0];
      if (better) { x->fmm[0] = d; x->fraw[0] = raw; }
    }
    return;
  }
  int64_t nn = 0, bm = 0;
  for (int64_t r = 0; r < n; r++) {
    if (vb && !bit_get(vb, r)) continue;
    int64_t i = get_i64(v, t, r);
    if (t == T_UINT64) {
      if (nn == 0 || (mn ? (uint64_t)i < (uint64_t)bm : (uint64_t)i > (uint64_t)bm)) bm = i;
    } else if (nn == 0 || (mn ? i < bm : i > bm)) bm = i;
    nn++;
  }
  if (nn > 0 && !x->valid[0]) {
    x->valid[0] = 1;
    x->imm[0] = bm;
    return;
  }
  const int64_t val = nn > 0 ? bm : 0; /* the null scalar's value */
  int take;
  if (t == T_UINT64) take = mn ? (uint64_t)x->imm[0] > (uint64_t)val : (uint64_t)x->imm[0] < (uint64_t)val;
  else take = mn ? x->imm[0] > val : x->imm[0] < val;
  if (take) x->imm[0] = val;
}

/* The global aggregate's AvgByCountAction::Evaluate (actions_impl.cc:3818-3838):
 * arrow Sum of the sum column and of the (integer) count column, each over its
 * own non-null rows. */
static void global_avg_by_count(gzo_agg* a, action_t* x, const gzo_batch* b) {
  const uint8_t* sb = (const uint8_t*)b->bufs[3 * x->in0];
  const void* sv = b->bufs[3 * x->in0 + 2];
  const uint8_t* cb = (const uint8_t*)b->bufs[3 * x->in1];
  const void* cv = b->bufs[3 * x->in1 + 2];
  const int32_t ct = a->types[x->in1];
  for (int64_t r = 0; r < b->num_rows; r++) {
    if (!sb || bit_get(sb, r)) x->dsum[0] += get_f64(sv, x->in_type, r);
    if (!cb || bit_get(cb, r)) x->cnt[0] += get_i64(cv, ct, r);
  }
}

int gzo_agg_process(gzo_agg* a, const gzo_batch* b) {
  int64_t n = b->num_rows;
  if (n == 0) return 0; /* ProcessAndCacheOne returns early on empty input (:808) */
  if (a->key_col < 0) {
    a->ngroups = 1;
    a->global_seen = 1;
    grow_groups(a, 1);
    for (int i = 0; i < a->nactions; i++) {
      action_t* x = &a->acts[i];
      if (x->kind == A_MIN || x->kind == A_MAX) global_min_max(x, b);
      else if (x->kind == A_AVG_BY_COUNT) global_avg_by_count(a, x, b);
      else for (int64_t r = 0; r < n; r++) update(x, b, r, 0);
    }
    return 0;
  }
  int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * n);
  if (a->nkeys > 1) {
    /* several keys: HashAggregateResultIterator<StringType> with an UnsafeRow
     * key (hash_aggregate_kernel.cc:970-1000) */
    for (int64_t r = 0; r < n; r++) {
      int32_t len = unsafe_row_key(a, b, r);
      idx[r] = bytes_get_or_insert(a, a->rowbuf, len);
    }
  } else if (is_binary(a->types[a->key_col])) {
    /* one utf8 key: LargeStringHashMap over the string's bytes, the null key
     * its own group (hash_aggregate_kernel.cc:1001-1020) */
    const uint8_t* kv = (const uint8_t*)b->bufs[3 * a->key_col];
    const int32_t* off = (const int32_t*)b->bufs[3 * a->key_col + 1];
    const uint8_t* data = (const uint8_t*)b->bufs[3 * a->key_col + 2];
    for (int64_t r = 0; r < n; r++) {
      if (kv && !bit_get(kv, r)) {
        if (a->null_id < 0) a->null_id = (int32_t)a->ngroups++;
        idx[r] = a->null_id;
      } else {
        idx[r] = bytes_get_or_insert(a, data + off[r], off[r + 1] - off[r]);
      }
    }
  } else {
    int32_t kt = a->types[a->key_col];
    const uint8_t* kv = (const uint8_t*)b->bufs[3 * a->key_col];
    const void* k = b->bufs[3 * a->key_col + 2];
    for (int64_t r = 0; r < n; r++) {
      if (kv && !bit_get(kv, r)) {
        if (a->null_id < 0) a->null_id = (int32_t)a->ngroups++;
        idx[r] = a->null_id;
      } else {
        int64_t key;
        /* SparseHashMap<float/double> (sparse_hash_map.h:86-160): every NaN is
         * one group (nan_index_); -0.0 and 0.0 hash alike (phmap Hash<double>,
         * phmap_utils.h:251-258) and compare equal, so they are one group too.
         * The group's key comes out as its first row's value (UniqueAction). */
        if (kt == T_DOUBLE) {
          double d = ((const double*)k)[r];
          if (isnan(d)) d = NAN;
          if (d == 0.0) d = 0.0;
          memcpy(&key, &d, 8);
        } else if (kt == T_FLOAT) {
          float f = ((const float*)k)[r];
          int32_t fb;
          if (isnan(f)) f = NAN;
          if (f == 0.0f) f = 0.0f;
          memcpy(&fb, &f, 4);
          key = fb;
        } else key = get_i64(k, kt, r);
        idx[r] = get_or_insert(a, key);
      }
    }
  }
  grow_groups(a, a->ngroups);
  /* the reference runs each action's lambda over all rows, action by action
   * (hash_aggregate_kernel.cc:886-890) */
  for (int i = 0; i < a->nactions; i++)
    for (int64_t r = 0; r < n; r++) update(&a->acts[i], b, r, idx[r]);
  free(idx);
  return 0;
}

int64_t gzo_agg_num_groups(gzo_agg* a) { return a->ngroups; }

int gzo_agg_num_results(gzo_agg* a) {
  int n = 0;
  for (int i = 0; i < a->nactions; i++) n += a->acts[i].kind == A_SUM_COUNT ? 2 : 1;
  return n;
}

static void locate(gzo_agg* a, int col, int* act, int* sub) {
  int n = 0;
  for (int i = 0; i < a->nactions; i++) {
    int w = a->acts[i].kind == A_SUM_COUNT ? 2 : 1;
    if (col < n + w) { *act = i; *sub = col - n; return; }
    n += w;
  }
  *act = -1; *sub = 0;
}

int gzo_agg_result_type(gzo_agg* a, int col) {
  int i, s;
  locate(a, col, &i, &s);
  if (i < 0) return -1;
  action_t* x = &a->acts[i];
  switch (x->kind) {
    case A_GROUPBY: case A_MIN: case A_MAX: return x->in_type;
    case A_COUNT: case A_COUNT_LITERAL: return T_INT64;
    case A_SUM: return is_float(x->in_type) ? T_DOUBLE : T_INT64; /* FindAccumulatorType (:34-51) */
    case A_AVG: case A_AVG_BY_COUNT: return T_DOUBLE;
    case A_SUM_COUNT: return s == 0 ? T_DOUBLE : T_INT64;
  }
  return -1;
}

void gzo_agg_result(gzo_agg* a, int col, void* values, uint8_t* valid) {
  int i, s;
  locate(a, col, &i, &s);
  action_t* x = &a->acts[i];
  int32_t rt = gzo_agg_result_type(a, col);
  int w = width_of(rt);
  for (int64_t g = 0; g < a->ngroups; g++) {
    uint8_t ok = 1;
    int64_t iv = 0;
    double dv = 0;
    int isf = 0;
    switch (x->kind) {
      case A_GROUPBY:
        ok = x->valid[g] && !x->nullf[g];
        iv = x->key[g];
        break;
      case A_COUNT: case A_COUNT_LITERAL: iv = x->isum[g]; break;
      case A_SUM:
        ok = x->valid[g];
        if (is_float(x->in_type)) { dv = x->dsum[g]; isf = 1; } else iv = x->isum[g];
        break;
      case A_AVG:
        /* grouped: null when no non-null input; the global aggregate is always
         * valid (AvgAction::Evaluate, :2744) */
        ok = a->key_col < 0 ? 1 : x->valid[g];
        dv = x->dsum[g] / (double)x->cnt[g];
        isf = 1;
        break;
      case A_SUM_COUNT:
        ok = x->valid[g];
        if (s == 0) { dv = x->dsum[g]; isf = 1; } else iv = x->cnt[g];
        break;
      case A_AVG_BY_COUNT:
        ok = x->cnt[g] != 0;
        dv = ok ? x->dsum[g] / (double)x->cnt[g] : 0;
        isf = 1;
        break;
      case A_MIN: case A_MAX:
        ok = x->valid[g];
        if (is_float(x->in_type)) { dv = x->fmm[g]; isf = 1; } else iv = x->imm[g];
        break;
    }
    valid[g] = ok;
    uint8_t* dst = (uint8_t*)values + g * w;
    if ((x->kind == A_MIN || x->kind == A_MAX) && is_float(x->in_type)) {
      memcpy(dst, &x->fraw[g], (size_t)w); /* the kept value's bits */
    } else if (x->kind == A_GROUPBY) {
      memcpy(dst, &iv, (size_t)w); /* the key's bits (a float key's as stored) */
    } else if (rt == T_DOUBLE) {
      double d = isf ? dv : (double)iv;
      memcpy(dst, &d, 8);
    } else if (rt == T_FLOAT) {
      float f = (float)dv;
      memcpy(dst, &f, 4);
    } else {
      memcpy(dst, &iv, (size_t)w); /* little-endian truncation to the result width */
    }
    (void)is_unsigned;
  }
}

int64_t gzo_agg_result_bytes(gzo_agg* a, int col) {
  int i, s;
  locate(a, col, &i, &s);
  if (i < 0 || a->acts[i].kind != A_GROUPBY || !is_binary(a->acts[i].in_type)) return -1;
  int64_t n = 0;
  for (int64_t g = 0; g < a->ngroups; g++)
    if (a->acts[i].valid[g] && !a->acts[i].nullf[g]) n += a->acts[i].slen[g];
  return n;
}

void gzo_agg_result_binary(gzo_agg* a, int col, int32_t* offsets, uint8_t* data, uint8_t* valid) {
  int i, s;
  locate(a, col, &i, &s);
  action_t* x = &a->acts[i];
  int32_t pos = 0;
  offsets[0] = 0;
  for (int64_t g = 0; g < a->ngroups; g++) {
    uint8_t ok = x->valid[g] && !x->nullf[g];
    valid[g] = ok;
    if (ok) {
      memcpy(data + pos, x->sarena + x->soff[g], (size_t)x->slen[g]);
      pos += x->slen[g];
    }
    offsets[g + 1] = pos;
  }
}

void gzo_agg_free(gzo_agg* a) {
  for (int i = 0; i < a->nactions; i++) {
    action_t* x = &a->acts[i];
    free(x->dsum); free(x->isum); free(x->cnt); free(x->imm); free(x->fmm); free(x->fraw);
    free(x->valid); free(x->nullf); free(x->key); free(x->soff); free(x->slen); free(x->sarena);
  }
  free(a->acts); free(a->types); free(a->mkeys); free(a->mids); free(a->mused);
  free(a->arena); free(a->soff); free(a->slen); free(a->sid); free(a->shash); free(a->rowbuf);
  free(a);
}

int gzo_agg_process_many(gzo_agg* a, const gzo_batch* batches, int nb) {
  for (int i = 0; i < nb; i++) {
    int rc = gzo_agg_process(a, &batches[i]);
    if (rc) return rc;
  }
  return 0;
}
