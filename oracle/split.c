/*
 * split.c -- CPU restatement of the reference shuffle Splitter's hash
 * partitioning and stable counting-sort split.  TEST INFRASTRUCTURE ONLY (the
 * checker and the CPU baseline); the product never calls it.
 */
#include <stdlib.h>
#include <string.h>

#include "gz_oracle.h"

/* ---- Spark Murmur3_x86_32 -------------------------------------------------
 * The mixing steps are those of murmurhash32.h:74-108 (k1 *= c1; rotl 15;
 * k1 *= c2; h1 ^= k1; rotl 13; h1 = h1*5 + 0xe6546b64; fmix), applied to the
 * low then the high 32-bit word of the key, exactly Spark's hashLong. */
static inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }
static inline uint32_t mix_k1(uint32_t k1) {
  k1 *= 0xcc9e2d51u;
  k1 = rotl32(k1, 15);
  return k1 * 0x1b873593u;
}
static inline uint32_t mix_h1(uint32_t h1, uint32_t k1) {
  h1 ^= k1;
  h1 = rotl32(h1, 13);
  return h1 * 5u + 0xe6546b64u;
}
static inline uint32_t fmix(uint32_t h1, uint32_t len) {
  h1 ^= len;
  h1 ^= h1 >> 16;
  h1 *= 0x85ebca6bu;
  h1 ^= h1 >> 13;
  h1 *= 0xc2b2ae35u;
  h1 ^= h1 >> 16;
  return h1;
}

int32_t gzo_hash_int(int32_t v, int32_t seed) {
  return (int32_t)fmix(mix_h1((uint32_t)seed, mix_k1((uint32_t)v)), 4);
}

int32_t gzo_hash_long(int64_t v, int32_t seed) {
  uint64_t u = (uint64_t)v;
  uint32_t h = mix_h1((uint32_t)seed, mix_k1((uint32_t)u));
  h = mix_h1(h, mix_k1((uint32_t)(u >> 32)));
  return (int32_t)fmix(h, 8);
}

int32_t gzo_hash_bytes(const uint8_t* p, int32_t len, int32_t seed) {
  int32_t aligned = len - len % 4;
  uint32_t h = (uint32_t)seed;
  for (int32_t i = 0; i < aligned; i += 4) {
    uint32_t w;
    memcpy(&w, p + i, 4);
    h = mix_h1(h, mix_k1(w));
  }
  for (int32_t i = aligned; i < len; i++) {
    int32_t half = (int32_t)(int8_t)p[i]; /* Java byte: sign-extended */
    h = mix_h1(h, mix_k1((uint32_t)half));
  }
  return (int32_t)fmix(h, (uint32_t)len);
}

int32_t gzo_hash32_numeric(double v) {
  int64_t bits;
  memcpy(&bits, &v, 8);
  return gzo_hash_long(bits, 0);
}

/* ---- column helpers ---- */
enum { T_NA = 0, T_BOOL = 1, T_UINT8 = 2, T_INT8 = 3, T_UINT16 = 4, T_INT16 = 5,
       T_UINT32 = 6, T_INT32 = 7, T_UINT64 = 8, T_INT64 = 9, T_FLOAT = 11, T_DOUBLE = 12,
       T_STRING = 13, T_BINARY = 14, T_DATE32 = 16, T_DATE64 = 17, T_TIMESTAMP = 18,
       T_DECIMAL128 = 23 };

static int width_of(int32_t t) {
  switch (t) {
    case T_UINT8: case T_INT8: return 1;
    case T_UINT16: case T_INT16: return 2;
    case T_UINT32: case T_INT32: case T_FLOAT: case T_DATE32: return 4;
    case T_UINT64: case T_INT64: case T_DOUBLE: case T_DATE64: case T_TIMESTAMP: return 8;
    case T_DECIMAL128: return 16;
    default: return 0;
  }
}
static int is_binary(int32_t t) { return t == T_STRING || t == T_BINARY; }

static inline int bit_get(const uint8_t* bm, int64_t i) { return (bm[i >> 3] >> (i & 7)) & 1; }
static inline void bit_clear(uint8_t* bm, int64_t i) { bm[i >> 3] &= (uint8_t)~(1u << (i & 7)); }

#define BUF(b, c, k) ((const uint8_t*)(b)->bufs[3 * (c) + (k)])

/* Spark hash of one key value (HashSplitter::CreateProjector, splitter.cc:1402-1438
 * picks hash32_spark / hash64_spark / hashbuf_spark by type). */
static int key_hash(const gzo_batch* b, int c, int32_t t, int64_t r, int32_t h, int* ok) {
  const uint8_t* val = BUF(b, c, 0);
  if (val && !bit_get(val, r)) return h; /* null: seed passes through */
  const uint8_t* v = BUF(b, c, 2);
  switch (t) {
    case T_NA: return h;
    case T_BOOL: return gzo_hash_int(bit_get(v, r), h);
    case T_INT8: return gzo_hash_int(((const int8_t*)v)[r], h);
    case T_INT16: return gzo_hash_int(((const int16_t*)v)[r], h);
    case T_INT32: case T_DATE32: return gzo_hash_int(((const int32_t*)v)[r], h);
    case T_INT64: return gzo_hash_long(((const int64_t*)v)[r], h);
    /* Spark hashes a float as hashInt(floatToIntBits(f)) and a double as
     * hashLong(doubleToLongBits(d)) (hash32_spark / hash64_spark).  -0.0 and NaN
     * are refused (status 10): whether they hash as 0.0 / one canonical NaN
     * depends on the Spark version and on the Gandiva fork, which is not here. */
    case T_FLOAT: {
      uint32_t bits = ((const uint32_t*)v)[r];
      if (bits == 0x80000000u || (bits & 0x7fffffffu) > 0x7f800000u) { *ok = 0; return h; }
      return gzo_hash_int((int32_t)bits, h);
    }
    case T_DOUBLE: {
      uint64_t bits = ((const uint64_t*)v)[r];
      if (bits == 0x8000000000000000ULL || (bits & 0x7fffffffffffffffULL) > 0x7ff0000000000000ULL) {
        *ok = 0;
        return h;
      }
      return gzo_hash_long((int64_t)bits, h);
    }
    case T_STRING: case T_BINARY: {
      const int32_t* off = (const int32_t*)BUF(b, c, 1);
      return gzo_hash_bytes(v + off[r], off[r + 1] - off[r], h);
    }
    default: *ok = 0; return h;
  }
}

int gzo_split_pids(int mode, const gzo_batch* b, int ncols, const int32_t* types,
                   const int32_t* keys, int nkeys, int P, int32_t seed, int32_t* rr_state,
                   int32_t* pids) {
  (void)ncols;
  int64_t n = b->num_rows;
  if (mode == 0) {
    for (int64_t r = 0; r < n; r++) {
      int ok = 1;
      int32_t h = seed;
      for (int k = 0; k < nkeys; k++) h = key_hash(b, keys[k], types[keys[k]], r, h, &ok);
      if (!ok) return 10;
      /* positive mod, as the cmovs fix-up of splitter.cc:1456-1465 */
      int32_t pid = h % P;
      if (pid < 0) pid += P;
      pids[r] = pid;
    }
  } else if (mode == 1 || mode == 3) {
    int32_t s = *rr_state;
    for (int64_t r = 0; r < n; r++) {
      pids[r] = mode == 3 ? 0 : s;
      s = (s + 1 == P) ? 0 : s + 1;
    }
    *rr_state = s;
  } else {
    const int32_t* src = (const int32_t*)BUF(b, 0, 2);
    for (int64_t r = 0; r < n; r++) {
      if (src[r] >= P || src[r] < 0) return 4;
      pids[r] = src[r];
    }
  }
  return 0;
}

void gzo_split_layout(const gzo_batch* batches, int nb, int ncols, const int32_t* types,
                      int P, const int32_t* pids, int64_t* row_start, int64_t* byte_start) {
  memset(row_start, 0, sizeof(int64_t) * (P + 1));
  int nbin = 0;
  for (int c = 0; c < ncols; c++) nbin += is_binary(types[c]);
  if (nbin) memset(byte_start, 0, sizeof(int64_t) * (P + 1) * nbin);
  int64_t g = 0;
  for (int bi = 0; bi < nb; bi++) {
    const gzo_batch* b = &batches[bi];
    for (int64_t r = 0; r < b->num_rows; r++) {
      int p = pids[g + r];
      row_start[p + 1]++;
      int j = 0;
      for (int c = 0; c < ncols; c++) {
        if (!is_binary(types[c])) continue;
        const int32_t* off = (const int32_t*)BUF(b, c, 1);
        byte_start[j * (P + 1) + p + 1] += off[r + 1] - off[r];
        j++;
      }
    }
    g += b->num_rows;
  }
  for (int p = 0; p < P; p++) row_start[p + 1] += row_start[p];
  for (int j = 0; j < nbin; j++)
    for (int p = 0; p < P; p++) byte_start[j * (P + 1) + p + 1] += byte_start[j * (P + 1) + p];
}

void gzo_split_scatter(const gzo_batch* batches, int nb, int ncols, const int32_t* types,
                       int P, const int32_t* pids, const int64_t* row_start,
                       const int64_t* byte_start, void** out_values, uint8_t** out_validity,
                       int32_t** out_offsets, uint8_t** out_data) {
  /* partition_buffer_idx_base_ (splitter.h:230): rows already in each partition */
  int64_t* idx_base = (int64_t*)calloc(P, sizeof(int64_t));
  int nbin = 0;
  for (int c = 0; c < ncols; c++) nbin += is_binary(types[c]);
  int64_t* value_offset = (int64_t*)calloc((size_t)P * (nbin ? nbin : 1), sizeof(int64_t));
  uint32_t* cnt = (uint32_t*)malloc(sizeof(uint32_t) * P);
  uint32_t* roo = (uint32_t*)malloc(sizeof(uint32_t) * (P + 1));
  int64_t g = 0;
  for (int bi = 0; bi < nb; bi++) {
    const gzo_batch* b = &batches[bi];
    int64_t n = b->num_rows;
    const int32_t* pid = pids + g;
    /* DoSplit (splitter.cc:855-875): partition counts, exclusive scan into
     * reducer_offset_offset_, stable counting sort of row ids. */
    memset(cnt, 0, sizeof(uint32_t) * P);
    for (int64_t r = 0; r < n; r++) cnt[pid[r]]++;
    uint32_t* reducer_offsets = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
    roo[0] = 0;
    for (int p = 1; p <= P; p++) roo[p] = roo[p - 1] + cnt[p - 1];
    for (int64_t r = 0; r < n; r++) reducer_offsets[roo[pid[r]]++] = (uint32_t)r;
    for (int p = 0; p < P; p++) roo[p] -= cnt[p];

    int j = 0;
    for (int c = 0; c < ncols; c++) {
      int32_t t = types[c];
      if (t == T_NA) continue;
      const uint8_t* src_valid = BUF(b, c, 0);
      /* SplitValidityBuffer (:1179-1207): bit copy into the partition bitmap */
      if (src_valid && out_validity[c]) {
        for (int p = 0; p < P; p++)
          for (uint32_t x = roo[p]; x < roo[p + 1]; x++) {
            uint32_t r = reducer_offsets[x];
            if (!bit_get(src_valid, r))
              bit_clear(out_validity[c], row_start[p] + idx_base[p] + (x - roo[p]));
          }
      }
      if (t == T_BOOL) {
        /* SplitBoolType (:1085-1177) on the value bits */
        const uint8_t* src = BUF(b, c, 2);
        uint8_t* dst = (uint8_t*)out_values[c];
        for (int p = 0; p < P; p++)
          for (uint32_t x = roo[p]; x < roo[p + 1]; x++)
            if (!bit_get(src, reducer_offsets[x]))
              bit_clear(dst, row_start[p] + idx_base[p] + (x - roo[p]));
      } else if (is_binary(t)) {
        /* SplitBinaryType (:1209-1278): per partition, offsets accumulate the
         * running value_offset; bytes are appended in order. */
        const int32_t* soff = (const int32_t*)BUF(b, c, 1);
        const uint8_t* sdata = BUF(b, c, 2);
        for (int p = 0; p < P; p++) {
          int32_t* doff = out_offsets[c] + row_start[p] + p + idx_base[p];
          uint8_t* ddata = out_data[c] + byte_start[j * (P + 1) + p];
          int64_t vo = value_offset[j * P + p];
          if (idx_base[p] == 0) doff[0] = 0;
          for (uint32_t x = roo[p]; x < roo[p + 1]; x++) {
            uint32_t r = reducer_offsets[x];
            int32_t len = soff[r + 1] - soff[r];
            memcpy(ddata + vo, sdata + soff[r], (size_t)len);
            vo += len;
            doff[x - roo[p] + 1] = (int32_t)vo;
          }
          value_offset[j * P + p] = vo;
        }
        j++;
      } else {
        /* SplitFixedType<T> (:939-958): dst[p][base_p + k] = src[reducer_offsets[r]] */
        int w = width_of(t);
        const uint8_t* src = BUF(b, c, 2);
        uint8_t* dst = (uint8_t*)out_values[c];
        for (int p = 0; p < P; p++)
          for (uint32_t x = roo[p]; x < roo[p + 1]; x++)
            memcpy(dst + (size_t)(row_start[p] + idx_base[p] + (x - roo[p])) * w,
                   src + (size_t)reducer_offsets[x] * w, (size_t)w);
      }
    }
    /* partition_buffer_idx_base_[pid] += partition_id_cnt_[pid] (:933-935) */
    for (int p = 0; p < P; p++) idx_base[p] += cnt[p];
    free(reducer_offsets);
    g += n;
  }
  free(idx_base);
  free(value_offset);
  free(cnt);
  free(roo);
}

/* ---- whole per-batch splitter for the CPU baseline ---- */
struct gzo_splitter {
  int ncols, P, nkeys;
  int32_t seed;
  int32_t* types;
  int32_t* keys;
  int64_t* rows;       /* rows per partition */
  int64_t* cap;        /* capacity (rows) per partition */
  uint8_t*** vals;     /* [col][p] value buffers */
  uint8_t*** valid;    /* [col][p] validity bitmaps */
  int32_t*** offs;     /* [col][p] offsets */
  uint8_t*** data;     /* [col][p] data */
  int64_t** dcap;      /* [col][p] data capacity */
  int32_t* pid;
  uint32_t* reducer_offsets;
  int64_t scratch_rows;
};

gzo_splitter* gzo_splitter_make(int ncols, const int32_t* types, int P, const int32_t* keys,
                                int nkeys, int32_t seed) {
  gzo_splitter* s = (gzo_splitter*)calloc(1, sizeof(gzo_splitter));
  s->ncols = ncols; s->P = P; s->nkeys = nkeys; s->seed = seed;
  s->types = (int32_t*)malloc(sizeof(int32_t) * ncols);
  memcpy(s->types, types, sizeof(int32_t) * ncols);
  s->keys = (int32_t*)malloc(sizeof(int32_t) * (nkeys ? nkeys : 1));
  memcpy(s->keys, keys, sizeof(int32_t) * nkeys);
  s->rows = (int64_t*)calloc(P, sizeof(int64_t));
  s->cap = (int64_t*)calloc(P, sizeof(int64_t));
  s->vals = (uint8_t***)calloc(ncols, sizeof(void*));
  s->valid = (uint8_t***)calloc(ncols, sizeof(void*));
  s->offs = (int32_t***)calloc(ncols, sizeof(void*));
  s->data = (uint8_t***)calloc(ncols, sizeof(void*));
  s->dcap = (int64_t**)calloc(ncols, sizeof(void*));
  for (int c = 0; c < ncols; c++) {
    s->vals[c] = (uint8_t**)calloc(P, sizeof(void*));
    s->valid[c] = (uint8_t**)calloc(P, sizeof(void*));
    s->offs[c] = (int32_t**)calloc(P, sizeof(void*));
    s->data[c] = (uint8_t**)calloc(P, sizeof(void*));
    s->dcap[c] = (int64_t*)calloc(P, sizeof(int64_t));
  }
  return s;
}

static void grow_partition(gzo_splitter* s, int p, int64_t need) {
  int64_t nc = s->cap[p] ? s->cap[p] : 4096;
  while (nc < need) nc += nc >> 1;
  for (int c = 0; c < s->ncols; c++) {
    int32_t t = s->types[c];
    if (t == T_NA) continue;
    int64_t vb = t == T_BOOL ? (nc + 7) / 8 : is_binary(t) ? 0 : nc * width_of(t);
    if (vb) s->vals[c][p] = (uint8_t*)realloc(s->vals[c][p], (size_t)vb);
    int64_t ob = (s->cap[p] + 7) / 8;
    s->valid[c][p] = (uint8_t*)realloc(s->valid[c][p], (size_t)(nc + 7) / 8);
    memset(s->valid[c][p] + ob, 0xff, (size_t)((nc + 7) / 8 - ob));
    if (t == T_BOOL) memset(s->vals[c][p] + ob, 0xff, (size_t)((nc + 7) / 8 - ob));
    if (is_binary(t)) s->offs[c][p] = (int32_t*)realloc(s->offs[c][p], sizeof(int32_t) * (nc + 1));
  }
  s->cap[p] = nc;
}

void gzo_splitter_split(gzo_splitter* s, const gzo_batch* b) {
  int64_t n = b->num_rows;
  int P = s->P;
  if (n > s->scratch_rows) {
    s->pid = (int32_t*)realloc(s->pid, sizeof(int32_t) * n);
    s->reducer_offsets = (uint32_t*)realloc(s->reducer_offsets, sizeof(uint32_t) * n);
    s->scratch_rows = n;
  }
  int32_t rr = 0;
  gzo_split_pids(0, b, s->ncols, s->types, s->keys, s->nkeys, P, s->seed, &rr, s->pid);
  uint32_t cnt_stack[1024];
  uint32_t* cnt = P + 1 <= 1024 ? cnt_stack : (uint32_t*)malloc(sizeof(uint32_t) * (P + 1));
  uint32_t* roo = (uint32_t*)malloc(sizeof(uint32_t) * (P + 1));
  memset(cnt, 0, sizeof(uint32_t) * P);
  for (int64_t r = 0; r < n; r++) cnt[s->pid[r]]++;
  roo[0] = 0;
  for (int p = 1; p <= P; p++) roo[p] = roo[p - 1] + cnt[p - 1];
  for (int64_t r = 0; r < n; r++) s->reducer_offsets[roo[s->pid[r]]++] = (uint32_t)r;
  for (int p = 0; p < P; p++) roo[p] -= cnt[p];
  for (int p = 0; p < P; p++)
    if (cnt[p] && s->rows[p] + cnt[p] > s->cap[p]) grow_partition(s, p, s->rows[p] + cnt[p]);
  for (int c = 0; c < s->ncols; c++) {
    int32_t t = s->types[c];
    if (t == T_NA) continue;
    const uint8_t* sv = BUF(b, c, 0);
    if (sv)
      for (int p = 0; p < P; p++)
        for (uint32_t x = roo[p]; x < roo[p + 1]; x++)
          if (!bit_get(sv, s->reducer_offsets[x])) bit_clear(s->valid[c][p], s->rows[p] + x - roo[p]);
    if (t == T_BOOL) {
      const uint8_t* src = BUF(b, c, 2);
      for (int p = 0; p < P; p++)
        for (uint32_t x = roo[p]; x < roo[p + 1]; x++)
          if (!bit_get(src, s->reducer_offsets[x])) bit_clear(s->vals[c][p], s->rows[p] + x - roo[p]);
    } else if (is_binary(t)) {
      const int32_t* soff = (const int32_t*)BUF(b, c, 1);
      const uint8_t* sd = BUF(b, c, 2);
      for (int p = 0; p < P; p++) {
        if (!cnt[p]) continue;
        int32_t* doff = s->offs[c][p] + s->rows[p];
        if (s->rows[p] == 0) doff[0] = 0;
        int64_t vo = doff[0];
        for (uint32_t x = roo[p]; x < roo[p + 1]; x++) {
          uint32_t r = s->reducer_offsets[x];
          int32_t len = soff[r + 1] - soff[r];
          if (vo + len > s->dcap[c][p]) {
            int64_t nc = s->dcap[c][p] + (s->dcap[c][p] >> 1);
            if (nc < vo + len) nc = vo + len + 4096;
            s->data[c][p] = (uint8_t*)realloc(s->data[c][p], (size_t)nc);
            s->dcap[c][p] = nc;
          }
          memcpy(s->data[c][p] + vo, sd + soff[r], (size_t)len);
          vo += len;
          doff[x - roo[p] + 1] = (int32_t)vo;
        }
      }
    } else {
      int w = width_of(t);
      const uint8_t* src = BUF(b, c, 2);
      for (int p = 0; p < P; p++) {
        uint8_t* dst = s->vals[c][p] + (size_t)s->rows[p] * w;
        for (uint32_t x = roo[p]; x < roo[p + 1]; x++, dst += w)
          memcpy(dst, src + (size_t)s->reducer_offsets[x] * w, (size_t)w);
      }
    }
  }
  for (int p = 0; p < P; p++) s->rows[p] += cnt[p];
  if (cnt != cnt_stack) free(cnt);
  free(roo);
}

int64_t gzo_splitter_rows(gzo_splitter* s, int p) { return s->rows[p]; }

void gzo_splitter_free(gzo_splitter* s) {
  for (int c = 0; c < s->ncols; c++) {
    for (int p = 0; p < s->P; p++) {
      free(s->vals[c][p]); free(s->valid[c][p]); free(s->offs[c][p]); free(s->data[c][p]);
    }
    free(s->vals[c]); free(s->valid[c]); free(s->offs[c]); free(s->data[c]); free(s->dcap[c]);
  }
  free(s->vals); free(s->valid); free(s->offs); free(s->data); free(s->dcap);
  free(s->rows); free(s->cap); free(s->types); free(s->keys); free(s->pid);
  free(s->reducer_offsets);
  free(s);
}

void gzo_splitter_split_many(gzo_splitter* s, const gzo_batch* batches, int nb) {
  for (int i = 0; i < nb; i++) gzo_splitter_split(s, &batches[i]);
}
