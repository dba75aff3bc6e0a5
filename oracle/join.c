/*
 * join.c -- CPU restatement of the reference hash join: HashRelation build on an
 * unsafeHashMap and the inner / semi / anti probe functions.  TEST
 * INFRASTRUCTURE ONLY (checker + CPU baseline).
 *
 * Table layout and probing follow third_party/row_wise_memory/hashMap.h:
 * slots of [value-or-chain][hash32][key], empty = -1, triangular probing
 * pos = (pos + step++) & mask (:322-363, :782-887); a repeated key keeps its
 * rows in insertion order (the bytesMap chain, :826-876); builder type 3 skips
 * repeated keys (appendNewKey, :1176-1250).  The capacity is
 * 2^(ceil(log2 rows) + 1), or 128 for at most 32 rows
 * (hash_relation_kernel.cc:205-224).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "gz_oracle.h"

enum { T_BOOL = 1, T_UINT8 = 2, T_INT8 = 3, T_UINT16 = 4, T_INT16 = 5, T_UINT32 = 6,
       T_INT32 = 7, T_UINT64 = 8, T_INT64 = 9, T_FLOAT = 11, T_DOUBLE = 12, T_DATE32 = 16,
       T_DATE64 = 17, T_TIMESTAMP = 18 };

typedef struct {
  int32_t head;   /* first build row of the key (index into rows) */
  int32_t hash;
  int64_t key;    /* value bits widened to 64 */
  int32_t used;
} slot_t;

struct gzo_rel {
  int32_t key_type, builder_type;
  int64_t* keys;    /* appended build keys (widened) */
  uint8_t* nulls;
  int64_t n, cap_rows;
  slot_t* slots;
  int64_t nslots, nkeys;
  int32_t* next;    /* chain: next build row of the same key, -1 at the end */
  int32_t* tail;    /* per slot: last row of the chain */
  int has_null;
};

static int is_float(int32_t t) { return t == T_FLOAT || t == T_DOUBLE; }

static int64_t widen(const void* v, int32_t t, int64_t r) {
  switch (t) {
    case T_BOOL: return (((const uint8_t*)v)[r >> 3] >> (r & 7)) & 1;
    case T_INT8: return ((const int8_t*)v)[r];
    case T_UINT8: return ((const uint8_t*)v)[r];
    case T_INT16: return ((const int16_t*)v)[r];
    case T_UINT16: return ((const uint16_t*)v)[r];
    case T_INT32: case T_DATE32: return ((const int32_t*)v)[r];
    case T_UINT32: return ((const uint32_t*)v)[r];
    case T_FLOAT: { double d = ((const float*)v)[r]; int64_t b; memcpy(&b, &d, 8); return b; }
    default: return ((const int64_t*)v)[r];
  }
}

/* the numeric key as the double the reference hashes (murmurhash32.h:262) */
static double as_double(int64_t w, int32_t t) {
  if (is_float(t)) { double d; memcpy(&d, &w, 8); return d; }
  if (t == T_UINT64) return (double)(uint64_t)w;
  return (double)w;
}

/* key equality of the typed keys (keyRow == *(CType*)...) */
static int key_eq(int64_t a, int64_t b, int32_t t) {
  if (is_float(t)) {
    double x, y;
    memcpy(&x, &a, 8); memcpy(&y, &b, 8);
    return x == y;
  }
  return a == b;
}

gzo_rel* gzo_rel_make(int32_t key_type, int32_t builder_type) {
  gzo_rel* r = (gzo_rel*)calloc(1, sizeof(gzo_rel));
  r->key_type = key_type;
  r->builder_type = builder_type;
  return r;
}

void gzo_rel_append(gzo_rel* r, const void* keys, const uint8_t* validity, int64_t n) {
  if (r->n + n > r->cap_rows) {
    int64_t nc = r->cap_rows ? r->cap_rows : 1024;
    while (nc < r->n + n) nc *= 2;
    r->keys = (int64_t*)realloc(r->keys, sizeof(int64_t) * nc);
    r->nulls = (uint8_t*)realloc(r->nulls, (size_t)nc);
    r->cap_rows = nc;
  }
  for (int64_t i = 0; i < n; i++) {
    int isnull = validity && !((validity[i >> 3] >> (i & 7)) & 1);
    r->keys[r->n + i] = isnull ? 0 : widen(keys, r->key_type, i);
    r->nulls[r->n + i] = (uint8_t)isnull;
  }
  r->n += n;
}

static int64_t lookup(const gzo_rel* r, int64_t key, int32_t hash) {
  int64_t mask = r->nslots - 1;
  int64_t pos = hash & mask, step = 1;
  while (1) {
    const slot_t* s = &r->slots[pos];
    if (!s->used) return -1;
    if (s->hash == hash && key_eq(s->key, key, r->key_type)) return pos;
    pos = (pos + step) & mask;
    step++;
  }
}

void gzo_rel_finish(gzo_rel* r) {
  int64_t cap = 128;
  if (r->n > 32) cap = (int64_t)pow(2.0, ceil(log2((double)r->n)) + 1);
  r->nslots = cap;
  r->slots = (slot_t*)calloc((size_t)cap, sizeof(slot_t));
  r->next = (int32_t*)malloc(sizeof(int32_t) * (r->n ? r->n : 1));
  r->tail = (int32_t*)malloc(sizeof(int32_t) * cap);
  int64_t mask = cap - 1;
  for (int64_t i = 0; i < r->n; i++) {
    r->next[i] = -1;
    if (r->nulls[i]) { r->has_null = 1; continue; } /* InsertNull (hash_relation.h:636-644) */
    int64_t key = r->keys[i];
    int32_t h = gzo_hash32_numeric(as_double(key, r->key_type));
    int64_t pos = h & mask, step = 1;
    while (1) {
      slot_t* s = &r->slots[pos];
      if (!s->used) {
        s->used = 1; s->hash = h; s->key = key; s->head = (int32_t)i;
        r->tail[pos] = (int32_t)i;
        r->nkeys++;
        break;
      }
      if (s->hash == h && key_eq(s->key, key, r->key_type)) {
        if (r->builder_type != 3) { /* append keeps insertion order */
          r->next[r->tail[pos]] = (int32_t)i;
          r->tail[pos] = (int32_t)i;
        }
        break;
      }
      pos = (pos + step) & mask;
      step++;
    }
  }
}

int64_t gzo_rel_num_keys(gzo_rel* r) { return r->nkeys; }
int gzo_rel_has_null(gzo_rel* r) { return r->has_null; }

int64_t gzo_probe(gzo_rel* r, int join_type, int null_aware, const void* keys,
                  const uint8_t* validity, int64_t n, int64_t* out_probe, int64_t* out_build) {
  int64_t out = 0;
  for (int64_t i = 0; i < n; i++) {
    int isnull = validity && !((validity[i >> 3] >> (i & 7)) & 1);
    int64_t pos = -1;
    if (!isnull) {
      int64_t key = widen(keys, r->key_type, i);
      pos = lookup(r, key, gzo_hash32_numeric(as_double(key, r->key_type)));
    }
    if (join_type == 0) {
      /* UnsafeInnerProbeFunction (conditioned_probe_kernel.cc:546-665): every
       * build row of the key, in insertion order; null keys never match */
      if (pos < 0) continue;
      for (int32_t b = r->slots[pos].head; b >= 0; b = r->next[b]) {
        if (out_probe) { out_probe[out] = i; out_build[out] = b; }
        out++;
      }
    } else if (join_type == 3) {
      /* UnsafeSemiProbeFunction (:1038-1173) */
      if (pos >= 0) { if (out_probe) out_probe[out] = i; out++; }
    } else if (join_type == 2) {
      /* UnsafeAntiProbeFunction (:861-1036) */
      int emit;
      if (!null_aware) emit = pos < 0;  /* a null key is emitted */
      else if (r->nkeys == 0 && !r->has_null) emit = 1;
      else if (r->has_null) emit = 0;
      else emit = !isnull && pos < 0;
      if (emit) { if (out_probe) out_probe[out] = i; out++; }
    }
  }
  return out;
}

void gzo_rel_free(gzo_rel* r) {
  free(r->keys); free(r->nulls); free(r->slots); free(r->next); free(r->tail);
  free(r);
}
