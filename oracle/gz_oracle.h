/*
 * gz_oracle.h -- CPU restatement of the reference's hot-path algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path and
 * the timed CPU baseline; the product never links, loads or calls it.
 *
 * Every function follows a named reference function (file:line under
 * native-sql-engine/cpp/src/); the pins that check each one are listed in
 * oracle/README.md.
 */
#ifndef GZ_ORACLE_H_
#define GZ_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One batch: three buffer slots per column (validity, offsets, values); NULL
 * where a column has no such buffer. */
typedef struct {
  int64_t num_rows;
  const void* const* bufs;
} gzo_batch;

/* Spark Murmur3_x86_32 hashInt / hashLong (the arithmetic of
 * third_party/murmurhash/murmurhash32.h:74-108 with a caller seed). */
int32_t gzo_hash_int(int32_t v, int32_t seed);
int32_t gzo_hash_long(int64_t v, int32_t seed);
/* Spark hashUnsafeBytes (legacy per-byte tail). */
int32_t gzo_hash_bytes(const uint8_t* p, int32_t len, int32_t seed);
/* Join-internal hash: murmurhash32::hash32 over the double bits of a numeric
 * key, seed 0 (murmurhash32.h:251-265). */
int32_t gzo_hash32_numeric(double v);

/* Partition id of every row (HashSplitter::ComputeAndCountPartitionId,
 * splitter.cc:1440-1470; RoundRobinSplitter :1369-1387; FallbackRangeSplitter
 * :1502-1521).  mode: 0 hash, 1 rr, 2 range, 3 single.  *rr_state carries the
 * round-robin cursor across calls.  Returns 0, or 4 (Invalid) on a bad pid. */
int gzo_split_pids(int mode, const gzo_batch* b, int ncols, const int32_t* types,
                   const int32_t* keys, int nkeys, int P, int32_t seed, int32_t* rr_state,
                   int32_t* pids);

/* Counting-sort layout of a list of batches whose pids are given (concatenated
 * over batches): row_start[P+1] and, for each utf8/binary column j (in column
 * order), byte_start[j*(P+1) + p]. */
void gzo_split_layout(const gzo_batch* batches, int nb, int ncols, const int32_t* types,
                      int P, const int32_t* pids, int64_t* row_start, int64_t* byte_start);

/* Splitter::DoSplit + SplitFixedType / SplitBoolType / SplitValidityBuffer /
 * SplitBinaryType (splitter.cc:855-1303) into the concatenated partition layout.
 * out_values[c]: fixed values (bool: bitmap) at global row positions;
 * out_validity[c]: bitmap or NULL; out_offsets[c]: N+P int32; out_data[c]: bytes.
 * Bitmaps must be pre-filled with 0xFF by the caller (the reference initializes
 * partition bitmaps to all-true, splitter.cc:1187-1200). */
void gzo_split_scatter(const gzo_batch* batches, int nb, int ncols, const int32_t* types,
                       int P, const int32_t* pids, const int64_t* row_start,
                       const int64_t* byte_start, void** out_values, uint8_t** out_validity,
                       int32_t** out_offsets, uint8_t** out_data);

/* Whole reference split of one batch at a time, for the CPU baseline: pids,
 * reducer offsets and per-partition copies into growable partition buffers,
 * exactly the per-batch loop of Splitter::Split.  Returns the rows split. */
typedef struct gzo_splitter gzo_splitter;
gzo_splitter* gzo_splitter_make(int ncols, const int32_t* types, int P, const int32_t* keys,
                                int nkeys, int32_t seed);
void gzo_splitter_split(gzo_splitter* s, const gzo_batch* b);
/* gzo_splitter_split over batches[0..nb) in order (a task's batch loop, timed
 * without a caller between batches) */
void gzo_splitter_split_many(gzo_splitter* s, const gzo_batch* batches, int nb);
int64_t gzo_splitter_rows(gzo_splitter* s, int p);
void gzo_splitter_free(gzo_splitter* s);

/* ---- hash aggregate ---- */
typedef struct gzo_agg gzo_agg;
/* kinds: as GZ_ACTION_* of include/gazelle_amd.h.  Several GROUPBY actions
 * group by all their columns together (an UnsafeRow key of at most
 * GZO_MAX_KEYS columns); NULL if there are more.  A utf8 key column groups by
 * its bytes. */
#define GZO_MAX_KEYS 8
gzo_agg* gzo_agg_make(int ncols, const int32_t* types, int nactions, const int32_t* kinds,
                      const int32_t* in0, const int32_t* in1, const int32_t* nan_check);
int gzo_agg_process(gzo_agg* a, const gzo_batch* b);
/* gzo_agg_process over batches[0..nb) in order */
int gzo_agg_process_many(gzo_agg* a, const gzo_batch* batches, int nb);
int64_t gzo_agg_num_groups(gzo_agg* a);
int gzo_agg_num_results(gzo_agg* a);
int gzo_agg_result_type(gzo_agg* a, int col);
/* values: element width of the result type; valid: one byte per group */
void gzo_agg_result(gzo_agg* a, int col, void* values, uint8_t* valid);
/* A utf8 key's result column: its byte count (-1 if the column is not one),
 * then offsets (num_groups + 1), bytes and validity. */
int64_t gzo_agg_result_bytes(gzo_agg* a, int col);
void gzo_agg_result_binary(gzo_agg* a, int col, int32_t* offsets, uint8_t* data, uint8_t* valid);
void gzo_agg_free(gzo_agg* a);

/* ---- hash join ---- */
typedef struct gzo_rel gzo_rel;
gzo_rel* gzo_rel_make(int32_t key_type, int32_t builder_type);
/* key column of one build batch (values + optional validity) */
void gzo_rel_append(gzo_rel* r, const void* keys, const uint8_t* validity, int64_t n);
void gzo_rel_finish(gzo_rel* r);
int64_t gzo_rel_num_keys(gzo_rel* r);
int gzo_rel_has_null(gzo_rel* r);
/* Probe one batch.  Inner: writes (probe_row, build_row) pairs, build_row is
 * the global build row index (batch order then row order).  Semi/anti: writes
 * probe_row only (build_row untouched).  Returns the number of output rows;
 * call first with out=NULL to size. */
int64_t gzo_probe(gzo_rel* r, int join_type, int null_aware, const void* keys,
                  const uint8_t* validity, int64_t n, int64_t* out_probe, int64_t* out_build);
void gzo_rel_free(gzo_rel* r);

#ifdef __cplusplus
}
#endif
#endif
