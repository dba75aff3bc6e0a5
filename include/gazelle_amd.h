/*
 * gazelle_amd.h -- C ABI of the MI355X (gfx950) columnar operator hot path:
 * the shuffle hash Splitter, the hash aggregate and the hash join.
 *
 * Every entry point is plain C: handles are opaque pointers, column data cross
 * the boundary as device addresses plus sizes, and no exception leaves the
 * library.  Each function returns a status code whose values are those of
 * arrow::StatusCode; gz_last_error() returns the message of the last failure on
 * the calling thread.
 *
 * Record batches use the JNI buffer convention of the reference
 * (native-sql-engine/cpp/src/jni/jni_common.h:309-340, MakeArrayData): for each
 * column in schema order, a validity buffer (address 0 or size 0 = no nulls),
 * then an int32 offsets buffer for utf8/binary columns, then the values buffer.
 * A null-typed column contributes two (empty) buffers.  Buffers carry no
 * offset: row 0 is bit 0 / element 0.
 *
 * Input buffers are borrowed in stream order: a call validates its arguments
 * and queues device work on the context's stream, and that work reads the
 * batch after the call returns.  The caller keeps every input buffer allocated
 * and unchanged until the stream has passed the call: after
 * gz_context_synchronize, or in work that the caller orders after it on the
 * same stream (gz_context_set_stream).  The reference's Splitter copies a batch into pool
 * memory during Split (splitter.cc:336-356); here the copy into the partition
 * segments is itself the queued work.
 *
 * Thread safety: a context may be shared by threads; a Splitter, aggregate,
 * hash relation or probe handle is used by one thread at a time (one Spark task
 * per handle, as in the reference).
 */
#ifndef GAZELLE_AMD_H_
#define GAZELLE_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (values of arrow::StatusCode) ---------------------------- */
#define GZ_OK 0
#define GZ_OUT_OF_MEMORY 1
#define GZ_KEY_ERROR 2
#define GZ_TYPE_ERROR 3
#define GZ_INVALID 4
#define GZ_IO_ERROR 5
#define GZ_CAPACITY_ERROR 6
#define GZ_INDEX_ERROR 7
#define GZ_UNKNOWN_ERROR 9
#define GZ_NOT_IMPLEMENTED 10
#define GZ_HIP_ERROR 100 /* a HIP runtime call failed; message has the HIP error */

/* ---- column types (values of arrow::Type::type) ---------------------------- */
#define GZ_TYPE_NA 0
#define GZ_TYPE_BOOL 1
#define GZ_TYPE_UINT8 2
#define GZ_TYPE_INT8 3
#define GZ_TYPE_UINT16 4
#define GZ_TYPE_INT16 5
#define GZ_TYPE_UINT32 6
#define GZ_TYPE_INT32 7
#define GZ_TYPE_UINT64 8
#define GZ_TYPE_INT64 9
#define GZ_TYPE_FLOAT 11
#define GZ_TYPE_DOUBLE 12
#define GZ_TYPE_STRING 13
#define GZ_TYPE_BINARY 14
#define GZ_TYPE_DATE32 16
#define GZ_TYPE_DATE64 17
#define GZ_TYPE_TIMESTAMP 18
#define GZ_TYPE_DECIMAL128 23

/* ---- context ----------------------------------------------------------------
 * A context owns one device, one HIP stream and an HBM budget.  All device
 * memory the library allocates (partition buffers, hash tables, aggregate
 * states, join outputs) is charged to the budget; an allocation beyond it fails
 * with GZ_OUT_OF_MEMORY (the analogue of the Splitter's capped memory pool,
 * reference shuffle_split_test.cc:47-95).
 */
typedef struct gz_context gz_context;

/* Replaces: the per-task arrow::MemoryPool handed to the kernels
 * (jni_wrapper.cc:1081-1085).  hbm_budget_bytes <= 0 means "no cap". */
int gz_context_create(int device, int64_t hbm_budget_bytes, gz_context** out);
/* Frees the context and all device memory of its pool: every handle made on it
 * (splitter, aggregate, relation, probe) is released first or not at all. */
int gz_context_destroy(gz_context* ctx);
/* Launch on a caller-owned HIP stream (hipStream_t) instead of the context's. */
int gz_context_set_stream(gz_context* ctx, void* hip_stream);
void* gz_context_stream(gz_context* ctx);
int gz_context_synchronize(gz_context* ctx);
int64_t gz_context_bytes_in_use(gz_context* ctx);
int64_t gz_context_peak_bytes(gz_context* ctx);
/* Kernel timing with HIP events on the launch stream (off by default). */
int gz_context_set_profiling(gz_context* ctx, int enabled);
/* Accumulated time (ms) and launch count of a named kernel family: "split_count",
 * "split_scatter", "agg_update", "join_build", "join_probe", ... */
int gz_context_kernel_time(gz_context* ctx, const char* name, double* total_ms,
                           int64_t* launches);
int gz_context_reset_kernel_times(gz_context* ctx);
const char* gz_last_error(void);

/* One record batch in the JNI buffer convention (device addresses). */
typedef struct {
  int64_t num_rows;
  const int64_t* buf_addrs;
  const int64_t* buf_sizes;
  int32_t num_bufs;
} gz_batch;

/* ---- shuffle Splitter -------------------------------------------------------
 * Replaces sparkcolumnarplugin::shuffle::Splitter (cpp/src/shuffle/splitter.h:38-295)
 * as driven by ShuffleSplitterJniWrapper (jni_wrapper.cc:1031-1234).
 */
typedef struct gz_splitter gz_splitter;

typedef struct {
  int32_t buffer_size;        /* SplitOptions.buffer_size (type.h:41), default 4096 */
  int32_t prefer_spill;       /* SplitOptions.prefer_spill (type.h:45), default 1 */
  int64_t offheap_per_task;   /* SplitOptions.offheap_per_task (type.h:40) */
  int32_t tile_rows;          /* rows per workgroup tile: 4096, 2048 or 1024; 0 = the
                                 default, 2048 (1024-row tiles are too small for LDS
                                 staging above 1024 partitions and become 2048) */
  int32_t hash_seed;          /* 42, the seed Spark uses (splitter.cc:1405) */
  int32_t append_row_ids;     /* 1 = add a last output column, int32, holding each
                                 row's index among the rows of its split call */
  int32_t batch_compress_threshold; /* SplitOptions.batch_compress_threshold
                                 (type.h:43), default 256: with LZ4_FRAME, record
                                 batches of at most this many rows are written
                                 uncompressed (splitter.cc:586-590) */
} gz_split_options;

void gz_split_options_defaults(gz_split_options* opts);

/* Replaces Splitter::Make(short_name, schema, num_partitions, expr_vector, options)
 * (splitter.h:55-58, splitter.cc:190-214).  short_name: "hash", "rr", "range",
 * "single".  The Gandiva key expressions become column indices (field(k)
 * references, as ColumnarShuffleExchangeExec.scala:307-321 builds them).  For
 * "range" column 0 is the int32 partition id and is removed from the output
 * (splitter.cc:1497-1500).  Limits: up to 128 columns and 16 hash keys; 1 to
 * 65536 partitions, as the reference (splitter.cc:218-221; more is Invalid).
 * Up to 2048 partitions, with num_partitions x (1 + utf8/binary columns)
 * counters of at most 128 KiB, the LDS-staged kernels split; past that, a
 * sorted path (partition ids, a stable radix sort by id, gathers) gives the
 * same layout.  A call's batches sit in tiles of 4096 rows on the sorted path,
 * and at most 2^20 tiles fit one call (CapacityError beyond). */
int gz_split_make(gz_context* ctx, const char* short_name, int32_t num_cols,
                  const int32_t* col_types, int32_t num_partitions,
                  const int32_t* key_cols, int32_t num_keys,
                  const gz_split_options* opts, gz_splitter** out);

/* Replaces Splitter::Split(const RecordBatch&) (splitter.cc:387-393) /
 * ShuffleSplitterJniWrapper.split (jni_wrapper.cc:1156-1195): one batch (the
 * plain per-batch path).  The call checks the buffers and copies them, with one
 * kernel on the context's stream, into a device arena of the splitter (up to
 * 128 MiB, about 32 batches of the first one's size); the staged batches are
 * split together, as by gz_split_batches, when the arena fills or at the next
 * gz_split_batches / gz_split_stop.  A batch above a quarter of the arena, the
 * "range" splitter (its bad partition ids are the call's error) and
 * append_row_ids run at once.  The buffers are borrowed in stream order (see
 * the top of this file): the copy is the queued work that reads them. */
int gz_split_batch(gz_splitter* h, int64_t num_rows, const int64_t* buf_addrs,
                   const int64_t* buf_sizes, int32_t num_bufs);

/* Batched form: many (uneven) batches split in one launch sequence; the result
 * is the same as calling gz_split_batch on each in order. */
int gz_split_batches(gz_splitter* h, int32_t num_batches, const gz_batch* batches);

/* Replaces Splitter::Stop (splitter.cc:395-437): completes all device work.
 * The partitioned output stays resident and is read with the calls below. */
int gz_split_stop(gz_splitter* h);

/* Replaces Splitter::PartitionLengths in rows (splitter.h:114); out[P]. */
int gz_split_partition_rows(gz_splitter* h, int64_t* out);
int gz_split_num_partitions(gz_splitter* h);
/* Output columns (the "range" pid column excluded) and their types. */
int gz_split_num_output_columns(gz_splitter* h);

/* The output is a list of segments (one per gz_split_batches call or run of
 * staged gz_split_batch calls, or per gz_split_batch that ran at once).
 * Inside a segment, the rows of partition p are contiguous, in input
 * order, at row positions [row_start[p], row_start[p+1]).  Per-partition
 * concatenation of the segments in order is the reference's partition stream.
 *   values:   fixed-width values at global row positions (bool: bitmap)
 *   validity: bitmap at global row positions, trailing bits 1, or NULL when
 *             the column has no validity in the input
 *   offsets:  utf8/binary only; partition p's int32 offsets, starting at 0,
 *             occupy entries [row_start[p] + p, row_start[p+1] + p]
 *   data:     utf8/binary bytes; partition p's run starts at byte_start[p]
 * All pointers are device addresses owned by the splitter. */
int gz_split_num_segments(gz_splitter* h);
int gz_split_segment_rows(gz_splitter* h, int32_t seg, int64_t* row_start /* P+1 */);
int gz_split_segment_column(gz_splitter* h, int32_t seg, int32_t col, void** values,
                            void** validity, void** offsets, void** data,
                            int64_t* byte_start /* P+1, may be NULL */,
                            int64_t* null_count);
/* Per-batch, per-partition row counts of a segment (counts[b * P + p]); the host
 * side uses them to reproduce the reference's output batch boundaries. */
int gz_split_segment_batch_counts(gz_splitter* h, int32_t seg, int32_t* num_batches,
                                  int32_t* counts /* may be NULL */);
/* Replaces Splitter::TotalComputePidTime (splitter.h:112), in nanoseconds;
 * measured only when profiling is enabled on the context. */
int64_t gz_split_compute_pid_time(gz_splitter* h);
int gz_split_release(gz_splitter* h);

/* Body compression of the partition writer, as arrow::Compression::type values. */
#define GZ_COMPRESSION_UNCOMPRESSED 0
#define GZ_COMPRESSION_LZ4_FRAME 6
/* Replaces Splitter::SetCompressType (splitter.cc:371-385) for the partition
 * writer: GZ_COMPRESSION_LZ4_FRAME makes gz_split_ipc_prepare compress every
 * non-empty body buffer on the device (int64 uncompressed length + one LZ4
 * frame of independent 32 KiB blocks, a block that does not shrink stored) and
 * mark record batches with BodyCompression{LZ4_FRAME, BUFFER}, as Arrow's IPC
 * writer does with that codec; other codecs are NotImplemented. */
int gz_split_set_compress_type(gz_splitter* h, int32_t codec);

/* Partition writer: replaces the data file Splitter::Stop writes through its
 * PartitionWriters (splitter.cc:70-185, :395-437) with the default
 * compression_type UNCOMPRESSED (shuffle/type.h:44).  The stream holds, partition
 * after partition, each non-empty partition's Arrow IPC stream: the caller's
 * encapsulated schema message when write_schema (SplitOptions.write_schema,
 * type.h:47; the bytes the JNI layer receives as schema_arr), the record batches
 * cut as CacheRecordBatch cuts them (:462-608, buffer policy :889-925), and the
 * end-of-stream marker; an empty partition writes nothing.  prepare (after
 * gz_split_stop) lays the stream out and returns its size and the per-partition
 * lengths (Splitter::PartitionLengths, splitter.h:114); write assembles it into
 * a device buffer of at least total_bytes.  With offheap_per_task > 0 the
 * buffer rows come from CalculateSplitBatchSize (gz_split_batch_rows). */
int gz_split_ipc_prepare(gz_splitter* h, const uint8_t* schema_msg, int64_t schema_len, int32_t write_schema,
                         int64_t* total_bytes, int64_t* partition_lengths /* P */);
int gz_split_ipc_write(gz_splitter* h, void* dst_device, int64_t dst_bytes);

/* Replaces Splitter::RawPartitionLengths (splitter.h:120-122), after
 * gz_split_ipc_prepare: per partition the in-memory bytes (batch_nbytes,
 * splitter.cc:438-453) of the record batches the reference caches, including the
 * validity buffers its partition buffers hold (allocated while a column's
 * input_has_null_ flag is set, or when an input batch with nulls adds rows). */
int gz_split_raw_partition_lengths(gz_splitter* h, int64_t* out /* P */);
/* Rows of a partition buffer used for the record batch cuts: buffer_size, or with
 * offheap_per_task > 0 CalculateSplitBatchSize (splitter.cc:800-853) from the first
 * input batch's row size (a binary column empty in that batch: NotImplemented). */
int gz_split_batch_rows(gz_splitter* h, int64_t* out);

/* The metrics of ShuffleSplitterJniWrapper.stop's SplitResult (jni_wrapper.cc:
 * 1197-1225; Splitter::TotalComputePidTime ... TotalBytesSpilled, splitter.h:105-118). */
typedef struct {
  int64_t total_compute_pid_time; /* ns of the partition id passes on the device */
  int64_t total_write_time;       /* ns laying out and assembling the data file */
  int64_t total_spill_time;       /* 0: the device split keeps partitions in HBM */
  int64_t total_compress_time;    /* ns compressing body buffers (LZ4_FRAME) */
  int64_t total_bytes_written;    /* bytes of the data file (after gz_split_ipc_prepare) */
  int64_t total_bytes_spilled;    /* 0 */
} gz_split_metrics;
int gz_split_get_metrics(gz_splitter* h, gz_split_metrics* out);
/* The encapsulated IPC message (continuation token, metadata length, Arrow
 * Message flatbuffer with a RecordBatch header, MetadataVersion V5, padding) of
 * a record batch with the given field nodes (length, null_count pairs) and
 * buffers (offset, length pairs); host only.  *out_length receives the size;
 * out may be NULL to query it. */
int gz_ipc_record_batch_message(int64_t num_rows, int32_t num_nodes, const int64_t* nodes, int32_t num_buffers,
                                const int64_t* buffers, int64_t body_length, uint8_t* out, int64_t out_capacity,
                                int64_t* out_length);

/* Shuffle read: replaces the reading side of the shuffle
 * (SchemaAwareArrowCompressedStreamReader + ShuffleDecompressionJniWrapper.make /
 * .decompress, jni_wrapper.cc:1236-1322, DecompressBuffers jni_common.h:471-560).
 * Parses an Arrow IPC stream (host copy in `stream`: one partition's bytes of a
 * data file) and lays out each record batch in the gz_batch buffer convention.
 * Buffers of uncompressed batches are addressed inside the device copy of the
 * same bytes, which starts at device_base (zero copy).  Buffers of LZ4_FRAME
 * compressed batches (the Spark "lz4" codec) are addressed in a decode arena
 * starting at arena_base, of *arena_bytes bytes, which gz_ipc_decompress fills;
 * a buffer whose length prefix is -1 is stored raw and stays in the stream.  A
 * zero-length buffer gives address 0.
 * Per batch: batch_rows[b], then num_bufs addresses and sizes at
 * buf_addrs/buf_sizes[b * num_bufs ...], where num_bufs counts 3 buffers per
 * STRING/BINARY column and 2 per other column (NA: two empty buffers).
 * At most max_batches are written; *num_batches receives the stream's count.
 * Schema messages are skipped; dictionaries, ZSTD and LZ4 frames with a
 * dictionary id are NotImplemented; malformed streams are IOError.
 * Host-only: no device work. */
int gz_ipc_read_batches(const uint8_t* stream, int64_t len, int64_t device_base, int64_t arena_base,
                        int32_t num_cols, const int32_t* col_types, int32_t max_batches, int64_t* batch_rows,
                        int64_t* buf_addrs, int64_t* buf_sizes, int32_t* num_batches, int64_t* arena_bytes);

/* Decodes every LZ4_FRAME compressed buffer of the stream (host copy `stream`,
 * device copy at device_base) into `arena` (device, arena_capacity >= the
 * *arena_bytes gz_ipc_read_batches reports), on the context's stream, and
 * waits for it.  Like DecompressBuffer (jni_common.h:471-492): corrupt LZ4 data
 * is IOError; a buffer that decodes to another length than its prefix is
 * Invalid ("Failed to fully decompress buffer ...").  Frame and block
 * checksums, when a frame has them, are skipped, not verified. */
int gz_ipc_decompress(gz_context* ctx, const uint8_t* stream, int64_t len, int64_t device_base, int32_t num_cols,
                      const int32_t* col_types, void* arena, int64_t arena_capacity);

/* gz_ipc_decompress without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The caller passes the same inputs and the output arena; the decode's task
 * table, block table and decoded lengths live in a library-owned 8 MiB device
 * array per GPU (allocated with the library, kept until the process ends), so
 * a stream may hold at most as many LZ4 buffers and blocks as fit in it
 * (32 B per compressed buffer + 16 B per LZ4 block, each table 256-byte
 * aligned): more is CapacityError.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other threads, other streams) race. */
int gz_ipc_decompress_noalloc(gz_context* ctx, const uint8_t* stream, int64_t len, int64_t device_base,
                              int32_t num_cols, const int32_t* col_types, void* arena, int64_t arena_capacity);

/* ---- hash aggregate ----------------------------------------------------------
 * Replaces HashAggregateKernel + HashAggregateResultIterator<number>
 * (ext/hash_aggregate_kernel.cc:51-928, kernels_ext.h:285-313).  The Gandiva
 * action nodes become an action list with the reference's action names.
 */
typedef struct gz_agg gz_agg;

#define GZ_ACTION_GROUPBY 0      /* action_groupby (UniqueAction, actions_impl.cc:132-290) */
#define GZ_ACTION_COUNT 1        /* action_count (CountAction, :293-454) */
#define GZ_ACTION_COUNT_LITERAL 2 /* action_countLiteral_<n> (:634-755) */
#define GZ_ACTION_SUM 3          /* action_sum (SumAction, :1982-2143) */
#define GZ_ACTION_AVG 4          /* action_avg (AvgAction, :2656-2829) */
#define GZ_ACTION_MIN 5          /* action_min[_true] (MinAction, :756-1041) */
#define GZ_ACTION_MAX 6          /* action_max[_true] (MaxAction, :1372-1654) */
#define GZ_ACTION_SUM_COUNT 7    /* action_sum_count (SumCountAction, :3012-3200) */
#define GZ_ACTION_AVG_BY_COUNT 8 /* action_avgByCount (AvgByCountAction, :3746-3925) */

typedef struct {
  int32_t kind;       /* GZ_ACTION_* */
  int32_t input0;     /* input column index, -1 if none */
  int32_t input1;     /* second input (avgByCount count column), -1 if none */
  int32_t nan_check;  /* action_min_true / action_max_true */
} gz_agg_action;

typedef struct {
  int32_t placement;      /* 0 = chosen per instance from its estimated groups (LDS
                             first level, partitioned, or HBM table), 1 = LDS group
                             table, 2 = HBM table, 3 = partitioned */
  int32_t tile_rows;      /* rows per update tile; 0 = 4096 (512 for per-call updates
                             of fewer than 2^20 rows) */
  double load_factor;     /* HBM table max load, 0 = 0.5 */
  int64_t expected_groups;/* sizing hint, 0 = unknown */
} gz_agg_options;

void gz_agg_options_defaults(gz_agg_options* opts);

/* Replaces HashAggregateKernel::Make + MakeResultIterator (hash_aggregate_kernel.cc:1238-1260,
 * :147-224).  A plan has one GZ_ACTION_GROUPBY (an integer or utf8 key), none
 * (the global aggregate), or up to 8 of them: then rows group by all those
 * columns together with the reference's UnsafeRow key equality (the same keys
 * null, identical bytes in the others: hash_aggregate_kernel.cc:970-1000), each
 * key fixed-width numeric (integers, dates, timestamps, float, double) except
 * that one, in any position, may be utf8 (two utf8 keys are NotImplemented: the
 * UnsafeRow joins their bytes with no length).  A utf8 key groups by its bytes,
 * the null key its own group (:1001-1020).  The gz_agg_set_* forms take one
 * integer key only. */
int gz_agg_make(gz_context* ctx, int32_t num_cols, const int32_t* col_types,
                int32_t num_actions, const gz_agg_action* actions,
                const gz_agg_options* opts, gz_agg** out);
/* Replaces ResultIterator::ProcessAndCacheOne (hash_aggregate_kernel.cc:805-893). */
int gz_agg_process(gz_agg* h, int64_t num_rows, const int64_t* buf_addrs,
                   const int64_t* buf_sizes, int32_t num_bufs);
int gz_agg_process_batches(gz_agg* h, int32_t num_batches, const gz_batch* batches);
/* Number of groups so far (max_group_id + 1), completes device work. */
int gz_agg_num_groups(gz_agg* h, int64_t* out);
/* Number of result columns and their types (sum_count yields two). */
int gz_agg_num_result_columns(gz_agg* h);
int gz_agg_result_type(gz_agg* h, int32_t col);
/* Replaces HasNext/Next (hash_aggregate_kernel.cc:895-913): finalizes the
 * group order (first-seen, as the reference's SparseHashMap assigns ids) and
 * exposes device result columns of num_groups rows: values (element width of
 * the result type) and validity as one byte per group (1 = valid). */
int gz_agg_finish(gz_agg* h);
int gz_agg_result_column(gz_agg* h, int32_t col, void** values, void** valid_bytes);
/* A utf8 result column (a utf8 group-by key): int32 offsets (num_groups + 1),
 * the bytes, validity one byte per group; gz_agg_result_column gives the
 * offsets as its values.  Replaces the StringType UniqueAction's Finish
 * (actions_impl.cc:132-290). */
int gz_agg_result_binary(gz_agg* h, int32_t col, void** offsets, void** data, void** valid_bytes);
/* The result column's validity as an Arrow bitmap (LSB first, 1 = valid, padded
 * with 1 bits to a multiple of 64 rows), as ResultIterator::Next's RecordBatch
 * holds it (result_iterator.h:37-82): a device buffer owned by the handle, packed
 * on the context's stream, valid until the next finish or release. */
int gz_agg_result_bitmap(gz_agg* h, int32_t col, void** bitmap);
int gz_agg_release(gz_agg* h);

/* Aggregate sets: many independent GROUP BYs (instances) sharing one plan,
 * processed in one launch sequence -- the batched form.  A set is a gz_agg
 * handle; gz_agg_finish and gz_agg_release apply to it.  No counterpart in the
 * reference, which runs one HashAggregateResultIterator per task. */
int gz_agg_set_make(gz_context* ctx, int32_t num_cols, const int32_t* col_types,
                    int32_t num_actions, const gz_agg_action* actions,
                    const gz_agg_options* opts, int32_t num_instances, gz_agg** out);
int gz_agg_set_process(gz_agg* h, int32_t num_batches, const gz_batch* batches,
                       const int32_t* instance_of_batch);
int gz_agg_set_num_groups(gz_agg* h, int32_t inst, int64_t* out);
int gz_agg_set_result_column(gz_agg* h, int32_t inst, int32_t col, void** values,
                             void** valid_bytes);
/* Placement of an instance in the current plan: 1 = LDS first level, 2 = HBM,
 * 3 = partitioned. */
int gz_agg_set_placement(gz_agg* h, int32_t inst, int32_t* out);
/* gz_agg_result_bitmap of one instance. */
int gz_agg_set_result_bitmap(gz_agg* h, int32_t inst, int32_t col, void** bitmap);

/* Planned, replayable form of a set.  gz_agg_set_prepare sizes every
 * instance's table for the given batches (cardinality estimate, allocation,
 * placement) and records the launch plan; gz_agg_set_run executes it (table
 * reset and update kernels only: no allocation, no host synchronization, no
 * memset); gz_agg_set_capture records that sequence into a HIP graph and
 * gz_agg_set_launch replays it (or runs it when nothing was captured).  After a
 * run or a launch, gz_agg_finish produces that run's results (and reports
 * GZ_CAPACITY_ERROR if an estimated table overflowed).  The batches must stay
 * valid while the plan is used.  No counterpart in the reference. */
int gz_agg_set_prepare(gz_agg* h, int32_t num_batches, const gz_batch* batches,
                       const int32_t* instance_of_batch);
int gz_agg_set_run(gz_agg* h);
int gz_agg_set_capture(gz_agg* h);
int gz_agg_set_launch(gz_agg* h);
/* Autotuner: a coordinate descent over the plan's knobs (update tile rows, LDS
 * placement threshold, groups per partition -- the partitioned fan-out --,
 * table load factor, workgroup size of the partitioned update), every
 * candidate prepared and timed on the given batches; the fastest stays
 * prepared and *chosen receives its number in the order tried (0: the default
 * plan).  No counterpart in the reference. */
int gz_agg_set_autotune(gz_agg* h, int32_t num_batches, const gz_batch* batches,
                        const int32_t* instance_of_batch, int32_t* chosen);

/* ---- hash join -----------------------------------------------------------------
 * Build replaces HashRelationKernel (ext/hash_relation_kernel.cc:50-344) and
 * HashRelation (codegen/common/hash_relation.h:123-645); probe replaces
 * ConditionedProbeKernel / ConditionedProbeResultIterator
 * (ext/conditioned_probe_kernel.cc:54-1772).
 */
typedef struct gz_relation gz_relation;
typedef struct gz_probe gz_probe;

#define GZ_JOIN_INNER 0
#define GZ_JOIN_ANTI 2
#define GZ_JOIN_SEMI 3

/* builder_type 1 = all rows (unsafeHashMap append), 3 = skip duplicate keys
 * (InsertSkipDup, for semi joins) -- hash_relation_kernel.cc:79-99. */
int gz_relation_make(gz_context* ctx, int32_t num_cols, const int32_t* col_types,
                     int32_t key_col, int32_t builder_type, gz_relation** out);
/* Replaces HashRelationKernel::Evaluate (hash_relation_kernel.cc:170-198): the
 * batch is copied into relation-owned HBM. */
int gz_relation_append(gz_relation* r, int64_t num_rows, const int64_t* buf_addrs,
                       const int64_t* buf_sizes, int32_t num_bufs);
int gz_relation_append_batches(gz_relation* r, int32_t num_batches, const gz_batch* batches);
/* Replaces FinishInternal (hash_relation_kernel.cc:200-281). */
int gz_relation_finish(gz_relation* r);
int gz_relation_num_keys(gz_relation* r, int64_t* out);
int gz_relation_has_null_key(gz_relation* r, int32_t* out);
int gz_relation_release(gz_relation* r);

/* Replaces ConditionedProbeKernel::Make + SetDependencies
 * (conditioned_probe_kernel.cc:1745-1772, :380-465).  result_cols lists the
 * output columns: side 0 = build ("left" in the reference), side 1 = probe
 * ("right"), with the column index on that side. */
int gz_probe_make(gz_relation* r, int32_t num_probe_cols, const int32_t* probe_types,
                  int32_t probe_key_col, int32_t join_type, int32_t null_aware_anti,
                  int32_t num_result_cols, const int32_t* result_sides,
                  const int32_t* result_indices, gz_probe** out);
/* Replaces ConditionedProbeResultIterator::Process (:468-519): probes one batch
 * and materializes one output batch in probe-row order (matches of one probe row
 * in build insertion order).  Output columns are device buffers owned by the
 * probe handle, valid until the next call: values + validity bytes (1 = valid);
 * for utf8 columns offsets + data. */
int gz_probe_process(gz_probe* p, int64_t num_rows, const int64_t* buf_addrs,
                     const int64_t* buf_sizes, int32_t num_bufs, int64_t* out_rows);
/* Batched form: probes many batches in one launch sequence; the output is the
 * concatenation of the per-batch outputs, out_batch_rows[b] rows each. */
int gz_probe_process_batches(gz_probe* p, int32_t num_batches, const gz_batch* batches,
                             int64_t* out_rows, int64_t* out_batch_rows);
int gz_probe_result_column(gz_probe* p, int32_t col, void** values, void** valid_bytes,
                           void** offsets, void** data);
/* The probe's stream and device memory: the relation's context by default.
 * SetDependencies hands one HashRelation to the probe of every task that joins
 * against it (a broadcast relation, conditioned_probe_kernel.cc:380-465); task
 * threads probing it at once give each probe its own context, so their calls
 * run on separate streams.  Call before the probe's first process call; it
 * waits for the relation's stream, so the finished relation is visible to the
 * new one.  The context must be on the relation's device and outlive the probe. */
int gz_probe_set_context(gz_probe* p, gz_context* ctx);

/* Several join key columns (HashRelationKernel / ConditionedProbe with
 * UnsafeRow keys: hash_relation_kernel.cc:264-277, hash_relation.h:176-198,
 * conditioned_probe_kernel.cc:609-640), or one utf8 key.  A build row with a
 * null key is never inserted; a probe row matches build rows whose keys have
 * identical bytes (fixed-width numeric keys, and a utf8 key only as the last
 * one); the probe's key types must equal the build's.  Build and probe take
 * batches, results and release as the one-key forms; one fixed-width key
 * falls back to gz_relation_make / gz_probe_make.  A null-aware anti join with
 * several keys is the plain anti join, as in the reference, whose UnsafeRow anti
 * probe never reads the null-aware flag (conditioned_probe_kernel.cc:957-982);
 * one utf8 key keeps the one-key null-aware rules (a null build key is the
 * relation's null key, a null probe key is null: :1016-1031).  The fused
 * join -> aggregate over such a relation is NotImplemented. */
int gz_relation_make_keys(gz_context* ctx, int32_t num_cols, const int32_t* col_types, int32_t num_keys,
                          const int32_t* key_cols, int32_t builder_type, gz_relation** out);
int gz_probe_make_keys(gz_relation* r, int32_t num_probe_cols, const int32_t* probe_types, int32_t num_keys,
                       const int32_t* probe_key_cols, int32_t join_type, int32_t null_aware_anti,
                       int32_t num_result_cols, const int32_t* result_sides, const int32_t* result_indices,
                       gz_probe** out);
/* Spark's join-key "normalize" (NormalizeFloatingNumbers; the reference's
 * normalize_nan_zero, precompile/gandiva.h:214-222, applied to the key
 * expressions, tests/arrow_compute_test_join_wocg.cc:1349-1462): for each key
 * with normalize[j] != 0 (float or double only), every NaN becomes one NaN and
 * |x| < 1e-7 becomes 0.0 before the key is hashed and compared, on the build
 * and the probe side.  Output columns keep the original values.  With several
 * keys NaN then matches NaN (key rows compare by bytes); one key still compares
 * with == after normalisation, so a NaN key never matches (hash_relation.h:
 * 368-383).  num_keys is the relation's key count; called before the first
 * append. */
int gz_relation_normalize_keys(gz_relation* r, int32_t num_keys, const int32_t* normalize);
/* The last call's output column validity as an Arrow bitmap (as
 * gz_agg_result_bitmap), owned by the handle, valid until the next call. */
int gz_probe_result_bitmap(gz_probe* p, int32_t col, void** bitmap);
int gz_probe_release(gz_probe* p);

/* ---- fused join -> aggregate -------------------------------------------------
 * Replaces whole-stage codegen's inner probe feeding a hash aggregate
 * (WholeStageCodeGenKernel, ext/whole_stage_codegen_kernel.cc:48-330, with
 * conditionedProbeArraysInner and hashAggregateArrays children; golden
 * WSCGTestInnerJoinWithGroupbyAggregate, tests/arrow_compute_test_wscg.cc:4838-5000).
 * The aggregate reads num_inputs join result columns (input_sides[c] 0 = build
 * relation column, 1 = probe column; input_indices[c] the column on that side)
 * and groups by the join key (its GROUPBY input must be the probe key column or
 * the relation's key column); the joined rows are never materialized.  Returns
 * a gz_agg handle: gz_agg_num_groups / finish / result_column / release apply,
 * with groups in the order of the join output (first probe row per key).  ctx
 * null = the relation's context.  Integer key types; no join condition (the
 * reference's condition expression is not evaluated). */
int gz_join_agg_make(gz_relation* rel, gz_context* ctx, int32_t num_probe_cols, const int32_t* probe_types,
                     int32_t probe_key_col, int32_t num_inputs, const int32_t* input_sides,
                     const int32_t* input_indices, int32_t num_actions, const gz_agg_action* actions,
                     const gz_agg_options* opts, gz_agg** out);
/* Probes the batches and updates the aggregate (ProcessAndCacheOne over a
 * probe batch of the fused stage). */
int gz_join_agg_process_batches(gz_agg* h, int32_t num_batches, const gz_batch* batches);

/* ---- serving -----------------------------------------------------------------
 * A fragment server: query fragments (an aggregate, a join followed by an
 * aggregate, a shuffle split) arrive on a timeline and are admitted, first come
 * first served, when one of the server's slots is idle and the fragment's HBM
 * estimate fits what the admitted fragments leave of the budget.  Each slot has
 * its own context (HIP stream and memory pool), so admitted fragments overlap
 * on the device; registered hash relations stay resident and serve the probes
 * of every join fragment.  No counterpart in the reference, where one Spark
 * task drives one kernel instance; the operators are the ones above.
 */
typedef struct gz_server gz_server;

typedef struct {
  int32_t num_cols;
  const int32_t* col_types;
  int32_t num_actions;
  const gz_agg_action* actions;
} gz_agg_plan;

typedef struct {
  int32_t num_cols;
  const int32_t* col_types;
  int32_t num_partitions;
  const int32_t* key_cols;
  int32_t num_keys;
} gz_split_plan;

typedef struct {
  int32_t num_probe_cols;
  const int32_t* probe_types;
  int32_t probe_key_col;
  int32_t join_type;
  int32_t num_result_cols;
  const int32_t* result_sides;
  const int32_t* result_indices;
} gz_probe_plan;

#define GZ_FRAG_AGGREGATE 0
#define GZ_FRAG_JOIN_AGGREGATE 1
#define GZ_FRAG_SPLIT 2

typedef struct {
  int32_t kind;             /* GZ_FRAG_* */
  int32_t relation;         /* join fragments: id from gz_server_add_relation */
  int32_t num_batches;
  const gz_batch* batches;  /* device buffers (the probe side for joins) */
  double arrival_s;         /* arrival, seconds after the run starts */
} gz_fragment;

typedef struct {
  double arrival_s, admitted_s, done_s;  /* seconds after the run starts */
  int64_t rows;                          /* input rows */
  int64_t out_rows;                      /* groups, or rows split */
  int32_t status;                        /* GZ_OK or the failure code */
  int32_t slot;
} gz_fragment_stats;

/* Plans may be null for fragment kinds the server will not see.  The join
 * aggregate reads the join result columns, in result order (fixed-width). */
int gz_server_create(int32_t device, int64_t hbm_budget, int32_t num_slots, const gz_agg_plan* agg,
                     const gz_split_plan* split, const gz_probe_plan* probe, const gz_agg_plan* join_agg,
                     gz_server** out);
/* A finished relation, resident from now on (the caller keeps ownership and
 * releases it after the server). */
int gz_server_add_relation(gz_server* s, gz_relation* r, int32_t* id);
/* Runs fragments sorted by arrival; returns when all are done, with per-fragment
 * times in stats.  keep_results != 0 keeps each output for the accessors below
 * until the next run or gz_server_release_results. */
int gz_server_run(gz_server* s, int32_t num_fragments, const gz_fragment* frags, int32_t keep_results,
                  gz_fragment_stats* stats);
/* Finished aggregate of an aggregate or join fragment / stopped splitter of a
 * split fragment; owned by the server. */
int gz_server_fragment_agg(gz_server* s, int32_t i, gz_agg** out);
int gz_server_fragment_split(gz_server* s, int32_t i, gz_splitter** out);
int gz_server_release_results(gz_server* s);
int gz_server_destroy(gz_server* s);

/* Validity bytes (1 = valid) of n rows to an Arrow bitmap in a caller-provided
 * device buffer of (n + 63) / 64 * 8 bytes (1 bits past row n), on the context's
 * stream: for result columns taken as bytes (gz_agg_result_column,
 * gz_probe_result_column).  No counterpart in the reference, whose builders
 * produce bitmaps. */
int gz_validity_bitmap(gz_context* ctx, const void* valid_bytes, int64_t n, void* bitmap);

#ifdef __cplusplus
}
#endif

#endif /* GAZELLE_AMD_H_ */
