// Host-side internals shared by the C ABI implementation files: status and
// error text, the per-context HBM pool with its budget, kernel timing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/gazelle_amd.h"

namespace gz {

// ---- errors ---------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define GZ_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return ::gz::fail(GZ_HIP_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define GZ_TRY(call)          \
  do {                        \
    int rc_ = (call);         \
    if (rc_ != GZ_OK) return rc_; \
  } while (0)

// ---- type helpers ---------------------------------------------------------------
inline int type_width(int t) {
  switch (t) {
    case GZ_TYPE_UINT8: case GZ_TYPE_INT8: return 1;
    case GZ_TYPE_UINT16: case GZ_TYPE_INT16: return 2;
    case GZ_TYPE_UINT32: case GZ_TYPE_INT32: case GZ_TYPE_FLOAT: case GZ_TYPE_DATE32: return 4;
    case GZ_TYPE_UINT64: case GZ_TYPE_INT64: case GZ_TYPE_DOUBLE: case GZ_TYPE_DATE64:
    case GZ_TYPE_TIMESTAMP: return 8;
    case GZ_TYPE_DECIMAL128: return 16;
    default: return 0;
  }
}
inline bool type_is_binary(int t) { return t == GZ_TYPE_STRING || t == GZ_TYPE_BINARY; }
inline bool type_supported(int t) {
  return t == GZ_TYPE_NA || t == GZ_TYPE_BOOL || type_is_binary(t) || type_width(t) > 0;
}
// Buffers a column contributes in the JNI convention.
inline int type_num_bufs(int t) { return type_is_binary(t) ? 3 : 2; }

// ---- device memory ----------------------------------------------------------------
// A caching pool: blocks are rounded (power of two up to 64 MiB, 2 MiB multiples
// above) and kept on free lists; the budget counts live blocks only, and cached
// blocks are returned to HIP before an allocation would exceed it, when a
// hipMalloc fails, and when the cache outgrows kCacheLimit (so that pools of
// several contexts on one device do not hold its memory between them).
class DevicePool {
 public:
  DevicePool(int device, int64_t budget) : device_(device), budget_(budget) {}
  ~DevicePool();
  int alloc(size_t bytes, void** out);
  void free(void* p);
  int64_t in_use() const { return in_use_; }
  int64_t peak() const { return peak_; }
  int64_t budget() const { return budget_; }
  int trim();

 private:
  static size_t round(size_t bytes);
  static constexpr int64_t kCacheLimit = 16ll << 30;
  int device_;
  int64_t budget_;
  int64_t in_use_ = 0, peak_ = 0, cached_ = 0;
  std::multimap<size_t, void*> free_;
  std::unordered_map<void*, size_t> live_;
  std::mutex mu_;
};

// A device buffer owned by a handle.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// ---- kernel timing ------------------------------------------------------------------
struct KernelStat {
  double ms = 0;
  int64_t launches = 0;
};

}  // namespace gz

struct gz_context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t own_stream = nullptr;
  gz::DevicePool* pool = nullptr;
  int profiling = 0;
  int num_cus = 0;
  std::mutex stat_mu;
  std::map<std::string, gz::KernelStat> stats;
  // events recorded but not yet resolved: (name, start, stop)
  struct Pending {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> event_cache;
  // pinned staging ring for host->device descriptor uploads
  struct Staging {
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
  };
  Staging staging[8];
  int staging_next = 0;
  size_t staging_hwm = 0;  // largest slot size so far: every slot grows to it at once
  std::mutex staging_mu;
};

namespace gz {

int ctx_alloc(gz_context* ctx, size_t bytes, DBuf* out);
void ctx_free(gz_context* ctx, DBuf* b);
int ctx_ensure(gz_context* ctx, DBuf* b, size_t bytes);  // grow-only scratch

// Brackets a launch with HIP events when profiling is on.
class KernelTimer {
 public:
  KernelTimer(gz_context* ctx, const char* name);
  ~KernelTimer();

 private:
  gz_context* ctx_;
  const char* name_;
  hipEvent_t a_ = nullptr, b_ = nullptr;
};
void ctx_resolve_timings(gz_context* ctx, bool wait);

// Stage a host array into device scratch (synchronous with respect to the host
// buffer: it may be reused as soon as this returns).
int upload(gz_context* ctx, DBuf* dst, const void* src, size_t bytes);

// Several uploads as one: the host arrays are packed into one blob, copied to
// the device in one transfer and scattered to their buffers by one kernel, all
// on the context's stream.  add() sizes the destination at once (its address is
// final when add returns) and copies the source; the data lands at flush().
// A destination is added at most once per batch.
class UploadBatch {
 public:
  int add(gz_context* ctx, DBuf* dst, const void* src, size_t bytes);
  int flush(gz_context* ctx, DBuf* blob);  // blob: device scratch the batch is staged in

 private:
  struct Item {
    uint8_t* dst;
    size_t off, bytes;
  };
  std::vector<uint8_t> host_;
  std::vector<Item> items_;
};

// Device view of a finished relation (the join's build side) for kernels
// outside the join: its hash table and its row-indexed column stores.
struct RelationView {
  gz_context* ctx;
  const int64_t* keys;     // [cap] slot keys
  const int32_t* start;    // [cap] first row (unique keys) or rowlist start
  const int32_t* count;    // [cap]
  const int32_t* rowlist;  // rows grouped by key, or null (unique keys)
  uint64_t cap;
  int64_t special_start, special_count, nkeys, rows;
  int key_col, ncols;
  int types[32];
  const uint8_t* valid[32];  // validity bytes per row (1 = valid)
  const void* values[32];    // fixed width: width bytes per row; bool: one byte per row
};
int relation_view(gz_relation* r, RelationView* v);
// Per-batch staging (gz_stage.hip): a per-batch call copies its borrowed buffers
// into the handle's device arena (one gather-copy kernel on the context's
// stream) and the handle runs its batched path over the staged batches later.
struct BatchStage {
  DBuf arena;
  int64_t cap = 0, used = 0;
  std::vector<int64_t> addrs, sizes;  // staged buffer lists, batch after batch
  std::vector<size_t> first;          // index of each batch's first buffer
  std::vector<int64_t> rows;
  std::vector<int32_t> nbufs;
  bool empty() const { return rows.empty(); }
};
// The buffer-size checks of the batched paths for one batch of a schema.
int stage_check(int32_t ncols, const int32_t* types, const gz_batch& b);
// Stages b when it fits (*staged); a batch above a quarter of the arena, or no
// arena under the HBM budget, is left to the caller to run directly.
int stage_add(gz_context* ctx, BatchStage* s, const gz_batch& b, bool* staged);
// gz_batch views of the staged batches (valid until the next stage_add / clear).
void stage_view(const BatchStage& s, std::vector<gz_batch>* out);
void stage_clear(BatchStage* s);
void stage_release(gz_context* ctx, BatchStage* s);

// Validity bytes (1 = valid) to an Arrow bitmap of (n + 63) / 64 * 8 bytes (1 bits
// past row n), on the context's stream (gz_serve.hip).
int pack_validity(gz_context* ctx, const uint8_t* valid, int64_t n, uint8_t* bits);

// Hash aggregate grouped by several key columns or a utf8 key (gz_keys.hip): a device key
// dictionary maps every row's key tuple to a dense code, and a single-key
// aggregate groups by the code.
struct KeyDict;
// Per-call scratch of a dictionary's user (its codes: one int64 per row).
struct KeyScratch {
  DBuf rowslot, d_kb, d_prefix;
};
int kd_make(gz_context* ctx, int32_t K, const int32_t* key_types, KeyDict** out);
// Order of K key columns for the key dictionary, which takes a utf8 key last:
// perm[i] = the caller's key at position i (the one utf8 key, if any, moved to
// the end).  With equal null and empty masks, the UnsafeRow bytes of two key
// rows (unsafe_row.h:126-137: fixed-width values, then the utf8 bytes with no
// length) can only be equal when their single utf8 key has the same length,
// so their equality is field by field in any key order.
std::vector<int> kd_key_order(int K, const int32_t* key_types);
void kd_release(KeyDict* d);
void kd_scratch_free(gz_context* ctx, KeyScratch* sc);
// Codes of the batches' key tuples (key_cols of a ncols/types schema), on
// ctx's stream: mode 0 inserts every tuple (a null key is part of the tuple),
// 1 inserts tuples without nulls (a null key: code -1), 2 only looks tuples up
// (a miss or a null key: code -2).  *codes (device, one per row of the
// batches in order) stays valid until the scratch's next use.
int kd_encode(KeyDict* d, gz_context* ctx, KeyScratch* sc, int32_t ncols, const int32_t* types,
              const int32_t* key_cols, int32_t nb, const gz_batch* batches, int mode, const int64_t** codes,
              int64_t* rows_out);
// Normalise the float/double keys whose bit is set in mask (Spark's join-key
// "normalize") before they are hashed and compared; before the first insert.
int kd_set_normalize(KeyDict* d, uint32_t mask);
// The batches' buffer lists with the int64 code column appended; valid_col >= 0
// gives the code column that column's validity (a one-key dictionary: a null
// key is then a null code).
void with_code_column(int32_t ncols, const int32_t* types, int32_t nb, const gz_batch* batches, const int64_t* codes,
                      std::vector<int64_t>& addrs, std::vector<int64_t>& sizes, std::vector<gz_batch>& out,
                      int valid_col = -1);
struct MultiKey;
int mk_make(gz_context* ctx, int32_t ncols, const int32_t* types, int32_t nact, const gz_agg_action* acts,
            const gz_agg_options* opts, MultiKey** out);
int mk_process(MultiKey* m, int32_t nb, const gz_batch* batches);
int mk_num_groups(MultiKey* m, int64_t* out);
int mk_num_result_columns(MultiKey* m);
int mk_result_type(MultiKey* m, int col);
int mk_finish(MultiKey* m);
bool mk_finished(const MultiKey* m);
int mk_result_column(MultiKey* m, int col, void** values, void** valid);
int mk_result_binary(MultiKey* m, int col, void** offsets, void** data, void** valid);
void mk_release(MultiKey* m);

// Probe on another context than the relation's (resident relations).
void probe_set_context(gz_probe* h, gz_context* ctx);
// Wait for the relation's build (its context's stream).
int relation_synchronize(gz_relation* r);

}  // namespace gz
