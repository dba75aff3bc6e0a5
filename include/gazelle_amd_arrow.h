// gazelle_amd_arrow.h -- the reference's C++ operator interfaces over Arrow C++,
// backed by the C ABI of gazelle_amd.h.
//
// Class names, methods and argument meaning follow the reference so that a
// caller written against it switches by namespace:
//   Splitter                 shuffle/splitter.h:38-295 (Make / Split / Stop and the metrics)
//   ResultIterator<T>        codegen/common/result_iterator.h:37-82
//   KernalBase               codegen/arrow_compute/ext/kernels_ext.h:43-122
//   HashAggregateKernel      kernels_ext.h:285-313
//   HashRelationKernel       kernels_ext.h:535-558
//   ConditionedProbeKernel   kernels_ext.h:583-619
// The Gandiva node lists of the reference's factories are replaced by typed
// plans (action names with input column indices, key column indices), since
// the Gandiva fork is not available; an arrow::compute::ExecContext becomes a
// device Context.
//
// Record batches are ordinary (host) Arrow arrays, as the JNI layer hands them
// over.  Each call copies its batch into device memory, runs the HIP
// operators, and results come back as host Arrow arrays with validity bitmaps.
#ifndef GAZELLE_AMD_ARROW_H_
#define GAZELLE_AMD_ARROW_H_

#include <arrow/api.h>
#include <arrow/util/compression.h>

#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "gazelle_amd.h"

namespace gazelle_amd {

using ArrayList = std::vector<std::shared_ptr<arrow::Array>>;

// A non-OK gz status as an arrow::Status of the same code with the library's message.
arrow::Status FromGz(int code, const std::string& where);

// One GPU, its HIP stream and an HBM budget (0 = unlimited): the gz_context
// that the reference's arrow::compute::ExecContext stands for here.
class Context {
 public:
  static arrow::Result<std::shared_ptr<Context>> Make(int device = 0, int64_t hbm_budget = 0);
  ~Context();
  gz_context* handle() const { return ctx_; }
  int device() const { return device_; }
  arrow::Status Synchronize();

 private:
  Context() = default;
  gz_context* ctx_ = nullptr;
  int device_ = 0;
};

// A batch's buffers copied into device memory in the JNI buffer order
// (validity, [offsets], values per column), kept until the device work that
// reads them has run.
class DeviceBatch {
 public:
  static arrow::Result<std::unique_ptr<DeviceBatch>> Make(const ArrayList& columns, int64_t num_rows);
  ~DeviceBatch();
  gz_batch batch() const;
  int64_t num_rows() const { return num_rows_; }

 private:
  DeviceBatch() = default;
  int64_t num_rows_ = 0;
  void* dev_ = nullptr;
  std::vector<int64_t> addrs_, sizes_;
};

// ---- shuffle ------------------------------------------------------------------
// SplitOptions (shuffle/type.h:39-59).  num_sub_dirs, buffered_write,
// thread_id, task_attempt_id, memory_pool and ipc_write_options have no effect
// on the device splitter and are kept for the caller's sake.
struct SplitOptions {
  int64_t offheap_per_task = 0;
  int32_t buffer_size = 4096;
  int32_t num_sub_dirs = 64;
  int32_t batch_compress_threshold = 256;
  arrow::Compression::type compression_type = arrow::Compression::UNCOMPRESSED;
  bool prefer_spill = true;
  bool write_schema = true;
  bool buffered_write = false;
  std::string data_file;
  int64_t thread_id = -1;
  int64_t task_attempt_id = -1;
  static SplitOptions Defaults() { return SplitOptions(); }
};

class Splitter {
 public:
  // Splitter::Make (splitter.h:55-62): short_name "hash", "rr", "range" or
  // "single"; key_columns are the hash keys (the field references of the
  // reference's expression vector).
  static arrow::Result<std::shared_ptr<Splitter>> Make(const std::string& short_name,
                                                       std::shared_ptr<arrow::Schema> schema, int num_partitions,
                                                       const std::vector<int>& key_columns,
                                                       SplitOptions options = SplitOptions::Defaults(),
                                                       std::shared_ptr<Context> ctx = nullptr);
  static arrow::Result<std::shared_ptr<Splitter>> Make(const std::string& short_name,
                                                       std::shared_ptr<arrow::Schema> schema, int num_partitions,
                                                       SplitOptions options = SplitOptions::Defaults(),
                                                       std::shared_ptr<Context> ctx = nullptr);
  ~Splitter();

  const std::shared_ptr<arrow::Schema>& input_schema() const { return schema_; }
  arrow::Status Split(const arrow::RecordBatch& rb);
  // Body length of the batch's IPC payload with the current codec, or -1
  // (splitter.cc:358-369): ColumnarShuffleWriter's first-batch codec probe.
  // Computed on the host by Arrow's own IPC writer, as the reference does.
  int64_t CompressedSize(const arrow::RecordBatch& rb);
  // Writes the data file (options.data_file, or a temporary file).
  arrow::Status Stop();
  arrow::Status SetCompressType(arrow::Compression::type compressed_type);
  // The device splitter keeps its partitions in HBM: nothing to spill.
  arrow::Status SpillFixedSize(int64_t size, int64_t* actual);

  int64_t TotalBytesWritten() const { return metrics_.total_bytes_written; }
  int64_t TotalBytesSpilled() const { return metrics_.total_bytes_spilled; }
  int64_t TotalWriteTime() const { return metrics_.total_write_time; }
  int64_t TotalSpillTime() const { return metrics_.total_spill_time; }
  int64_t TotalCompressTime() const { return metrics_.total_compress_time; }
  int64_t TotalComputePidTime() const { return metrics_.total_compute_pid_time; }
  const std::vector<int64_t>& PartitionLengths() const { return partition_lengths_; }
  const std::vector<int64_t>& RawPartitionLengths() const { return raw_partition_lengths_; }
  int64_t RawPartitionBytes() const;
  const std::string& DataFile() const { return options_.data_file; }

 private:
  Splitter() = default;
  std::shared_ptr<Context> ctx_;
  gz_splitter* h_ = nullptr;
  std::shared_ptr<arrow::Schema> schema_;
  int num_partitions_ = 0;
  SplitOptions options_;
  std::vector<std::unique_ptr<DeviceBatch>> pending_;
  std::vector<int64_t> partition_lengths_, raw_partition_lengths_;
  gz_split_metrics metrics_{};
  bool stopped_ = false;
};

// ---- result iterators and kernels ---------------------------------------------
class ResultIteratorBase {
 public:
  virtual ~ResultIteratorBase() = default;
  virtual bool HasNext() { return false; }
  virtual std::string ToString() { return ""; }
  virtual arrow::Status ProcessAndCacheOne(const ArrayList& in,
                                           const std::shared_ptr<arrow::Array>& selection = nullptr) {
    return arrow::Status::NotImplemented("ResultIterator abstract ProcessAndCacheOne()");
  }
};

template <typename T>
class ResultIterator : public ResultIteratorBase {
 public:
  bool HasNext() override { return false; }
  virtual arrow::Status Next(std::shared_ptr<T>* out) {
    return arrow::Status::NotImplemented("ResultIterator abstract Next()");
  }
  virtual arrow::Status Process(const ArrayList& in, std::shared_ptr<T>* out,
                                const std::shared_ptr<arrow::Array>& selection = nullptr) {
    return arrow::Status::NotImplemented("ResultIterator abstract Process()");
  }
  virtual arrow::Status SetDependencies(const std::vector<std::shared_ptr<ResultIteratorBase>>&) {
    return arrow::Status::NotImplemented("ResultIterator abstract SetDependencies()");
  }
};

// The built hash relation a probe reads: resident in HBM, shared by the probe
// iterators that depend on it (codegen/common/hash_relation.h:123-645).
class HashRelation {
 public:
  explicit HashRelation(gz_relation* r, std::shared_ptr<Context> ctx) : r_(r), ctx_(std::move(ctx)) {}
  ~HashRelation();
  gz_relation* handle() const { return r_; }
  int64_t NumKeys() const;
  bool HasNullKey() const;

 private:
  gz_relation* r_;
  std::shared_ptr<Context> ctx_;
};

class KernalBase {
 public:
  virtual ~KernalBase() = default;
  virtual arrow::Status Evaluate(ArrayList& in) { return arrow::Status::NotImplemented("Evaluate"); }
  virtual arrow::Status MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                           std::shared_ptr<ResultIterator<arrow::RecordBatch>>* out) {
    return arrow::Status::NotImplemented("MakeResultIterator");
  }
  virtual arrow::Status MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                           std::shared_ptr<ResultIterator<HashRelation>>* out) {
    return arrow::Status::NotImplemented("MakeResultIterator");
  }
  std::string kernel_name_;
};

// One aggregate action as the reference names it (action_groupby,
// action_sum, action_avg, action_count, action_countLiteral_1, action_min,
// action_min_true, action_max, action_max_true, action_sum_count,
// action_avgByCount) and the input columns it reads.
struct Action {
  std::string name;
  std::vector<int> inputs;
};

class HashAggregateKernel : public KernalBase {
 public:
  // HashAggregateKernel::Make (kernels_ext.h:285-313): input_field_list are the
  // input columns, action_list the actions; result_field_list names and types
  // the output columns (one per result; sum_count gives two).
  static arrow::Status Make(std::shared_ptr<Context> ctx, const arrow::FieldVector& input_field_list,
                            const std::vector<Action>& action_list, const arrow::FieldVector& result_field_list,
                            std::shared_ptr<KernalBase>* out);
  // The iterator: ProcessAndCacheOne(columns) per batch; HasNext / Next give
  // the groups in first-seen order, NATIVESQL_BATCH_SIZE (10000) at a time.
  arrow::Status MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                   std::shared_ptr<ResultIterator<arrow::RecordBatch>>* out) override;

 private:
  std::shared_ptr<Context> ctx_;
  arrow::FieldVector inputs_, results_;
  std::vector<gz_agg_action> actions_;
};

class HashRelationKernel : public KernalBase, public std::enable_shared_from_this<HashRelationKernel> {
 public:
  // HashRelationKernel::Make (kernels_ext.h:535-558): the build side's columns,
  // its key columns (one, or several: the UnsafeRow-keyed relation) and the
  // builder type (1 = unsafeHashMap, 3 = skipping duplicate keys for semi joins).
  // normalize_keys[j] marks a float/double key wrapped in Spark's "normalize"
  // (gz_relation_normalize_keys); it applies to the probe keys as well.
  static arrow::Status Make(std::shared_ptr<Context> ctx, const arrow::FieldVector& input_field_list,
                            const std::vector<int>& key_indices, int builder_type,
                            std::shared_ptr<KernalBase>* out, const std::vector<bool>& normalize_keys = {});
  ~HashRelationKernel() override;
  // Evaluate (hash_relation_kernel.cc:170-198): one build batch's columns.
  arrow::Status Evaluate(ArrayList& in) override;
  // The iterator's Next gives the finished relation (FinishInternal, :200-281).
  arrow::Status MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                   std::shared_ptr<ResultIterator<HashRelation>>* out) override;

 private:
  std::shared_ptr<Context> ctx_;
  arrow::FieldVector inputs_;
  gz_relation* r_ = nullptr;
  std::vector<std::unique_ptr<DeviceBatch>> pending_;
  std::shared_ptr<HashRelation> relation_;
};

class ConditionedProbeKernel : public KernalBase {
 public:
  // ConditionedProbeKernel::Make (kernels_ext.h:583-619): right_key_list are the
  // probe key columns, right_schema_list the probe columns, join_type 0 inner,
  // 2 anti, 3 semi; result_schema lists (side, column) per output column (side
  // 0 build, 1 probe).  A condition is not built (the reference's route refuses
  // one too, conditioned_probe_kernel.cc:118-121): pass nullptr.
  static arrow::Status Make(std::shared_ptr<Context> ctx, const std::vector<int>& left_key_list,
                            const std::vector<int>& right_key_list, const arrow::FieldVector& left_schema_list,
                            const arrow::FieldVector& right_schema_list, const void* condition, int join_type,
                            bool is_null_aware_anti_join, const std::vector<std::pair<int, int>>& result_schema,
                            int hash_relation_idx, std::shared_ptr<KernalBase>* out);
  // The iterator: SetDependencies({relation iterator}), then Process(probe
  // columns, &out) gives one output RecordBatch per probe batch.
  arrow::Status MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                   std::shared_ptr<ResultIterator<arrow::RecordBatch>>* out) override;

 private:
  std::shared_ptr<Context> ctx_;
  std::vector<int> left_keys_, right_keys_;
  arrow::FieldVector left_, right_;
  int join_type_ = 0;
  bool null_aware_ = false;
  std::vector<std::pair<int, int>> result_;
  int relation_idx_ = 0;
};

}  // namespace gazelle_amd

#endif  // GAZELLE_AMD_ARROW_H_
