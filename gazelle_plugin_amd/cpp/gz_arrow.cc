// The reference's C++ operator interfaces over Arrow C++ (gazelle_amd_arrow.h):
// host Arrow arrays in, the C ABI's device operators, host Arrow arrays out.
#include "gazelle_amd_arrow.h"

#include <arrow/io/file.h>
#include <arrow/ipc/api.h>
#include <arrow/util/bitmap_ops.h>
#include <hip/hip_runtime_api.h>
#include <stdlib.h>
#include <unistd.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>

namespace gazelle_amd {

namespace {

constexpr int64_t kAlign = 256;
// Input batches kept on the device for stream-ordered reads; above this many
// bytes the owner waits for its stream and frees them.
constexpr int64_t kPendingLimit = int64_t(1) << 30;

int64_t AlignUp(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

arrow::Status Hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return arrow::Status::OK();
  return arrow::Status::UnknownError(what, ": ", hipGetErrorString(e));
}

#define GZ_RETURN(call, where)                              \
  do {                                                      \
    int _rc = (call);                                       \
    if (_rc != GZ_OK) return ::gazelle_amd::FromGz(_rc, where); \
  } while (0)

// Bytes per value of a fixed-width column type; 0 for bool (bits), utf8/binary and null.
int Width(int t) {
  switch (t) {
    case GZ_TYPE_UINT8: case GZ_TYPE_INT8: return 1;
    case GZ_TYPE_UINT16: case 5 /* INT16 */: return 2;
    case GZ_TYPE_UINT32: case GZ_TYPE_INT32: case GZ_TYPE_FLOAT: case GZ_TYPE_DATE32: return 4;
    case GZ_TYPE_UINT64: case GZ_TYPE_INT64: case GZ_TYPE_DOUBLE: case GZ_TYPE_DATE64: case GZ_TYPE_TIMESTAMP:
      return 8;
    case GZ_TYPE_DECIMAL128: return 16;
    default: return 0;
  }
}

bool IsBinary(int t) { return t == GZ_TYPE_STRING || t == GZ_TYPE_BINARY; }

// The C ABI's type id of an Arrow type (the ids are arrow::Type::type values).
arrow::Result<int32_t> GzType(const std::shared_ptr<arrow::DataType>& type) {
  int t = type->id();
  if (t == GZ_TYPE_NA || t == GZ_TYPE_BOOL || IsBinary(t) || Width(t) > 0) return int32_t(t);
  return arrow::Status::NotImplemented("column type ", type->ToString(), " is not supported");
}

arrow::Result<std::vector<int32_t>> GzTypes(const arrow::FieldVector& fields) {
  std::vector<int32_t> out;
  for (const auto& f : fields) {
    ARROW_ASSIGN_OR_RAISE(auto t, GzType(f->type()));
    out.push_back(t);
  }
  return out;
}

// One host buffer of a batch, either borrowed from the array or rebuilt with
// row 0 at bit / element 0.
struct HostBuf {
  const uint8_t* p = nullptr;
  int64_t n = 0;
  std::vector<uint8_t> own;
};

void Bits(const uint8_t* src, int64_t offset, int64_t rows, HostBuf* out) {
  out->n = (rows + 7) / 8;
  if (offset % 8 == 0) {
    out->p = src + offset / 8;
    return;
  }
  out->own.assign(out->n, 0);
  for (int64_t i = 0; i < rows; i++)
    if (arrow::bit_util::GetBit(src, offset + i)) arrow::bit_util::SetBit(out->own.data(), i);
  out->p = out->own.data();
}

arrow::Status ColumnBuffers(const arrow::Array& arr, std::vector<HostBuf>* bufs) {
  const arrow::ArrayData& d = *arr.data();
  const int64_t n = d.length, off = d.offset;
  ARROW_ASSIGN_OR_RAISE(int t, GzType(d.type));
  HostBuf valid;
  if (t != GZ_TYPE_NA && arr.null_count() > 0 && d.buffers[0]) Bits(d.buffers[0]->data(), off, n, &valid);
  bufs->push_back(std::move(valid));
  if (t == GZ_TYPE_NA) {
    bufs->emplace_back();
    return arrow::Status::OK();
  }
  if (IsBinary(t)) {
    HostBuf offs, data;
    const int32_t* o = d.GetValues<int32_t>(1);  // includes the array offset
    offs.n = (n + 1) * 4;
    if (o[0] == 0) {
      offs.p = reinterpret_cast<const uint8_t*>(o);
    } else {
      offs.own.resize(offs.n);
      int32_t* w = reinterpret_cast<int32_t*>(offs.own.data());
      for (int64_t i = 0; i <= n; i++) w[i] = o[i] - o[0];
      offs.p = offs.own.data();
    }
    data.n = o[n] - o[0];
    data.p = d.buffers[2] ? d.buffers[2]->data() + o[0] : nullptr;
    bufs->push_back(std::move(offs));
    bufs->push_back(std::move(data));
    return arrow::Status::OK();
  }
  HostBuf vals;
  if (t == GZ_TYPE_BOOL) {
    Bits(d.buffers[1]->data(), off, n, &vals);
  } else {
    int w = Width(t);
    vals.p = d.buffers[1]->data() + off * w;
    vals.n = n * w;
  }
  bufs->push_back(std::move(vals));
  return arrow::Status::OK();
}

arrow::Result<std::shared_ptr<arrow::Buffer>> FromDevice(const void* src, int64_t n) {
  ARROW_ASSIGN_OR_RAISE(std::shared_ptr<arrow::Buffer> b, arrow::AllocateBuffer(n));
  if (n > 0) ARROW_RETURN_NOT_OK(Hip(hipMemcpy(b->mutable_data(), src, n, hipMemcpyDeviceToHost), "hipMemcpy"));
  return b;
}

// Validity bitmap (LSB first) of n rows from a device bitmap; null when every row is valid.
arrow::Result<std::shared_ptr<arrow::Buffer>> Bitmap(const void* dev_bits, int64_t n, int64_t* null_count) {
  *null_count = 0;
  if (n == 0 || dev_bits == nullptr) return nullptr;
  ARROW_ASSIGN_OR_RAISE(auto b, FromDevice(dev_bits, (n + 7) / 8));
  *null_count = n - arrow::internal::CountSetBits(b->data(), 0, n);
  if (*null_count == 0) return nullptr;
  return b;
}

// A column of n rows from device buffers: fixed values (bool: one byte per row),
// or utf8/binary offsets + data, with a device validity bitmap.
arrow::Result<std::shared_ptr<arrow::Array>> MakeColumn(const std::shared_ptr<arrow::DataType>& type, int64_t n,
                                                        const void* values, const void* offsets, const void* data,
                                                        const void* dev_bits) {
  int t = type->id();
  int64_t nulls = 0;
  ARROW_ASSIGN_OR_RAISE(auto bitmap, Bitmap(dev_bits, n, &nulls));
  if (t == GZ_TYPE_NA) return arrow::MakeArrayOfNull(arrow::null(), n);
  std::vector<std::shared_ptr<arrow::Buffer>> bufs{bitmap};
  if (IsBinary(t)) {
    ARROW_ASSIGN_OR_RAISE(auto offs, FromDevice(offsets, (n + 1) * 4));
    int64_t nbytes = n ? reinterpret_cast<const int32_t*>(offs->data())[n] : 0;
    ARROW_ASSIGN_OR_RAISE(auto bytes, FromDevice(data, nbytes));
    bufs.push_back(offs);
    bufs.push_back(bytes);
  } else if (t == GZ_TYPE_BOOL) {
    ARROW_ASSIGN_OR_RAISE(auto bytes, FromDevice(values, n));
    ARROW_ASSIGN_OR_RAISE(std::shared_ptr<arrow::Buffer> bits, arrow::AllocateBuffer((n + 7) / 8));
    std::memset(bits->mutable_data(), 0, bits->size());
    for (int64_t i = 0; i < n; i++)
      if (bytes->data()[i]) arrow::bit_util::SetBit(bits->mutable_data(), i);
    bufs.push_back(bits);
  } else {
    ARROW_ASSIGN_OR_RAISE(auto vals, FromDevice(values, n * Width(t)));
    bufs.push_back(vals);
  }
  return arrow::MakeArray(arrow::ArrayData::Make(type, n, std::move(bufs), nulls));
}

int64_t NumRows(const ArrayList& in) { return in.empty() ? 0 : in[0]->length(); }

arrow::Status CheckColumns(const ArrayList& in, const arrow::FieldVector& fields, const char* who) {
  if (in.size() != fields.size())
    return arrow::Status::Invalid(who, ": ", in.size(), " columns, the schema has ", fields.size());
  for (size_t c = 0; c < in.size(); c++) {
    if (!in[c]->type()->Equals(*fields[c]->type()))
      return arrow::Status::TypeError(who, ": column ", c, " is ", in[c]->type()->ToString(), ", expected ",
                                      fields[c]->type()->ToString());
    if (in[c]->length() != in[0]->length()) return arrow::Status::Invalid(who, ": ragged columns");
  }
  return arrow::Status::OK();
}

// Input batches borrowed by queued device work, freed once the stream has passed them.
class Pending {
 public:
  void Add(std::unique_ptr<DeviceBatch> b, int64_t bytes) {
    bytes_ += bytes;
    batches_.push_back(std::move(b));
  }
  bool Full() const { return bytes_ > kPendingLimit; }
  arrow::Status Release(Context* ctx) {
    if (batches_.empty()) return arrow::Status::OK();
    ARROW_RETURN_NOT_OK(ctx->Synchronize());
    batches_.clear();
    bytes_ = 0;
    return arrow::Status::OK();
  }

 private:
  std::vector<std::unique_ptr<DeviceBatch>> batches_;
  int64_t bytes_ = 0;
};

int64_t BatchBytes(const ArrayList& in) {
  int64_t s = 0;
  for (const auto& a : in)
    for (const auto& b : a->data()->buffers)
      if (b) s += b->size();
  return s;
}

}  // namespace

arrow::Status FromGz(int code, const std::string& where) {
  const char* msg = gz_last_error();
  std::string text = where + ": " + (msg ? msg : "");
  if (code == GZ_HIP_ERROR) return arrow::Status::UnknownError(text);  // a HIP runtime failure
  return arrow::Status(static_cast<arrow::StatusCode>(code), text);
}

// ---- Context ---------------------------------------------------------------------
arrow::Result<std::shared_ptr<Context>> Context::Make(int device, int64_t hbm_budget) {
  std::shared_ptr<Context> c(new Context());
  c->device_ = device;
  GZ_RETURN(gz_context_create(device, hbm_budget, &c->ctx_), "gz_context_create");
  return c;
}

Context::~Context() {
  if (ctx_) gz_context_destroy(ctx_);
}

arrow::Status Context::Synchronize() {
  GZ_RETURN(gz_context_synchronize(ctx_), "gz_context_synchronize");
  return arrow::Status::OK();
}

// ---- DeviceBatch -------------------------------------------------------------------
arrow::Result<std::unique_ptr<DeviceBatch>> DeviceBatch::Make(const ArrayList& columns, int64_t num_rows) {
  std::vector<HostBuf> bufs;
  for (const auto& a : columns) {
    if (a->length() != num_rows) return arrow::Status::Invalid("column of ", a->length(), " rows in a batch of ", num_rows);
    ARROW_RETURN_NOT_OK(ColumnBuffers(*a, &bufs));
  }
  std::unique_ptr<DeviceBatch> out(new DeviceBatch());
  out->num_rows_ = num_rows;
  int64_t total = 0;
  for (const auto& b : bufs) total += AlignUp(b.n);
  if (total > 0) ARROW_RETURN_NOT_OK(Hip(hipMalloc(&out->dev_, total), "hipMalloc"));
  int64_t pos = 0;
  for (const auto& b : bufs) {
    uint8_t* dst = static_cast<uint8_t*>(out->dev_) + pos;
    if (b.n > 0) ARROW_RETURN_NOT_OK(Hip(hipMemcpy(dst, b.p, b.n, hipMemcpyHostToDevice), "hipMemcpy"));
    out->addrs_.push_back(b.n > 0 ? reinterpret_cast<int64_t>(dst) : 0);
    out->sizes_.push_back(b.n);
    pos += AlignUp(b.n);
  }
  return out;
}

DeviceBatch::~DeviceBatch() {
  if (dev_) (void)hipFree(dev_);
}

gz_batch DeviceBatch::batch() const {
  gz_batch b;
  b.num_rows = num_rows_;
  b.buf_addrs = addrs_.data();
  b.buf_sizes = sizes_.data();
  b.num_bufs = static_cast<int32_t>(addrs_.size());
  return b;
}

// ---- Splitter ----------------------------------------------------------------------
arrow::Result<std::shared_ptr<Splitter>> Splitter::Make(const std::string& short_name,
                                                        std::shared_ptr<arrow::Schema> schema, int num_partitions,
                                                        SplitOptions options, std::shared_ptr<Context> ctx) {
  return Make(short_name, std::move(schema), num_partitions, {}, std::move(options), std::move(ctx));
}

arrow::Result<std::shared_ptr<Splitter>> Splitter::Make(const std::string& short_name,
                                                        std::shared_ptr<arrow::Schema> schema, int num_partitions,
                                                        const std::vector<int>& key_columns, SplitOptions options,
                                                        std::shared_ptr<Context> ctx) {
  if (!schema) return arrow::Status::Invalid("Splitter::Make: null schema");
  if (!ctx) {
    ARROW_ASSIGN_OR_RAISE(ctx, Context::Make(0));
  }
  ARROW_ASSIGN_OR_RAISE(auto types, GzTypes(schema->fields()));
  std::shared_ptr<Splitter> s(new Splitter());
  s->ctx_ = ctx;
  s->schema_ = schema;
  s->num_partitions_ = num_partitions;
  s->options_ = options;
  gz_split_options o;
  gz_split_options_defaults(&o);
  o.buffer_size = options.buffer_size;
  o.prefer_spill = options.prefer_spill ? 1 : 0;
  o.offheap_per_task = options.offheap_per_task;
  o.batch_compress_threshold = options.batch_compress_threshold;
  std::vector<int32_t> keys(key_columns.begin(), key_columns.end());
  GZ_RETURN(gz_split_make(ctx->handle(), short_name.c_str(), static_cast<int32_t>(types.size()), types.data(),
                          num_partitions, keys.data(), static_cast<int32_t>(keys.size()), &o, &s->h_),
            "Splitter::Make");
  ARROW_RETURN_NOT_OK(s->SetCompressType(options.compression_type));
  return s;
}

Splitter::~Splitter() {
  if (ctx_) (void)ctx_->Synchronize();
  pending_.clear();
  if (h_) gz_split_release(h_);
}

arrow::Status Splitter::Split(const arrow::RecordBatch& rb) {
  if (stopped_) return arrow::Status::Invalid("Splitter::Split after Stop");
  ARROW_RETURN_NOT_OK(CheckColumns(rb.columns(), schema_->fields(), "Splitter::Split"));
  ARROW_ASSIGN_OR_RAISE(auto db, DeviceBatch::Make(rb.columns(), rb.num_rows()));
  gz_batch b = db->batch();
  GZ_RETURN(gz_split_batch(h_, b.num_rows, b.buf_addrs, b.buf_sizes, b.num_bufs), "Splitter::Split");
  pending_.push_back(std::move(db));
  if (pending_.size() >= 64) {  // the splitter holds the split rows; the inputs go once queued work has read them
    ARROW_RETURN_NOT_OK(ctx_->Synchronize());
    pending_.clear();
  }
  return arrow::Status::OK();
}

int64_t Splitter::CompressedSize(const arrow::RecordBatch& rb) {
  arrow::ipc::IpcWriteOptions opts = arrow::ipc::IpcWriteOptions::Defaults();
  if (options_.compression_type != arrow::Compression::UNCOMPRESSED) {
    auto codec = arrow::util::Codec::Create(options_.compression_type);
    if (!codec.ok()) return -1;
    opts.codec = std::shared_ptr<arrow::util::Codec>(std::move(*codec));
  }
  arrow::ipc::IpcPayload payload;
  if (!arrow::ipc::GetRecordBatchPayload(rb, opts, &payload).ok()) return -1;
  return payload.body_length;
}

arrow::Status Splitter::SetCompressType(arrow::Compression::type compressed_type) {
  int32_t codec;
  if (compressed_type == arrow::Compression::UNCOMPRESSED) {
    codec = GZ_COMPRESSION_UNCOMPRESSED;
  } else if (compressed_type == arrow::Compression::LZ4_FRAME) {
    codec = GZ_COMPRESSION_LZ4_FRAME;
  } else {
    return arrow::Status::NotImplemented("compression ", arrow::util::Codec::GetCodecAsString(compressed_type),
                                         " is not implemented on the device");
  }
  GZ_RETURN(gz_split_set_compress_type(h_, codec), "Splitter::SetCompressType");
  options_.compression_type = compressed_type;
  return arrow::Status::OK();
}

arrow::Status Splitter::SpillFixedSize(int64_t size, int64_t* actual) {
  *actual = 0;
  return arrow::Status::OK();
}

int64_t Splitter::RawPartitionBytes() const {
  return std::accumulate(raw_partition_lengths_.begin(), raw_partition_lengths_.end(), int64_t(0));
}

arrow::Status Splitter::Stop() {
  if (stopped_) return arrow::Status::Invalid("Splitter::Stop called twice");
  stopped_ = true;
  GZ_RETURN(gz_split_stop(h_), "Splitter::Stop");
  pending_.clear();  // gz_split_stop completed the device work that read them
  // the output schema: the "range" splitter's leading pid column is not written
  arrow::FieldVector out_fields = schema_->fields();
  int ncol_out = gz_split_num_output_columns(h_);
  if (ncol_out < static_cast<int>(out_fields.size())) out_fields.erase(out_fields.begin());
  auto out_schema = arrow::schema(out_fields, schema_->metadata());
  ARROW_ASSIGN_OR_RAISE(auto schema_msg, arrow::ipc::SerializeSchema(*out_schema));
  partition_lengths_.assign(num_partitions_, 0);
  int64_t total = 0;
  GZ_RETURN(gz_split_ipc_prepare(h_, schema_msg->data(), schema_msg->size(), options_.write_schema ? 1 : 0, &total,
                                 partition_lengths_.data()),
            "Splitter::Stop");
  void* dev = nullptr;
  if (total > 0) ARROW_RETURN_NOT_OK(Hip(hipMalloc(&dev, total), "hipMalloc"));
  int rc = gz_split_ipc_write(h_, dev, total);
  if (rc == GZ_OK) rc = gz_context_synchronize(ctx_->handle());
  arrow::Result<std::shared_ptr<arrow::Buffer>> host =
      rc == GZ_OK ? FromDevice(dev, total) : arrow::Result<std::shared_ptr<arrow::Buffer>>(FromGz(rc, "Splitter::Stop"));
  if (dev) (void)hipFree(dev);
  ARROW_RETURN_NOT_OK(host.status());
  raw_partition_lengths_.assign(num_partitions_, 0);
  GZ_RETURN(gz_split_raw_partition_lengths(h_, raw_partition_lengths_.data()), "Splitter::Stop");
  GZ_RETURN(gz_split_get_metrics(h_, &metrics_), "Splitter::Stop");

  auto t0 = std::chrono::steady_clock::now();
  if (options_.data_file.empty()) {
    const char* dir = getenv("TMPDIR");
    std::string tmpl = std::string(dir && *dir ? dir : "/tmp") + "/gazelle_amd_shuffle_XXXXXX";
    std::vector<char> path(tmpl.begin(), tmpl.end());
    path.push_back('\0');
    int fd = mkstemp(path.data());
    if (fd < 0) return arrow::Status::IOError("cannot create a temporary data file in ", tmpl);
    close(fd);
    options_.data_file = path.data();
  }
  ARROW_ASSIGN_OR_RAISE(auto file, arrow::io::FileOutputStream::Open(options_.data_file));
  ARROW_RETURN_NOT_OK(file->Write((*host)->data(), total));
  ARROW_RETURN_NOT_OK(file->Close());
  metrics_.total_write_time +=
      std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
  return arrow::Status::OK();
}

// ---- aggregate -------------------------------------------------------------------
namespace {

arrow::Result<gz_agg_action> ParseAction(const Action& a) {
  gz_agg_action g{-1, a.inputs.size() > 0 ? a.inputs[0] : -1, a.inputs.size() > 1 ? a.inputs[1] : -1, 0};
  const std::string& n = a.name;
  auto is = [&](const char* prefix) { return n == prefix || n.rfind(std::string(prefix) + "_", 0) == 0; };
  if (n == "action_groupby") {
    g.kind = GZ_ACTION_GROUPBY;
    g.input1 = -1;
  } else if (n == "action_count") {
    g.kind = GZ_ACTION_COUNT;
  } else if (n.rfind("action_countLiteral_", 0) == 0) {
    g.kind = GZ_ACTION_COUNT_LITERAL;
    g.input0 = g.input1 = -1;
  } else if (n == "action_sum") {
    g.kind = GZ_ACTION_SUM;
  } else if (n == "action_avg") {
    g.kind = GZ_ACTION_AVG;
  } else if (n == "action_sum_count") {
    g.kind = GZ_ACTION_SUM_COUNT;
  } else if (n == "action_avgByCount") {
    g.kind = GZ_ACTION_AVG_BY_COUNT;
  } else if (is("action_min") || is("action_max")) {
    bool is_min = n.rfind("action_min", 0) == 0;
    g.kind = is_min ? GZ_ACTION_MIN : GZ_ACTION_MAX;
    g.nan_check = n.size() > 10 && n.substr(11) == "true";
    g.input1 = -1;
  } else {
    return arrow::Status::NotImplemented(n, " is not implemented.");
  }
  if (g.kind != GZ_ACTION_AVG_BY_COUNT && g.kind != GZ_ACTION_COUNT) g.input1 = -1;
  return g;
}

class AggregateIterator : public ResultIterator<arrow::RecordBatch> {
 public:
  AggregateIterator(std::shared_ptr<Context> ctx, gz_agg* h, arrow::FieldVector inputs,
                    std::shared_ptr<arrow::Schema> schema)
      : ctx_(std::move(ctx)), h_(h), inputs_(std::move(inputs)), schema_(std::move(schema)) {}
  ~AggregateIterator() override {
    (void)pending_.Release(ctx_.get());
    gz_agg_release(h_);
  }

  arrow::Status ProcessAndCacheOne(const ArrayList& in, const std::shared_ptr<arrow::Array>& selection) override {
    if (selection) return arrow::Status::NotImplemented("HashAggregate: selection vectors");
    if (finished_) return arrow::Status::Invalid("HashAggregate: ProcessAndCacheOne after Next");
    ARROW_RETURN_NOT_OK(CheckColumns(in, inputs_, "HashAggregate::ProcessAndCacheOne"));
    ARROW_ASSIGN_OR_RAISE(auto db, DeviceBatch::Make(in, NumRows(in)));
    gz_batch b = db->batch();
    GZ_RETURN(gz_agg_process(h_, b.num_rows, b.buf_addrs, b.buf_sizes, b.num_bufs),
              "HashAggregate::ProcessAndCacheOne");
    pending_.Add(std::move(db), BatchBytes(in));
    if (pending_.Full()) ARROW_RETURN_NOT_OK(pending_.Release(ctx_.get()));
    return arrow::Status::OK();
  }

  // A failure while finishing is reported by the Next that follows.
  bool HasNext() override {
    if (!finished_) status_ = Finish();
    return !status_.ok() || offset_ < result_->num_rows();
  }

  arrow::Status Next(std::shared_ptr<arrow::RecordBatch>* out) override {
    if (!finished_) status_ = Finish();
    ARROW_RETURN_NOT_OK(status_);
    if (offset_ >= result_->num_rows()) return arrow::Status::Invalid("HashAggregate: Next past the last batch");
    int64_t bs = 10000;
    if (const char* e = getenv("NATIVESQL_BATCH_SIZE")) bs = std::max<int64_t>(1, atoll(e));
    *out = result_->Slice(offset_, bs);
    offset_ += (*out)->num_rows();
    return arrow::Status::OK();
  }

 private:
  arrow::Status Finish() {
    finished_ = true;
    GZ_RETURN(gz_agg_finish(h_), "HashAggregate::Next");
    int64_t g = 0;
    GZ_RETURN(gz_agg_num_groups(h_, &g), "HashAggregate::Next");
    int ncol = gz_agg_num_result_columns(h_);
    if (ncol != schema_->num_fields())
      return arrow::Status::Invalid("HashAggregate: ", ncol, " result columns, the result fields name ",
                                    schema_->num_fields());
    std::vector<void*> bits(ncol);
    for (int c = 0; c < ncol; c++) {
      int t = gz_agg_result_type(h_, c);
      if (t != schema_->field(c)->type()->id())
        return arrow::Status::TypeError("HashAggregate: result ", c, " is type id ", t, ", the result field is ",
                                        schema_->field(c)->type()->ToString());
      GZ_RETURN(gz_agg_result_bitmap(h_, c, &bits[c]), "HashAggregate::Next");
    }
    ARROW_RETURN_NOT_OK(ctx_->Synchronize());
    ARROW_RETURN_NOT_OK(pending_.Release(ctx_.get()));
    ArrayList cols;
    for (int c = 0; c < ncol; c++) {
      auto type = schema_->field(c)->type();
      void *vals = nullptr, *valid = nullptr, *offs = nullptr, *data = nullptr;
      if (IsBinary(type->id())) {
        GZ_RETURN(gz_agg_result_binary(h_, c, &offs, &data, &valid), "HashAggregate::Next");
      } else {
        GZ_RETURN(gz_agg_result_column(h_, c, &vals, &valid), "HashAggregate::Next");
      }
      ARROW_ASSIGN_OR_RAISE(auto col, MakeColumn(type, g, vals, offs, data, bits[c]));
      cols.push_back(col);
    }
    result_ = arrow::RecordBatch::Make(schema_, g, cols);
    return arrow::Status::OK();
  }

  std::shared_ptr<Context> ctx_;
  gz_agg* h_;
  arrow::FieldVector inputs_;
  std::shared_ptr<arrow::Schema> schema_;
  Pending pending_;
  bool finished_ = false;
  arrow::Status status_;
  std::shared_ptr<arrow::RecordBatch> result_;
  int64_t offset_ = 0;
};

}  // namespace

arrow::Status HashAggregateKernel::Make(std::shared_ptr<Context> ctx, const arrow::FieldVector& input_field_list,
                                        const std::vector<Action>& action_list,
                                        const arrow::FieldVector& result_field_list,
                                        std::shared_ptr<KernalBase>* out) {
  if (!ctx) {
    ARROW_ASSIGN_OR_RAISE(ctx, Context::Make(0));
  }
  ARROW_RETURN_NOT_OK(GzTypes(input_field_list).status());
  auto k = std::make_shared<HashAggregateKernel>();
  k->ctx_ = ctx;
  k->inputs_ = input_field_list;
  k->results_ = result_field_list;
  for (const auto& a : action_list) {
    ARROW_ASSIGN_OR_RAISE(auto g, ParseAction(a));
    for (int i : {g.input0, g.input1})
      if (i >= static_cast<int>(input_field_list.size()))
        return arrow::Status::IndexError(a.name, ": input column ", i, " of ", input_field_list.size());
    k->actions_.push_back(g);
  }
  k->kernel_name_ = "HashAggregateKernel";
  *out = k;
  return arrow::Status::OK();
}

arrow::Status HashAggregateKernel::MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                                      std::shared_ptr<ResultIterator<arrow::RecordBatch>>* out) {
  ARROW_ASSIGN_OR_RAISE(auto types, GzTypes(inputs_));
  gz_agg_options o;
  gz_agg_options_defaults(&o);
  gz_agg* h = nullptr;
  GZ_RETURN(gz_agg_make(ctx_->handle(), static_cast<int32_t>(types.size()), types.data(),
                        static_cast<int32_t>(actions_.size()), actions_.data(), &o, &h),
            "HashAggregateKernel::MakeResultIterator");
  *out = std::make_shared<AggregateIterator>(ctx_, h, inputs_, arrow::schema(results_));
  return arrow::Status::OK();
}

// ---- hash relation ---------------------------------------------------------------
HashRelation::~HashRelation() {
  if (r_) gz_relation_release(r_);
}

int64_t HashRelation::NumKeys() const {
  int64_t n = -1;
  gz_relation_num_keys(r_, &n);
  return n;
}

bool HashRelation::HasNullKey() const {
  int32_t v = 0;
  gz_relation_has_null_key(r_, &v);
  return v != 0;
}

namespace {

class RelationIterator : public ResultIterator<HashRelation> {
 public:
  bool HasNext() override { return !delivered_; }
  // Every call gives the same finished relation (several probes may depend on it).
  arrow::Status Next(std::shared_ptr<HashRelation>* out) override {
    ARROW_RETURN_NOT_OK(Finish(out));
    delivered_ = true;
    return arrow::Status::OK();
  }
  std::function<arrow::Status(std::shared_ptr<HashRelation>*)> Finish;

 private:
  bool delivered_ = false;
};

}  // namespace

arrow::Status HashRelationKernel::Make(std::shared_ptr<Context> ctx, const arrow::FieldVector& input_field_list,
                                       const std::vector<int>& key_indices, int builder_type,
                                       std::shared_ptr<KernalBase>* out, const std::vector<bool>& normalize_keys) {
  if (!ctx) {
    ARROW_ASSIGN_OR_RAISE(ctx, Context::Make(0));
  }
  ARROW_ASSIGN_OR_RAISE(auto types, GzTypes(input_field_list));
  auto k = std::make_shared<HashRelationKernel>();
  k->ctx_ = ctx;
  k->inputs_ = input_field_list;
  std::vector<int32_t> keys(key_indices.begin(), key_indices.end());
  GZ_RETURN(gz_relation_make_keys(ctx->handle(), static_cast<int32_t>(types.size()), types.data(),
                                  static_cast<int32_t>(keys.size()), keys.data(), builder_type, &k->r_),
            "HashRelationKernel::Make");
  if (std::find(normalize_keys.begin(), normalize_keys.end(), true) != normalize_keys.end()) {
    std::vector<int32_t> flags(key_indices.size(), 0);
    for (size_t j = 0; j < flags.size() && j < normalize_keys.size(); j++) flags[j] = normalize_keys[j] ? 1 : 0;
    GZ_RETURN(gz_relation_normalize_keys(k->r_, static_cast<int32_t>(flags.size()), flags.data()),
              "HashRelationKernel::Make");
  }
  k->kernel_name_ = "HashRelationKernel";
  *out = k;
  return arrow::Status::OK();
}

HashRelationKernel::~HashRelationKernel() {
  if (ctx_) (void)ctx_->Synchronize();
  pending_.clear();
  if (r_) gz_relation_release(r_);
}

arrow::Status HashRelationKernel::Evaluate(ArrayList& in) {
  if (relation_) return arrow::Status::Invalid("HashRelationKernel::Evaluate after the relation was finished");
  ARROW_RETURN_NOT_OK(CheckColumns(in, inputs_, "HashRelationKernel::Evaluate"));
  ARROW_ASSIGN_OR_RAISE(auto db, DeviceBatch::Make(in, NumRows(in)));
  gz_batch b = db->batch();
  GZ_RETURN(gz_relation_append(r_, b.num_rows, b.buf_addrs, b.buf_sizes, b.num_bufs), "HashRelationKernel::Evaluate");
  pending_.push_back(std::move(db));
  if (pending_.size() >= 64) {  // the relation copied the rows; the inputs go once the copies have run
    ARROW_RETURN_NOT_OK(ctx_->Synchronize());
    pending_.clear();
  }
  return arrow::Status::OK();
}

arrow::Status HashRelationKernel::MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                                     std::shared_ptr<ResultIterator<HashRelation>>* out) {
  auto it = std::make_shared<RelationIterator>();
  it->Finish = [self = shared_from_this()](std::shared_ptr<HashRelation>* rel) -> arrow::Status {
    if (!self->relation_) {
      GZ_RETURN(gz_relation_finish(self->r_), "HashRelation::Next");
      ARROW_RETURN_NOT_OK(self->ctx_->Synchronize());
      self->pending_.clear();
      self->relation_ = std::make_shared<HashRelation>(self->r_, self->ctx_);
      self->r_ = nullptr;  // owned by the relation now
    }
    *rel = self->relation_;
    return arrow::Status::OK();
  };
  *out = it;
  return arrow::Status::OK();
}

// ---- probe -----------------------------------------------------------------------
namespace {

class ProbeIterator : public ResultIterator<arrow::RecordBatch> {
 public:
  ProbeIterator(std::shared_ptr<Context> ctx, std::vector<int> left_keys, std::vector<int> right_keys,
                arrow::FieldVector left, arrow::FieldVector right, int join_type, bool null_aware,
                std::vector<std::pair<int, int>> result, int relation_idx)
      : ctx_(std::move(ctx)),
        left_keys_(std::move(left_keys)),
        right_keys_(std::move(right_keys)),
        left_(std::move(left)),
        right_(std::move(right)),
        join_type_(join_type),
        null_aware_(null_aware),
        result_(std::move(result)),
        relation_idx_(relation_idx) {
    arrow::FieldVector f;
    for (auto [side, col] : result_) f.push_back(side == 0 ? left_[col] : right_[col]);
    schema_ = arrow::schema(f);
  }
  ~ProbeIterator() override {
    if (p_) gz_probe_release(p_);
  }

  arrow::Status SetDependencies(const std::vector<std::shared_ptr<ResultIteratorBase>>& deps) override {
    if (relation_idx_ < 0 || relation_idx_ >= static_cast<int>(deps.size()))
      return arrow::Status::IndexError("ConditionedProbe: hash relation index ", relation_idx_, " of ", deps.size());
    auto it = std::dynamic_pointer_cast<ResultIterator<HashRelation>>(deps[relation_idx_]);
    if (!it) return arrow::Status::TypeError("ConditionedProbe: dependency ", relation_idx_, " is no hash relation");
    ARROW_RETURN_NOT_OK(it->Next(&relation_));
    ARROW_ASSIGN_OR_RAISE(auto types, GzTypes(right_));
    std::vector<int32_t> keys(right_keys_.begin(), right_keys_.end()), sides, idx;
    for (auto [s, c] : result_) {
      sides.push_back(s);
      idx.push_back(c);
    }
    if (p_) gz_probe_release(p_);
    p_ = nullptr;
    GZ_RETURN(gz_probe_make_keys(relation_->handle(), static_cast<int32_t>(types.size()), types.data(),
                                 static_cast<int32_t>(keys.size()), keys.data(), join_type_, null_aware_ ? 1 : 0,
                                 static_cast<int32_t>(sides.size()), sides.data(), idx.data(), &p_),
              "ConditionedProbe::SetDependencies");
    return arrow::Status::OK();
  }

  arrow::Status Process(const ArrayList& in, std::shared_ptr<arrow::RecordBatch>* out,
                        const std::shared_ptr<arrow::Array>& selection) override {
    if (selection) return arrow::Status::NotImplemented("ConditionedProbe: selection vectors");
    if (!p_) return arrow::Status::Invalid("ConditionedProbe::Process before SetDependencies");
    ARROW_RETURN_NOT_OK(CheckColumns(in, right_, "ConditionedProbe::Process"));
    ARROW_ASSIGN_OR_RAISE(auto db, DeviceBatch::Make(in, NumRows(in)));
    gz_batch b = db->batch();
    int64_t n = 0;
    GZ_RETURN(gz_probe_process(p_, b.num_rows, b.buf_addrs, b.buf_sizes, b.num_bufs, &n), "ConditionedProbe::Process");
    std::vector<void*> bits(result_.size());
    for (size_t c = 0; c < result_.size(); c++)
      GZ_RETURN(gz_probe_result_bitmap(p_, static_cast<int32_t>(c), &bits[c]), "ConditionedProbe::Process");
    ARROW_RETURN_NOT_OK(ctx_->Synchronize());
    ArrayList cols;
    for (size_t c = 0; c < result_.size(); c++) {
      void *vals = nullptr, *valid = nullptr, *offs = nullptr, *data = nullptr;
      GZ_RETURN(gz_probe_result_column(p_, static_cast<int32_t>(c), &vals, &valid, &offs, &data),
                "ConditionedProbe::Process");
      ARROW_ASSIGN_OR_RAISE(auto col, MakeColumn(schema_->field(static_cast<int>(c))->type(), n, vals, offs, data,
                                                 bits[c]));
      cols.push_back(col);
    }
    *out = arrow::RecordBatch::Make(schema_, n, cols);
    return arrow::Status::OK();
  }

 private:
  std::shared_ptr<Context> ctx_;
  std::vector<int> left_keys_, right_keys_;
  arrow::FieldVector left_, right_;
  int join_type_;
  bool null_aware_;
  std::vector<std::pair<int, int>> result_;
  int relation_idx_;
  std::shared_ptr<arrow::Schema> schema_;
  std::shared_ptr<HashRelation> relation_;
  gz_probe* p_ = nullptr;
};

}  // namespace

arrow::Status ConditionedProbeKernel::Make(std::shared_ptr<Context> ctx, const std::vector<int>& left_key_list,
                                           const std::vector<int>& right_key_list,
                                           const arrow::FieldVector& left_schema_list,
                                           const arrow::FieldVector& right_schema_list, const void* condition,
                                           int join_type, bool is_null_aware_anti_join,
                                           const std::vector<std::pair<int, int>>& result_schema,
                                           int hash_relation_idx, std::shared_ptr<KernalBase>* out) {
  if (condition) return arrow::Status::NotImplemented("ConditionedProbe: join conditions");
  if (join_type != GZ_JOIN_INNER && join_type != GZ_JOIN_ANTI && join_type != GZ_JOIN_SEMI)
    return arrow::Status::NotImplemented("ConditionedProbe: join type ", join_type);
  if (left_key_list.size() != right_key_list.size() || right_key_list.empty())
    return arrow::Status::Invalid("ConditionedProbe: ", left_key_list.size(), " build keys and ",
                                  right_key_list.size(), " probe keys");
  for (auto [side, col] : result_schema) {
    const auto& f = side == 0 ? left_schema_list : right_schema_list;
    if ((side != 0 && side != 1) || col < 0 || col >= static_cast<int>(f.size()))
      return arrow::Status::IndexError("ConditionedProbe: result column (", side, ", ", col, ")");
  }
  if (!ctx) {
    ARROW_ASSIGN_OR_RAISE(ctx, Context::Make(0));
  }
  auto k = std::make_shared<ConditionedProbeKernel>();
  k->ctx_ = ctx;
  k->left_keys_ = left_key_list;
  k->right_keys_ = right_key_list;
  k->left_ = left_schema_list;
  k->right_ = right_schema_list;
  k->join_type_ = join_type;
  k->null_aware_ = is_null_aware_anti_join;
  k->result_ = result_schema;
  k->relation_idx_ = hash_relation_idx;
  k->kernel_name_ = "ConditionedProbeKernel";
  *out = k;
  return arrow::Status::OK();
}

arrow::Status ConditionedProbeKernel::MakeResultIterator(std::shared_ptr<arrow::Schema> schema,
                                                         std::shared_ptr<ResultIterator<arrow::RecordBatch>>* out) {
  *out = std::make_shared<ProbeIterator>(ctx_, left_keys_, right_keys_, left_, right_, join_type_, null_aware_,
                                         result_, relation_idx_);
  return arrow::Status::OK();
}

}  // namespace gazelle_amd
