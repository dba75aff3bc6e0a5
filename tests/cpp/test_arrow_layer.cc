// The C++ operator interfaces (gazelle_amd_arrow.h) driven as the reference's
// native code drives its own: Splitter::Make / Split / Stop with the data file
// read back by Arrow's IPC reader, HashAggregateKernel's ResultIterator
// (ProcessAndCacheOne, HasNext, Next), and HashRelationKernel feeding
// ConditionedProbeKernel through SetDependencies / Process.  Every result is
// checked against the oracle (the CPU restatement, oracle/gz_oracle.h) on the
// same seeded inputs.  Needs a GPU; prints one PASS line per case and exits 1
// on the first failure.
#include <arrow/api.h>
#include <arrow/array/concatenate.h>
#include <arrow/io/memory.h>
#include <arrow/ipc/api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>

#include "../../include/gazelle_amd_arrow.h"
#include "../../oracle/gz_oracle.h"

using gazelle_amd::ArrayList;

namespace {

#define EXPECT(cond, msg)                                                                   \
  do {                                                                                      \
    if (!(cond)) {                                                                          \
      std::ostringstream _os;                                                               \
      _os << msg;                                                                           \
      std::cerr << "FAIL " << __FILE__ << ":" << __LINE__ << ": " << _os.str() << std::endl; \
      return false;                                                                         \
    }                                                                                       \
  } while (0)

#define OK(expr)                                                                          \
  do {                                                                                    \
    arrow::Status _s = (expr);                                                            \
    EXPECT(_s.ok(), #expr << ": " << _s.ToString());                                      \
  } while (0)

template <typename T>
T Value(arrow::Result<T> r, const char* what) {
  if (!r.ok()) {
    std::cerr << "FAIL " << what << ": " << r.status().ToString() << std::endl;
    std::exit(1);
  }
  return std::move(r).ValueUnsafe();
}

// ---- seeded inputs ---------------------------------------------------------------
struct Gen {
  std::mt19937_64 rng;
  explicit Gen(uint64_t seed) : rng(seed) {}
  double U() { return std::uniform_real_distribution<double>(0, 1)(rng); }
  int64_t I(int64_t lo, int64_t hi) { return std::uniform_int_distribution<int64_t>(lo, hi - 1)(rng); }

  std::shared_ptr<arrow::Array> Column(const std::shared_ptr<arrow::DataType>& t, int64_t n, double null_frac,
                                       int64_t card) {
    switch (t->id()) {
      case arrow::Type::INT64: {
        arrow::Int64Builder b;
        for (int64_t i = 0; i < n; i++) U() < null_frac ? (void)b.AppendNull() : (void)b.Append(I(0, card));
        return Value(b.Finish(), "int64");
      }
      case arrow::Type::INT32: {
        arrow::Int32Builder b;
        for (int64_t i = 0; i < n; i++)
          U() < null_frac ? (void)b.AppendNull() : (void)b.Append(static_cast<int32_t>(I(1, card)));
        return Value(b.Finish(), "int32");
      }
      case arrow::Type::DOUBLE: {
        arrow::DoubleBuilder b;
        std::normal_distribution<double> nd(100.0, 25.0);
        for (int64_t i = 0; i < n; i++) U() < null_frac ? (void)b.AppendNull() : (void)b.Append(nd(rng));
        return Value(b.Finish(), "double");
      }
      case arrow::Type::BOOL: {
        arrow::BooleanBuilder b;
        for (int64_t i = 0; i < n; i++) U() < null_frac ? (void)b.AppendNull() : (void)b.Append(U() < 0.5);
        return Value(b.Finish(), "bool");
      }
      case arrow::Type::STRING: {
        arrow::StringBuilder b;
        for (int64_t i = 0; i < n; i++) {
          if (U() < null_frac) {
            (void)b.AppendNull();
          } else if (card > 0) {
            (void)b.Append("key-" + std::to_string(I(0, card)));
          } else {
            std::string s(static_cast<size_t>(I(0, 13)), 'a');
            for (auto& ch : s) ch = static_cast<char>('a' + I(0, 26));
            (void)b.Append(s);
          }
        }
        return Value(b.Finish(), "utf8");
      }
      default:
        std::cerr << "no generator for " << t->ToString() << std::endl;
        std::exit(1);
    }
  }

  std::shared_ptr<arrow::RecordBatch> Batch(const std::shared_ptr<arrow::Schema>& s, int64_t n,
                                            const std::vector<double>& null_frac, const std::vector<int64_t>& card) {
    ArrayList cols;
    for (int c = 0; c < s->num_fields(); c++) cols.push_back(Column(s->field(c)->type(), n, null_frac[c], card[c]));
    return arrow::RecordBatch::Make(s, n, cols);
  }
};

// The oracle's view of a batch: three host buffer slots per column, row 0 at
// bit / element 0 (slices are first copied so).
struct OracleBatch {
  std::shared_ptr<arrow::RecordBatch> rb;
  std::vector<const void*> bufs;
  gzo_batch b;
  explicit OracleBatch(const std::shared_ptr<arrow::RecordBatch>& in) {
    ArrayList cols;
    for (const auto& a : in->columns()) cols.push_back(Value(arrow::Concatenate({a}), "Concatenate"));
    rb = arrow::RecordBatch::Make(in->schema(), in->num_rows(), cols);
    for (const auto& a : cols) {
      const auto& d = *a->data();
      bufs.push_back(a->null_count() > 0 ? d.buffers[0]->data() : nullptr);
      bool bin = a->type_id() == arrow::Type::STRING;
      bufs.push_back(bin ? d.buffers[1]->data() : nullptr);
      bufs.push_back(a->type_id() == arrow::Type::NA ? nullptr : d.buffers[bin ? 2 : 1]->data());
    }
    b.num_rows = rb->num_rows();
    b.bufs = bufs.data();
  }
};

std::vector<int32_t> Types(const std::shared_ptr<arrow::Schema>& s) {
  std::vector<int32_t> t;
  for (const auto& f : s->fields()) t.push_back(f->type()->id());
  return t;
}

bool RowEquals(const arrow::RecordBatch& a, int64_t i, const arrow::RecordBatch& b, int64_t j, int skip_b = 0) {
  for (int c = 0; c < a.num_columns(); c++)
    if (!a.column(c)->RangeEquals(i, i + 1, j, *b.column(c + skip_b))) return false;
  return true;
}

std::string TmpPath(const std::string& name) {
  const char* d = getenv("TMPDIR");
  return std::string(d && *d ? d : "/tmp") + "/" + name;
}

// ---- shuffle ---------------------------------------------------------------------
// Every partition's stream in the data file holds the input rows of that
// partition id, in input order (a "range" split drops its pid column).
bool CheckDataFile(gazelle_amd::Splitter& sp, const std::vector<std::shared_ptr<arrow::RecordBatch>>& in,
                   const std::vector<std::vector<int32_t>>& pids, int P, int skip_cols) {
  std::ifstream f(sp.DataFile(), std::ios::binary);
  std::string bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  const auto& lengths = sp.PartitionLengths();
  int64_t total = 0;
  for (auto l : lengths) total += l;
  EXPECT(static_cast<int64_t>(bytes.size()) == total, "data file " << bytes.size() << " bytes, lengths sum " << total);
  EXPECT(sp.TotalBytesWritten() == total, "TotalBytesWritten " << sp.TotalBytesWritten());
  EXPECT(static_cast<int>(sp.RawPartitionLengths().size()) == P, "RawPartitionLengths size");
  int64_t pos = 0;
  for (int p = 0; p < P; p++) {
    std::vector<std::pair<int, int64_t>> want;
    for (size_t b = 0; b < in.size(); b++)
      for (int64_t r = 0; r < in[b]->num_rows(); r++)
        if (pids[b][r] == p) want.emplace_back(static_cast<int>(b), r);
    EXPECT((lengths[p] == 0) == want.empty(), "partition " << p << ": " << lengths[p] << " bytes, " << want.size()
                                                             << " rows");
    EXPECT((sp.RawPartitionLengths()[p] == 0) == want.empty(), "raw length of partition " << p);
    if (want.empty()) continue;
    auto buf = std::make_shared<arrow::Buffer>(reinterpret_cast<const uint8_t*>(bytes.data()) + pos, lengths[p]);
    auto reader = Value(arrow::ipc::RecordBatchStreamReader::Open(std::make_shared<arrow::io::BufferReader>(buf)),
                        "open partition stream");
    EXPECT(reader->schema()->num_fields() == in[0]->num_columns() - skip_cols, "partition schema");
    size_t k = 0;
    while (true) {
      std::shared_ptr<arrow::RecordBatch> rb;
      OK(reader->ReadNext(&rb));
      if (!rb) break;
      for (int64_t i = 0; i < rb->num_rows(); i++, k++) {
        EXPECT(k < want.size(), "partition " << p << " has extra rows");
        EXPECT(RowEquals(*rb, i, *in[want[k].first], want[k].second, skip_cols),
               "partition " << p << " row " << k << " differs from input batch " << want[k].first << " row "
                            << want[k].second);
      }
    }
    EXPECT(k == want.size(), "partition " << p << ": " << k << " rows, expected " << want.size());
    pos += lengths[p];
  }
  return true;
}

bool TestHashSplit(const std::shared_ptr<gazelle_amd::Context>& ctx, bool lz4) {
  auto schema = arrow::schema({arrow::field("k", arrow::int64()), arrow::field("d", arrow::float64()),
                               arrow::field("s", arrow::utf8()), arrow::field("q", arrow::int32()),
                               arrow::field("b", arrow::boolean())});
  Gen g(lz4 ? 11 : 7);
  std::vector<double> nulls{0.05, 0.1, 0.05, 0.0, 0.1};
  std::vector<int64_t> card{5000, 0, 0, 50, 0};
  std::vector<std::shared_ptr<arrow::RecordBatch>> in{g.Batch(schema, 3000, nulls, card), g.Batch(schema, 1, nulls, card),
                                                      g.Batch(schema, 4096, nulls, card),
                                                      g.Batch(schema, 5000, nulls, card)->Slice(123, 2500)};
  const int P = 37;
  const std::vector<int> keys{0, 2};
  gazelle_amd::SplitOptions opts;
  opts.buffer_size = 500;
  opts.data_file = TmpPath(lz4 ? "gz_arrow_split_lz4.data" : "gz_arrow_split.data");
  auto sp = Value(gazelle_amd::Splitter::Make("hash", schema, P, keys, opts, ctx), "Splitter::Make");
  if (lz4) OK(sp->SetCompressType(arrow::Compression::LZ4_FRAME));
  EXPECT(sp->CompressedSize(*in[0]) > 0, "CompressedSize");
  for (const auto& b : in) OK(sp->Split(*b));
  OK(sp->Stop());
  EXPECT(sp->TotalComputePidTime() > 0 && sp->TotalWriteTime() > 0, "split metrics");
  EXPECT(sp->RawPartitionBytes() > 0, "RawPartitionBytes");
  auto types = Types(schema);
  std::vector<int32_t> keys32(keys.begin(), keys.end());
  std::vector<std::vector<int32_t>> pids;
  for (const auto& b : in) {
    OracleBatch ob(b);
    std::vector<int32_t> pid(b->num_rows());
    int32_t rr = 0;
    EXPECT(gzo_split_pids(0, &ob.b, static_cast<int>(types.size()), types.data(), keys32.data(), 2, P, 42, &rr,
                          pid.data()) == 0,
           "oracle pids");
    pids.push_back(pid);
  }
  if (!CheckDataFile(*sp, in, pids, P, 0)) return false;
  std::remove(opts.data_file.c_str());
  return true;
}

bool TestRangeSplit(const std::shared_ptr<gazelle_amd::Context>& ctx) {
  auto schema = arrow::schema({arrow::field("pid", arrow::int32()), arrow::field("v", arrow::int64()),
                               arrow::field("s", arrow::utf8())});
  Gen g(3);
  const int P = 5;
  std::vector<std::shared_ptr<arrow::RecordBatch>> in;
  std::vector<std::vector<int32_t>> pids;
  for (int64_t n : {1000, 2000}) {
    arrow::Int32Builder pb;
    std::vector<int32_t> pid;
    for (int64_t i = 0; i < n; i++) {
      pid.push_back(static_cast<int32_t>(g.I(0, P)));
      (void)pb.Append(pid.back());
    }
    in.push_back(arrow::RecordBatch::Make(schema, n, {Value(pb.Finish(), "pid"), g.Column(arrow::int64(), n, 0.1, 1 << 30),
                                                      g.Column(arrow::utf8(), n, 0.1, 0)}));
    pids.push_back(pid);
  }
  gazelle_amd::SplitOptions opts;
  opts.data_file = TmpPath("gz_arrow_range.data");
  auto sp = Value(gazelle_amd::Splitter::Make("range", schema, P, opts, ctx), "Splitter::Make range");
  for (const auto& b : in) OK(sp->Split(*b));
  OK(sp->Stop());
  if (!CheckDataFile(*sp, in, pids, P, 1)) return false;
  std::remove(opts.data_file.c_str());
  return true;
}

// ---- aggregate -------------------------------------------------------------------
struct OracleAction {
  int32_t kind, in0, in1, nan;
};

// Result columns equal the oracle's: validity, then values by their bits except
// the FP64 sums and averages (free summation order: within 1e-12).
bool CheckAggregate(const std::shared_ptr<arrow::RecordBatch>& got, gzo_agg* o, const std::vector<bool>& approx) {
  int64_t g = gzo_agg_num_groups(o);
  EXPECT(got->num_rows() == g, "groups " << got->num_rows() << ", oracle " << g);
  EXPECT(got->num_columns() == gzo_agg_num_results(o), "result columns");
  for (int c = 0; c < got->num_columns(); c++) {
    auto col = got->column(c);
    std::vector<uint8_t> valid(g + 1);
    int64_t nbytes = gzo_agg_result_bytes(o, c);
    if (nbytes >= 0) {
      std::vector<int32_t> offs(g + 1);
      std::vector<uint8_t> data(nbytes + 1);
      gzo_agg_result_binary(o, c, offs.data(), data.data(), valid.data());
      auto s = std::static_pointer_cast<arrow::StringArray>(col);
      for (int64_t i = 0; i < g; i++) {
        EXPECT(s->IsValid(i) == (valid[i] != 0), "utf8 result " << c << " validity at " << i);
        if (!valid[i]) continue;
        std::string_view want(reinterpret_cast<const char*>(data.data()) + offs[i], offs[i + 1] - offs[i]);
        EXPECT(s->GetView(i) == want, "utf8 result " << c << " group " << i);
      }
      continue;
    }
    int w = std::static_pointer_cast<arrow::FixedWidthType>(col->type())->bit_width() / 8;
    std::vector<uint8_t> vals(g * w + 8);
    gzo_agg_result(o, c, vals.data(), valid.data());
    const uint8_t* mine = col->data()->buffers[1]->data() + col->offset() * w;
    for (int64_t i = 0; i < g; i++) {
      EXPECT(col->IsValid(i) == (valid[i] != 0), "result " << c << " validity at group " << i);
      if (!valid[i]) continue;
      if (approx[c]) {
        double a, b;
        std::memcpy(&a, mine + i * 8, 8);
        std::memcpy(&b, vals.data() + i * 8, 8);
        EXPECT(std::fabs(a - b) <= 1e-12 * std::fabs(b), "result " << c << " group " << i << ": " << a << " vs " << b);
      } else {
        EXPECT(std::memcmp(mine + i * w, vals.data() + i * w, w) == 0, "result " << c << " group " << i);
      }
    }
  }
  return true;
}

bool RunAggregate(const std::shared_ptr<gazelle_amd::Context>& ctx, const std::shared_ptr<arrow::Schema>& schema,
                  const std::vector<std::shared_ptr<arrow::RecordBatch>>& in,
                  const std::vector<gazelle_amd::Action>& actions, const std::vector<OracleAction>& oacts,
                  const arrow::FieldVector& results, const std::vector<bool>& approx, int64_t min_batches) {
  std::shared_ptr<gazelle_amd::KernalBase> k;
  OK(gazelle_amd::HashAggregateKernel::Make(ctx, schema->fields(), actions, results, &k));
  std::shared_ptr<gazelle_amd::ResultIterator<arrow::RecordBatch>> it;
  OK(k->MakeResultIterator(arrow::schema(results), &it));
  for (const auto& b : in) OK(it->ProcessAndCacheOne(b->columns()));
  std::vector<std::shared_ptr<arrow::RecordBatch>> out;
  while (it->HasNext()) {
    std::shared_ptr<arrow::RecordBatch> rb;
    OK(it->Next(&rb));
    EXPECT(rb->num_rows() <= 10000, "a Next batch of " << rb->num_rows() << " rows");
    out.push_back(rb);
  }
  EXPECT(static_cast<int64_t>(out.size()) >= min_batches, out.size() << " result batches");
  auto all = Value(arrow::Table::FromRecordBatches(arrow::schema(results), out), "table");
  auto merged = Value(all->CombineChunksToBatch(), "combine");

  auto types = Types(schema);
  std::vector<int32_t> kinds, in0, in1, nan;
  for (const auto& a : oacts) {
    kinds.push_back(a.kind);
    in0.push_back(a.in0);
    in1.push_back(a.in1);
    nan.push_back(a.nan);
  }
  gzo_agg* o = gzo_agg_make(static_cast<int>(types.size()), types.data(), static_cast<int>(kinds.size()),
                            kinds.data(), in0.data(), in1.data(), nan.data());
  EXPECT(o, "oracle aggregate");
  for (const auto& b : in) {
    OracleBatch ob(b);
    EXPECT(gzo_agg_process(o, &ob.b) == 0, "oracle process");
  }
  bool ok = CheckAggregate(merged, o, approx);
  gzo_agg_free(o);
  return ok;
}

bool TestAggregate(const std::shared_ptr<gazelle_amd::Context>& ctx) {
  auto schema = arrow::schema({arrow::field("k", arrow::int64()), arrow::field("price", arrow::float64()),
                               arrow::field("qty", arrow::int32()), arrow::field("name", arrow::utf8())});
  Gen g(21);
  std::vector<double> nulls{0.02, 0.1, 0.0, 0.05};
  std::vector<int64_t> card{25000, 0, 50, 300};
  std::vector<std::shared_ptr<arrow::RecordBatch>> in{g.Batch(schema, 30000, nulls, card),
                                                      g.Batch(schema, 30000, nulls, card)->Slice(77, 20000),
                                                      g.Batch(schema, 30000, nulls, card)};
  using A = gazelle_amd::Action;
  std::vector<A> actions{{"action_groupby", {0}}, {"action_sum", {1}},     {"action_count", {1}},
                         {"action_avg", {1}},     {"action_min", {2}},     {"action_max_true", {1}},
                         {"action_sum_count", {1}}, {"action_countLiteral_1", {}}};
  std::vector<OracleAction> oacts{{GZ_ACTION_GROUPBY, 0, -1, 0}, {GZ_ACTION_SUM, 1, -1, 0},
                                  {GZ_ACTION_COUNT, 1, -1, 0},   {GZ_ACTION_AVG, 1, -1, 0},
                                  {GZ_ACTION_MIN, 2, -1, 0},     {GZ_ACTION_MAX, 1, -1, 1},
                                  {GZ_ACTION_SUM_COUNT, 1, -1, 0}, {GZ_ACTION_COUNT_LITERAL, -1, -1, 0}};
  arrow::FieldVector results{arrow::field("k", arrow::int64()),      arrow::field("sum", arrow::float64()),
                             arrow::field("count", arrow::int64()),  arrow::field("avg", arrow::float64()),
                             arrow::field("min", arrow::int32()),    arrow::field("max", arrow::float64()),
                             arrow::field("psum", arrow::float64()), arrow::field("pcount", arrow::int64()),
                             arrow::field("rows", arrow::int64())};
  if (!RunAggregate(ctx, schema, in, actions, oacts, results,
                    {false, true, false, true, false, false, true, false, false}, 3))
    return false;
  // a utf8 group-by key
  std::vector<A> sact{{"action_groupby", {3}}, {"action_sum", {2}}, {"action_countLiteral_1", {}}};
  std::vector<OracleAction> soact{{GZ_ACTION_GROUPBY, 3, -1, 0}, {GZ_ACTION_SUM, 2, -1, 0},
                                  {GZ_ACTION_COUNT_LITERAL, -1, -1, 0}};
  arrow::FieldVector sres{arrow::field("name", arrow::utf8()), arrow::field("qty", arrow::int64()),
                          arrow::field("rows", arrow::int64())};
  if (!RunAggregate(ctx, schema, in, sact, soact, sres, {false, false, false}, 1)) return false;
  // the global aggregate: one row
  std::vector<A> gact{{"action_sum", {1}}, {"action_count", {1}}, {"action_min", {1}}, {"action_max", {2}}};
  std::vector<OracleAction> goact{{GZ_ACTION_SUM, 1, -1, 0}, {GZ_ACTION_COUNT, 1, -1, 0}, {GZ_ACTION_MIN, 1, -1, 0},
                                  {GZ_ACTION_MAX, 2, -1, 0}};
  arrow::FieldVector gres{arrow::field("sum", arrow::float64()), arrow::field("count", arrow::int64()),
                          arrow::field("min", arrow::float64()), arrow::field("max", arrow::int32())};
  return RunAggregate(ctx, schema, in, gact, goact, gres, {true, false, false, false}, 1);
}

// ---- join ------------------------------------------------------------------------
bool TestJoin(const std::shared_ptr<gazelle_amd::Context>& ctx) {
  auto bschema = arrow::schema({arrow::field("k", arrow::int64()), arrow::field("v", arrow::float64()),
                                arrow::field("s", arrow::utf8())});
  auto pschema = arrow::schema({arrow::field("k", arrow::int64()), arrow::field("x", arrow::int32())});
  Gen g(5);
  std::vector<std::shared_ptr<arrow::RecordBatch>> build{g.Batch(bschema, 1500, {0.03, 0.2, 0.1}, {3000, 0, 0}),
                                                         g.Batch(bschema, 700, {0.03, 0.2, 0.1}, {3000, 0, 0})};
  std::vector<std::shared_ptr<arrow::RecordBatch>> probe{g.Batch(pschema, 5000, {0.02, 0.3}, {4000, 9}),
                                                         g.Batch(pschema, 4000, {0.02, 0.3}, {4000, 9})->Slice(9, 3333)};
  std::shared_ptr<gazelle_amd::KernalBase> rk;
  OK(gazelle_amd::HashRelationKernel::Make(ctx, bschema->fields(), {0}, 1, &rk));
  for (const auto& b : build) {
    ArrayList cols = b->columns();
    OK(rk->Evaluate(cols));
  }
  std::shared_ptr<gazelle_amd::ResultIterator<gazelle_amd::HashRelation>> rel_it;
  OK(rk->MakeResultIterator(nullptr, &rel_it));

  gzo_rel* o = gzo_rel_make(GZ_TYPE_INT64, 1);
  std::vector<std::pair<int, int64_t>> build_row;  // global build row -> (batch, row)
  std::vector<OracleBatch> ob;
  for (size_t b = 0; b < build.size(); b++) {
    ob.emplace_back(build[b]);
    const auto& d = *ob.back().rb->column(0)->data();
    gzo_rel_append(o, d.buffers[1]->data(), ob.back().rb->column(0)->null_count() ? d.buffers[0]->data() : nullptr,
                   build[b]->num_rows());
    for (int64_t r = 0; r < build[b]->num_rows(); r++) build_row.emplace_back(static_cast<int>(b), r);
  }
  gzo_rel_finish(o);

  struct Case {
    const char* name;
    int type;
    bool null_aware;
  };
  for (Case cs : {Case{"inner", GZ_JOIN_INNER, false}, Case{"semi", GZ_JOIN_SEMI, false},
                  Case{"anti", GZ_JOIN_ANTI, false}, Case{"null-aware anti", GZ_JOIN_ANTI, true}}) {
    std::vector<std::pair<int, int>> result{{1, 0}, {1, 1}};
    if (cs.type == GZ_JOIN_INNER) result = {{1, 0}, {1, 1}, {0, 1}, {0, 2}};
    std::shared_ptr<gazelle_amd::KernalBase> pk;
    OK(gazelle_amd::ConditionedProbeKernel::Make(ctx, {0}, {0}, bschema->fields(), pschema->fields(), nullptr,
                                                 cs.type, cs.null_aware, result, 0, &pk));
    std::shared_ptr<gazelle_amd::ResultIterator<arrow::RecordBatch>> it;
    OK(pk->MakeResultIterator(nullptr, &it));
    OK(it->SetDependencies({rel_it}));
    for (const auto& pb : probe) {
      std::shared_ptr<arrow::RecordBatch> out;
      OK(it->Process(pb->columns(), &out));
      OracleBatch opb(pb);
      const auto& d = *opb.rb->column(0)->data();
      const uint8_t* pv = opb.rb->column(0)->null_count() ? d.buffers[0]->data() : nullptr;
      int64_t n = gzo_probe(o, cs.type, cs.null_aware, d.buffers[1]->data(), pv, pb->num_rows(), nullptr, nullptr);
      std::vector<int64_t> prow(n + 1), brow(n + 1);
      gzo_probe(o, cs.type, cs.null_aware, d.buffers[1]->data(), pv, pb->num_rows(), prow.data(), brow.data());
      EXPECT(out->num_rows() == n, cs.name << ": " << out->num_rows() << " rows, oracle " << n);
      for (int64_t i = 0; i < n; i++) {
        for (size_t c = 0; c < result.size(); c++) {
          auto [side, col] = result[c];
          bool eq = side == 1 ? out->column(c)->RangeEquals(i, i + 1, prow[i], *pb->column(col))
                              : out->column(c)->RangeEquals(i, i + 1, build_row[brow[i]].second,
                                                            *build[build_row[brow[i]].first]->column(col));
          EXPECT(eq, cs.name << ": output row " << i << " column " << c);
        }
      }
    }
    std::cout << "PASS join " << cs.name << std::endl;
  }
  gzo_rel_free(o);
  return true;
}

}  // namespace

int main() {
  auto ctx = Value(gazelle_amd::Context::Make(0), "Context::Make");
  struct T {
    const char* name;
    std::function<bool()> run;
  };
  std::vector<T> tests{{"split hash", [&] { return TestHashSplit(ctx, false); }},
                       {"split hash lz4", [&] { return TestHashSplit(ctx, true); }},
                       {"split range", [&] { return TestRangeSplit(ctx); }},
                       {"aggregate", [&] { return TestAggregate(ctx); }},
                       {"join", [&] { return TestJoin(ctx); }}};
  for (auto& t : tests) {
    std::cout << "RUN " << t.name << std::endl;
    if (!t.run()) {
      std::cout << "FAILED " << t.name << std::endl;
      return 1;
    }
    std::cout << "PASS " << t.name << std::endl;
  }
  std::cout << "ALL PASS" << std::endl;
  return 0;
}
