// Fragment server: continuous batching of query fragments on one MI355X.
//
// Fragments (an aggregate, a join followed by an aggregate, a shuffle split)
// arrive on a timeline.  A dispatcher admits each one, first come first served,
// when a slot is idle and the fragment's HBM estimate fits what the admitted
// fragments leave of the budget; each slot is a worker thread with its own
// context (HIP stream and memory pool), so admitted fragments overlap on the
// device.  Hash relations registered with the server stay resident and serve
// the probes of every join fragment, from every slot's stream.
//
// The reference runs one kernel instance per Spark task and has no server of
// its own; this is the serving form of the same operators (gz_agg_*, gz_probe_*,
// gz_split_*), called through their C entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

constexpr int kMaxJoinAggCols = 16;  // join result columns an aggregate may follow

// Validity bytes (1 = valid) of a join output column to an Arrow bitmap, one
// thread per output byte (8 rows); bytes from (n + 7) / 8 up to nbytes are 0xff.
__global__ void pack_valid_kernel(const uint8_t* valid, int64_t n, uint8_t* bits, int64_t nbytes) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nbytes;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint32_t b = 0;
    for (int k = 0; k < 8; k++) {
      int64_t r = i * 8 + k;
      if (r >= n || valid[r]) b |= 1u << k;
    }
    bits[i] = (uint8_t)b;
  }
}

// Validity bytes -> an Arrow bitmap (LSB first) padded with 1 bits to a multiple
// of 64 rows (Arrow's 8-byte buffer padding), on the context's stream.
int pack_validity(gz_context* ctx, const uint8_t* valid, int64_t n, uint8_t* bits) {
  if (n <= 0) return GZ_OK;
  const int64_t nb = (n + 63) / 64 * 8;
  const unsigned g = (unsigned)std::min<int64_t>((nb + 255) / 256, 4096);
  hipLaunchKernelGGL(pack_valid_kernel, dim3(g), dim3(256), 0, ctx->stream, valid, n, bits, nb);
  GZ_HIP(hipGetLastError());
  return GZ_OK;
}

}  // namespace gz

using namespace gz;

namespace {

struct AggPlanCopy {
  std::vector<int32_t> types;
  std::vector<gz_agg_action> actions;
  bool set = false;
  void copy(const gz_agg_plan* p) {
    if (!p) return;
    types.assign(p->col_types, p->col_types + p->num_cols);
    actions.assign(p->actions, p->actions + p->num_actions);
    set = true;
  }
  int num_lanes_bound() const { return 2 * (int)actions.size() + 2; }
};

using Clock = std::chrono::steady_clock;

struct Slot {
  gz_context* ctx = nullptr;
  std::thread th;
  int frag = -1;  // assigned fragment, -1 idle
  bool stop = false;
  DBuf bits[kMaxJoinAggCols];
  std::vector<DBuf> tmp;
};

}  // namespace

struct gz_server {
  int device = 0;
  int64_t budget = 0;
  AggPlanCopy agg, join_agg;
  std::vector<int32_t> split_types, split_keys;
  int32_t split_P = 0;
  bool split_set = false;
  std::vector<int32_t> probe_types, probe_sides, probe_idx;
  int32_t probe_key = 0, join_type = 0;
  bool probe_set = false;
  std::vector<gz_relation*> relations;
  std::vector<Slot> slots;
  // run state
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  const gz_fragment* frags = nullptr;
  gz_fragment_stats* stats = nullptr;
  std::vector<int64_t> est;
  int64_t in_use = 0;
  int done = 0;
  bool keep = false;
  Clock::time_point t0;
  std::vector<gz_agg*> agg_out;
  std::vector<gz_splitter*> split_out;
  int first_failed = -1;    // fragment of the first failure and its message
  std::string first_msg;
};

namespace {

double since(const gz_server* s) {
  return std::chrono::duration<double>(Clock::now() - s->t0).count();
}

int64_t rows_of(const gz_fragment& f) {
  int64_t n = 0;
  for (int b = 0; b < f.num_batches; b++) n += f.batches[b].num_rows;
  return n;
}

int64_t bytes_of(const gz_fragment& f) {
  int64_t n = 0;
  for (int b = 0; b < f.num_batches; b++)
    for (int k = 0; k < f.batches[b].num_bufs; k++) n += std::max<int64_t>(0, f.batches[b].buf_sizes[k]);
  return n;
}

// HBM a fragment may allocate: hash tables sized for every row being a group
// (load 0.5) plus partition records; split outputs of the input's size plus the
// per-tile counters; join outputs of one row per probe row (unique build keys
// bound the matches; larger outputs fail with an out-of-memory status).
int64_t estimate(const gz_server* s, const gz_fragment& f) {
  const int64_t rows = rows_of(f);
  auto agg_bytes = [&](const AggPlanCopy& a) {
    const int64_t lanes = a.num_lanes_bound();
    return rows * 2 * 8 * (2 + lanes) + rows * 8 * (2 + 8) + (64 << 20);
  };
  switch (f.kind) {
    case GZ_FRAG_AGGREGATE: return agg_bytes(s->agg);
    case GZ_FRAG_SPLIT: {
      const int64_t tiles = rows / 1024 + f.num_batches + 1;
      return bytes_of(f) + rows / 4 + tiles * (int64_t)s->split_P * 4 * 6 + (64 << 20);
    }
    case GZ_FRAG_JOIN_AGGREGATE: {
      // materialized join output (the fallback) plus a group table that may
      // round up to twice the aggregate's bound
      int64_t w = 0;
      for (size_t c = 0; c < s->probe_sides.size(); c++) w += 9;
      return rows * (w + 8) + 2 * agg_bytes(s->join_agg);
    }
  }
  return 0;
}

int run_agg(gz_server* s, Slot& sl, const AggPlanCopy& plan, int32_t nb, const gz_batch* batches,
            gz_agg** out, int64_t* groups) {
  gz_agg* a = nullptr;
  GZ_TRY(gz_agg_make(sl.ctx, (int32_t)plan.types.size(), plan.types.data(), (int32_t)plan.actions.size(),
                     plan.actions.data(), nullptr, &a));
  int rc = gz_agg_process_batches(a, nb, batches);
  if (!rc) rc = gz_agg_finish(a);
  if (!rc) rc = gz_agg_num_groups(a, groups);
  if (rc) {
    gz_agg_release(a);
    return rc;
  }
  *out = a;
  return GZ_OK;
}

int run_fragment(gz_server* s, Slot& sl, int i, int64_t* out_rows) {
  const gz_fragment& f = s->frags[i];
  GZ_HIP(hipSetDevice(s->device));
  switch (f.kind) {
    case GZ_FRAG_AGGREGATE: {
      if (!s->agg.set) return fail(GZ_INVALID, "no aggregate plan");
      gz_agg* a = nullptr;
      GZ_TRY(run_agg(s, sl, s->agg, f.num_batches, f.batches, &a, out_rows));
      if (s->keep) s->agg_out[i] = a;
      else gz_agg_release(a);
      return GZ_OK;
    }
    case GZ_FRAG_SPLIT: {
      if (!s->split_set) return fail(GZ_INVALID, "no split plan");
      gz_splitter* sp = nullptr;
      GZ_TRY(gz_split_make(sl.ctx, "hash", (int32_t)s->split_types.size(), s->split_types.data(), s->split_P,
                           s->split_keys.data(), (int32_t)s->split_keys.size(), nullptr, &sp));
      int rc = gz_split_batches(sp, f.num_batches, f.batches);
      if (!rc) rc = gz_split_stop(sp);
      if (rc) {
        gz_split_release(sp);
        return rc;
      }
      *out_rows = rows_of(f);
      if (s->keep) s->split_out[i] = sp;
      else gz_split_release(sp);
      return GZ_OK;
    }
    case GZ_FRAG_JOIN_AGGREGATE: {
      if (!s->probe_set || !s->join_agg.set) return fail(GZ_INVALID, "no join plan");
      if (f.relation < 0 || f.relation >= (int)s->relations.size()) return fail(GZ_INDEX_ERROR, "unknown relation");
      if (s->join_type == GZ_JOIN_INNER) {
        // the fused stage: the probe updates the aggregate, no join output
        gz_agg* a = nullptr;
        int rc = gz_join_agg_make(s->relations[f.relation], sl.ctx, (int32_t)s->probe_types.size(),
                                  s->probe_types.data(), s->probe_key, (int32_t)s->probe_sides.size(),
                                  s->probe_sides.data(), s->probe_idx.data(), (int32_t)s->join_agg.actions.size(),
                                  s->join_agg.actions.data(), nullptr, &a);
        if (rc == GZ_OK) {
          rc = gz_join_agg_process_batches(a, f.num_batches, f.batches);
          if (!rc) rc = gz_agg_finish(a);
          if (!rc) rc = gz_agg_num_groups(a, out_rows);
          if (rc) {
            gz_agg_release(a);
            return rc;
          }
          if (s->keep) s->agg_out[i] = a;
          else gz_agg_release(a);
          return GZ_OK;
        }
        if (rc != GZ_NOT_IMPLEMENTED) return rc;
        // plans the fused stage does not cover materialize the join output
      }
      gz_probe* pr = nullptr;
      GZ_TRY(gz_probe_make(s->relations[f.relation], (int32_t)s->probe_types.size(), s->probe_types.data(),
                           s->probe_key, s->join_type, 0, (int32_t)s->probe_sides.size(), s->probe_sides.data(),
                           s->probe_idx.data(), &pr));
      probe_set_context(pr, sl.ctx);
      int64_t n = 0;
      int rc = gz_probe_process_batches(pr, f.num_batches, f.batches, &n, nullptr);
      // the join output as one Arrow batch: validity bitmaps packed from bytes
      const int nres = (int)s->probe_sides.size();
      std::vector<int64_t> addrs, sizes;
      for (int c = 0; c < nres && !rc; c++) {
        void *vals = nullptr, *valid = nullptr, *offs = nullptr, *data = nullptr;
        rc = gz_probe_result_column(pr, c, &vals, &valid, &offs, &data);
        if (rc) break;
        rc = ctx_ensure(sl.ctx, &sl.bits[c], (size_t)std::max<int64_t>((n + 7) / 8, 8));
        if (rc) break;
        if (n > 0) {
          unsigned g = (unsigned)std::min<int64_t>(((n + 7) / 8 + 255) / 256, 4096);
          hipLaunchKernelGGL(pack_valid_kernel, dim3(g), dim3(256), 0, sl.ctx->stream, (const uint8_t*)valid, n,
                             (uint8_t*)sl.bits[c].p, (n + 7) / 8);
        }
        addrs.push_back((int64_t)(uintptr_t)sl.bits[c].p);
        sizes.push_back((n + 7) / 8);
        addrs.push_back((int64_t)(uintptr_t)vals);
        sizes.push_back(n * type_width(s->join_agg.types[c]));
      }
      if (!rc) {
        gz_batch b{n, addrs.data(), sizes.data(), (int32_t)addrs.size()};
        gz_agg* a = nullptr;
        rc = run_agg(s, sl, s->join_agg, n > 0 ? 1 : 0, &b, &a, out_rows);
        if (!rc) {
          if (s->keep) s->agg_out[i] = a;
          else gz_agg_release(a);
        }
      }
      gz_probe_release(pr);
      return rc;
    }
  }
  return fail(GZ_INVALID, "unknown fragment kind");
}

void worker(gz_server* s, int k) {
  Slot& sl = s->slots[k];
  hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] { return sl.stop || sl.frag >= 0; });
    if (sl.stop) return;
    const int i = sl.frag;
    lk.unlock();
    int64_t out_rows = 0;
    int rc = run_fragment(s, sl, i, &out_rows);
    if (rc == GZ_OK) {
      hipError_t e = hipStreamSynchronize(sl.ctx->stream);
      if (e != hipSuccess) rc = fail(GZ_HIP_ERROR, hipGetErrorString(e));
    }
    std::string msg = rc == GZ_OK ? std::string() : std::string(gz_last_error());
    lk.lock();
    if (rc != GZ_OK && s->first_failed < 0) {
      s->first_failed = i;
      s->first_msg = msg;
    }
    gz_fragment_stats& st = s->stats[i];
    st.done_s = since(s);
    st.status = rc;
    st.out_rows = out_rows;
    s->in_use -= s->est[i];
    s->done++;
    sl.frag = -1;
    s->cv_done.notify_all();
  }
}

}  // namespace

extern "C" {

int gz_server_create(int32_t device, int64_t hbm_budget, int32_t num_slots, const gz_agg_plan* agg,
                     const gz_split_plan* split, const gz_probe_plan* probe, const gz_agg_plan* join_agg,
                     gz_server** out) {
  if (!out) return fail(GZ_INVALID, "gz_server_create: null output");
  if (num_slots < 1 || num_slots > 32) return fail(GZ_INVALID, "slots must be in [1, 32]");
  if (probe && probe->num_result_cols > kMaxJoinAggCols)
    return fail(GZ_NOT_IMPLEMENTED, "too many join result columns for the following aggregate");
  if (probe && (!join_agg || join_agg->num_cols != probe->num_result_cols))
    return fail(GZ_INVALID, "the join aggregate reads the join result columns");
  if (join_agg)
    for (int c = 0; c < join_agg->num_cols; c++)
      if (type_width(join_agg->col_types[c]) <= 0)
        return fail(GZ_NOT_IMPLEMENTED, "join aggregates read fixed-width columns only");
  auto* s = new gz_server();
  s->device = device;
  s->budget = hbm_budget > 0 ? hbm_budget : INT64_MAX;
  s->agg.copy(agg);
  s->join_agg.copy(join_agg);
  if (split) {
    s->split_types.assign(split->col_types, split->col_types + split->num_cols);
    s->split_keys.assign(split->key_cols, split->key_cols + split->num_keys);
    s->split_P = split->num_partitions;
    s->split_set = true;
  }
  if (probe) {
    s->probe_types.assign(probe->probe_types, probe->probe_types + probe->num_probe_cols);
    s->probe_sides.assign(probe->result_sides, probe->result_sides + probe->num_result_cols);
    s->probe_idx.assign(probe->result_indices, probe->result_indices + probe->num_result_cols);
    s->probe_key = probe->probe_key_col;
    s->join_type = probe->join_type;
    s->probe_set = true;
  }
  s->slots.resize(num_slots);
  for (int k = 0; k < num_slots; k++) {
    int rc = gz_context_create(device, hbm_budget, &s->slots[k].ctx);
    if (rc) {
      for (int j = 0; j < k; j++) gz_context_destroy(s->slots[j].ctx);
      delete s;
      return rc;
    }
  }
  for (int k = 0; k < num_slots; k++) s->slots[k].th = std::thread(worker, s, k);
  *out = s;
  return GZ_OK;
}

int gz_server_add_relation(gz_server* s, gz_relation* r, int32_t* id) {
  if (!s || !r || !id) return fail(GZ_INVALID, "null argument");
  // the relation is read from every slot's stream: its build must be complete
  GZ_TRY(relation_synchronize(r));
  s->relations.push_back(r);
  *id = (int32_t)s->relations.size() - 1;
  return GZ_OK;
}

int gz_server_run(gz_server* s, int32_t n, const gz_fragment* frags, int32_t keep_results,
                  gz_fragment_stats* stats) {
  if (!s || (n > 0 && (!frags || !stats))) return fail(GZ_INVALID, "null argument");
  for (int i = 1; i < n; i++)
    if (frags[i].arrival_s < frags[i - 1].arrival_s) return fail(GZ_INVALID, "fragments must be sorted by arrival");
  GZ_TRY(gz_server_release_results(s));
  std::unique_lock<std::mutex> lk(s->mu);
  s->frags = frags;
  s->stats = stats;
  s->keep = keep_results != 0;
  s->est.assign(n, 0);
  s->agg_out.assign(n, nullptr);
  s->split_out.assign(n, nullptr);
  for (int i = 0; i < n; i++) {
    s->est[i] = estimate(s, frags[i]);
    gz_fragment_stats& st = stats[i];
    memset(&st, 0, sizeof(st));
    st.arrival_s = frags[i].arrival_s;
    st.rows = rows_of(frags[i]);
    st.slot = -1;
    st.status = -1;
    if (s->est[i] > s->budget) {
      lk.unlock();
      return fail(GZ_OUT_OF_MEMORY, "fragment " + std::to_string(i) + " needs more HBM than the budget");
    }
  }
  s->in_use = 0;
  s->done = 0;
  s->first_failed = -1;
  s->first_msg.clear();
  s->t0 = Clock::now();
  int next = 0;  // next fragment to admit (first come, first served)
  while (s->done < n) {
    // admit the head while it has arrived, a slot is idle and its estimate fits
    while (next < n && frags[next].arrival_s <= since(s)) {
      int idle = -1;
      for (int k = 0; k < (int)s->slots.size(); k++)
        if (s->slots[k].frag < 0) {
          idle = k;
          break;
        }
      if (idle < 0 || s->in_use + s->est[next] > s->budget) break;
      s->in_use += s->est[next];
      stats[next].admitted_s = since(s);
      stats[next].slot = idle;
      s->slots[idle].frag = next;
      next++;
      s->cv_work.notify_all();
    }
    if (s->done >= n) break;
    // sleep until the next arrival or a completion
    if (next < n && frags[next].arrival_s > since(s)) {
      auto wake = s->t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(frags[next].arrival_s));
      s->cv_done.wait_until(lk, wake);
    } else {
      const int d = s->done;
      s->cv_done.wait(lk, [&] { return s->done > d; });
    }
  }
  int first_error = GZ_OK;
  for (int i = 0; i < n; i++)
    if (stats[i].status != GZ_OK && first_error == GZ_OK) first_error = stats[i].status;
  const std::string msg = s->first_failed < 0 ? std::string()
                                              : "fragment " + std::to_string(s->first_failed) + " (kind " +
                                                    std::to_string(frags[s->first_failed].kind) + ", " +
                                                    std::to_string(rows_of(frags[s->first_failed])) +
                                                    " rows) failed first: " + s->first_msg;
  lk.unlock();
  if (first_error != GZ_OK) return fail(first_error, msg.empty() ? "a fragment failed (see its status)" : msg);
  return GZ_OK;
}

int gz_server_fragment_agg(gz_server* s, int32_t i, gz_agg** out) {
  if (!s || !out) return fail(GZ_INVALID, "null argument");
  if (i < 0 || i >= (int)s->agg_out.size() || !s->agg_out[i]) return fail(GZ_INDEX_ERROR, "no aggregate result kept");
  *out = s->agg_out[i];
  return GZ_OK;
}

int gz_server_fragment_split(gz_server* s, int32_t i, gz_splitter** out) {
  if (!s || !out) return fail(GZ_INVALID, "null argument");
  if (i < 0 || i >= (int)s->split_out.size() || !s->split_out[i]) return fail(GZ_INDEX_ERROR, "no split result kept");
  *out = s->split_out[i];
  return GZ_OK;
}

int gz_server_release_results(gz_server* s) {
  if (!s) return fail(GZ_INVALID, "null server");
  for (auto*& a : s->agg_out)
    if (a) {
      gz_agg_release(a);
      a = nullptr;
    }
  for (auto*& p : s->split_out)
    if (p) {
      gz_split_release(p);
      p = nullptr;
    }
  return GZ_OK;
}

int gz_validity_bitmap(gz_context* ctx, const void* valid_bytes, int64_t n, void* bitmap) {
  if (!ctx || n < 0 || (n > 0 && (!valid_bytes || !bitmap))) return fail(GZ_INVALID, "gz_validity_bitmap: bad argument");
  GZ_HIP(hipSetDevice(ctx->device));
  return pack_validity(ctx, (const uint8_t*)valid_bytes, n, (uint8_t*)bitmap);
}

int gz_server_destroy(gz_server* s) {
  if (!s) return GZ_OK;
  gz_server_release_results(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto& sl : s->slots) sl.stop = true;
    s->cv_work.notify_all();
  }
  for (auto& sl : s->slots)
    if (sl.th.joinable()) sl.th.join();
  for (auto& sl : s->slots) {
    for (auto& b : sl.bits) ctx_free(sl.ctx, &b);
    gz_context_destroy(sl.ctx);
  }
  delete s;
  return GZ_OK;
}

}  // extern "C"
