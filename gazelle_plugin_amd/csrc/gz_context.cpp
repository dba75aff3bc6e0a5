// Context, error reporting, the HBM pool and kernel timing of the C ABI.
#include <algorithm>
#include <cstring>

#include <iterator>

#include "gz_internal.h"

namespace gz {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ---- DevicePool -----------------------------------------------------------------
size_t DevicePool::round(size_t bytes) {
  if (bytes < 256) return 256;
  if (bytes <= (64ull << 20)) {
    size_t r = 256;
    while (r < bytes) r <<= 1;
    return r;
  }
  const size_t g = 2ull << 20;
  return (bytes + g - 1) / g * g;
}

DevicePool::~DevicePool() {
  hipSetDevice(device_);
  for (auto& kv : free_) hipFree(kv.second);
  for (auto& kv : live_) hipFree(kv.first);
}

int DevicePool::trim() {
  for (auto& kv : free_) hipFree(kv.second);
  free_.clear();
  cached_ = 0;
  return GZ_OK;
}

int DevicePool::alloc(size_t bytes, void** out) {
  std::lock_guard<std::mutex> g(mu_);
  size_t r = round(bytes);
  if (budget_ > 0 && in_use_ + (int64_t)r > budget_)
    return fail(GZ_OUT_OF_MEMORY, "HBM budget exceeded: " + std::to_string(in_use_) + " + " +
                                      std::to_string(r) + " > " + std::to_string(budget_));
  // best fit among cached blocks of at most 1.25x the request (sizes vary
  // between calls, so exact-size reuse alone would keep growing the cache)
  auto it = free_.lower_bound(r);
  void* p = nullptr;
  if (it != free_.end() && it->first <= r + r / 4) {
    p = it->second;
    r = it->first;
    free_.erase(it);
    cached_ -= r;
  } else {
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
      // release cached blocks and retry once.  The failed call left its error
      // as the thread's last error, which a later hipGetLastError (after a
      // kernel launch) would report: clear it, the retry decides.
      (void)hipGetLastError();
      for (auto& kv : free_) hipFree(kv.second);
      free_.clear();
      cached_ = 0;
      e = hipMalloc(&p, r);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(GZ_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
      }
    }
  }
  live_[p] = r;
  in_use_ += r;
  if (in_use_ > peak_) peak_ = in_use_;
  *out = p;
  return GZ_OK;
}

void DevicePool::free(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  size_t r = it->second;
  live_.erase(it);
  in_use_ -= r;
  free_.emplace(r, p);
  cached_ += r;
  if (cached_ > kCacheLimit) {
    // the largest blocks go back to HIP first, down to half the limit;
    // hipFree waits for the device, so work still reading a block has finished
    while (cached_ > kCacheLimit / 2 && !free_.empty()) {
      auto last = std::prev(free_.end());
      hipFree(last->second);
      cached_ -= (int64_t)last->first;
      free_.erase(last);
    }
  }
}

int ctx_alloc(gz_context* ctx, size_t bytes, DBuf* out) {
  void* p = nullptr;
  GZ_TRY(ctx->pool->alloc(bytes ? bytes : 1, &p));
  out->p = p;
  out->bytes = bytes;
  return GZ_OK;
}

void ctx_free(gz_context* ctx, DBuf* b) {
  if (b->p) ctx->pool->free(b->p);
  b->p = nullptr;
  b->bytes = 0;
}

int ctx_ensure(gz_context* ctx, DBuf* b, size_t bytes) {
  if (b->p && b->bytes >= bytes) return GZ_OK;
  if (b->p) {
    // the old block may still be read by queued work: free it only after the stream drains
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    ctx_free(ctx, b);
  }
  size_t want = bytes < 4096 ? 4096 : bytes + bytes / 4;
  return ctx_alloc(ctx, want, b);
}

int upload(gz_context* ctx, DBuf* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GZ_OK;
  GZ_TRY(ctx_ensure(ctx, dst, bytes));
  std::lock_guard<std::mutex> g(ctx->staging_mu);
  auto& s = ctx->staging[ctx->staging_next];
  ctx->staging_next = (ctx->staging_next + 1) % 8;
  if (s.pending) {
    GZ_HIP(hipEventSynchronize(s.done));
    s.pending = false;
  }
  if (s.cap < bytes || s.cap < ctx->staging_hwm) {
    if (s.host) hipHostFree(s.host);
    s.host = nullptr;
    // pinned allocations cost ~100 us each: size slots generously and to the
    // largest size any slot has needed, so the ring stops reallocating
    size_t want = std::max<size_t>({bytes + bytes / 2, (size_t)1 << 20, ctx->staging_hwm});
    ctx->staging_hwm = want;
    GZ_HIP(hipHostMalloc(&s.host, want, hipHostMallocDefault));
    s.cap = want;
  }
  if (!s.done) GZ_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  memcpy(s.host, src, bytes);
  GZ_HIP(hipMemcpyAsync(dst->p, s.host, bytes, hipMemcpyHostToDevice, ctx->stream));
  GZ_HIP(hipEventRecord(s.done, ctx->stream));
  s.pending = true;
  return GZ_OK;
}

// ---- timing ------------------------------------------------------------------------
static hipEvent_t take_event(gz_context* ctx) {
  {
    std::lock_guard<std::mutex> g(ctx->stat_mu);
    if (!ctx->event_cache.empty()) {
      hipEvent_t e = ctx->event_cache.back();
      ctx->event_cache.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

KernelTimer::KernelTimer(gz_context* ctx, const char* name) : ctx_(ctx), name_(name) {
  if (!ctx_->profiling) return;
  a_ = take_event(ctx_);
  b_ = take_event(ctx_);
  hipEventRecord(a_, ctx_->stream);
}

KernelTimer::~KernelTimer() {
  if (!a_) return;
  hipEventRecord(b_, ctx_->stream);
  size_t n;
  {
    std::lock_guard<std::mutex> g(ctx_->stat_mu);
    ctx_->pending.push_back({name_, a_, b_});
    n = ctx_->pending.size();
  }
  if (n > 4096) ctx_resolve_timings(ctx_, true);
}

void ctx_resolve_timings(gz_context* ctx, bool wait) {
  std::lock_guard<std::mutex> g(ctx->stat_mu);
  std::vector<gz_context::Pending> keep;
  for (auto& p : ctx->pending) {
    if (!wait && hipEventQuery(p.b) != hipSuccess) {
      keep.push_back(p);
      continue;
    }
    hipEventSynchronize(p.b);
    float ms = 0;
    hipEventElapsedTime(&ms, p.a, p.b);
    auto& s = ctx->stats[p.name];
    s.ms += ms;
    s.launches += 1;
    ctx->event_cache.push_back(p.a);
    ctx->event_cache.push_back(p.b);
  }
  ctx->pending.swap(keep);
}

}  // namespace gz

using namespace gz;

extern "C" {

const char* gz_last_error(void) { return g_last_error.c_str(); }

int gz_context_create(int device, int64_t hbm_budget_bytes, gz_context** out) {
  if (!out) return fail(GZ_INVALID, "out is null");
  int n = 0;
  GZ_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(GZ_INVALID, "device " + std::to_string(device) + " out of range");
  GZ_HIP(hipSetDevice(device));
  auto* ctx = new gz_context();
  ctx->device = device;
  hipError_t e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return fail(GZ_HIP_ERROR, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  ctx->stream = ctx->own_stream;
  ctx->pool = new DevicePool(device, hbm_budget_bytes);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
  if (ctx->num_cus <= 0) ctx->num_cus = 256;
  *out = ctx;
  return GZ_OK;
}

int gz_context_destroy(gz_context* ctx) {
  if (!ctx) return GZ_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  ctx_resolve_timings(ctx, true);
  for (auto e : ctx->event_cache) hipEventDestroy(e);
  for (auto& st : ctx->staging) {
    if (st.host) hipHostFree(st.host);
    if (st.done) hipEventDestroy(st.done);
  }
  delete ctx->pool;
  if (ctx->own_stream) hipStreamDestroy(ctx->own_stream);
  delete ctx;
  return GZ_OK;
}

int gz_context_set_stream(gz_context* ctx, void* s) {
  if (!ctx) return fail(GZ_INVALID, "null context");
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  return GZ_OK;
}

void* gz_context_stream(gz_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int gz_context_synchronize(gz_context* ctx) {
  if (!ctx) return fail(GZ_INVALID, "null context");
  GZ_HIP(hipStreamSynchronize(ctx->stream));
  return GZ_OK;
}

int64_t gz_context_bytes_in_use(gz_context* ctx) { return ctx ? ctx->pool->in_use() : 0; }
int64_t gz_context_peak_bytes(gz_context* ctx) { return ctx ? ctx->pool->peak() : 0; }

int gz_context_set_profiling(gz_context* ctx, int enabled) {
  if (!ctx) return fail(GZ_INVALID, "null context");
  ctx->profiling = enabled;
  return GZ_OK;
}

int gz_context_kernel_time(gz_context* ctx, const char* name, double* total_ms,
                           int64_t* launches) {
  if (!ctx || !name) return fail(GZ_INVALID, "null argument");
  ctx_resolve_timings(ctx, true);
  std::lock_guard<std::mutex> g(ctx->stat_mu);
  auto it = ctx->stats.find(name);
  if (total_ms) *total_ms = it == ctx->stats.end() ? 0.0 : it->second.ms;
  if (launches) *launches = it == ctx->stats.end() ? 0 : it->second.launches;
  return GZ_OK;
}

int gz_context_reset_kernel_times(gz_context* ctx) {
  if (!ctx) return fail(GZ_INVALID, "null context");
  ctx_resolve_timings(ctx, true);
  std::lock_guard<std::mutex> g(ctx->stat_mu);
  ctx->stats.clear();
  return GZ_OK;
}

}  // extern "C"
