// Staging of the per-batch entry points (gz_split_batch, gz_agg_process).
//
// A per-batch call from a JNI-style caller holds one small batch (4096 rows is
// the reference's default): launching the whole kernel sequence per batch costs
// tens of microseconds of host work for a fraction of a microsecond of device
// work.  Instead the call validates the batch, copies its buffers into a device
// arena owned by the handle with one gather-copy kernel on the context's stream
// (so the borrowed buffers are read in stream order, as the entry points
// promise), and returns.  The handle runs its batched path over the staged
// batches when the arena fills or when the results are needed.
#include <algorithm>
#include <cstring>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

namespace {

constexpr int kStageAlign = 256;
constexpr int kCopyJobs = 96;                   // buffers per copy launch
constexpr int64_t kArenaBatches = 256;          // the first arena holds this many batches like the first
constexpr int64_t kArenaMin = 4ll << 20;        // at least 4 MiB
constexpr int64_t kArenaMax = 128ll << 20;      // largest arena

struct CopyJob {
  const uint8_t* src;
  uint8_t* dst;
  int64_t n;
};

struct CopyList {
  int njobs;
  CopyJob job[kCopyJobs];
};

// blockIdx.y: the buffer; the x blocks stride over its bytes, 16 at a time when
// both ends are 16-byte aligned.
__global__ __launch_bounds__(256) void stage_copy_kernel(CopyList L) {
  const CopyJob J = L.job[blockIdx.y];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t done = 0;
  if ((((uintptr_t)J.src | (uintptr_t)J.dst) & 15) == 0) {
    const int64_t n16 = J.n >> 4;
    const uint4* s = (const uint4*)J.src;
    uint4* d = (uint4*)J.dst;
    for (int64_t i = t; i < n16; i += stride) d[i] = s[i];
    done = n16 << 4;
  }
  for (int64_t i = done + t; i < J.n; i += stride) J.dst[i] = J.src[i];
}

int64_t align_up(int64_t v) { return (v + kStageAlign - 1) / kStageAlign * kStageAlign; }

// Segment s of an upload blob: bytes [off, off + n) of the blob go to dst.
struct BlobSeg {
  uint8_t* dst;
  int64_t off, n;
};

// blockIdx.y: the segment (the segment table sits at blob + table); the x
// blocks stride over its 16-byte words: segment offsets and destinations are
// 16-byte aligned (kStageAlign and the pool's allocations), and sizes are
// rounded up to 16 within the destination's capacity and the blob.
__global__ __launch_bounds__(256) void blob_scatter_kernel(const uint8_t* blob, int64_t table, int nseg) {
  if ((int)blockIdx.y >= nseg) return;
  const BlobSeg S = ((const BlobSeg*)(blob + table))[blockIdx.y];
  const uint4* src = (const uint4*)(blob + S.off);
  uint4* dst = (uint4*)S.dst;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < S.n >> 4; i += stride) dst[i] = src[i];
}

}  // namespace

int UploadBatch::add(gz_context* ctx, DBuf* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GZ_OK;
  // destinations hold whole 16-byte words, so the copy may round up
  GZ_TRY(ctx_ensure(ctx, dst, (bytes + 15) / 16 * 16));
  const size_t off = (size_t)align_up((int64_t)host_.size());
  host_.resize(off + bytes);
  memcpy(host_.data() + off, src, bytes);
  items_.push_back(Item{(uint8_t*)dst->p, off, bytes});
  return GZ_OK;
}

int UploadBatch::flush(gz_context* ctx, DBuf* blob) {
  if (items_.empty()) return GZ_OK;
  // blob: [data | segment table]; the last segment's round-up stays inside the blob
  const size_t table = (size_t)align_up((int64_t)host_.size() + 16);
  host_.resize(table + sizeof(BlobSeg) * items_.size());
  BlobSeg* segs = (BlobSeg*)(host_.data() + table);
  for (size_t k = 0; k < items_.size(); k++)
    segs[k] = BlobSeg{items_[k].dst, (int64_t)items_[k].off, (int64_t)((items_[k].bytes + 15) / 16 * 16)};
  GZ_TRY(upload(ctx, blob, host_.data(), host_.size()));
  size_t most = 0;
  for (auto& it : items_) most = std::max(most, it.bytes);
  const unsigned gx = (unsigned)std::min<size_t>(64, (most + 4095) / 4096);  // 4 KiB per workgroup and step
  hipLaunchKernelGGL(blob_scatter_kernel, dim3(gx, (unsigned)items_.size()), dim3(256), 0, ctx->stream,
                     (const uint8_t*)blob->p, (int64_t)table, (int)items_.size());
  GZ_HIP(hipGetLastError());
  host_.clear();
  items_.clear();
  return GZ_OK;
}

int stage_check(int32_t ncols, const int32_t* types, const gz_batch& b) {
  if (b.num_rows < 0 || b.num_rows > INT32_MAX) return fail(GZ_INVALID, "bad num_rows");
  if (b.num_rows > 0 && (!b.buf_addrs || !b.buf_sizes)) return fail(GZ_INVALID, "buf_addrs can't be null");
  int need = 0;
  for (int c = 0; c < ncols; c++) need += type_num_bufs(types[c]);
  if (b.num_rows == 0) return GZ_OK;
  if (b.num_bufs < need) return fail(GZ_INVALID, "insufficient number of in_buf_addrs");
  const int64_t n = b.num_rows;
  int k = 0;
  for (int c = 0; c < ncols; c++) {
    const int t = types[c];
    if (t != GZ_TYPE_NA && b.buf_addrs[k] && b.buf_sizes[k] > 0 && b.buf_sizes[k] < (n + 7) / 8)
      return fail(GZ_INVALID, "validity buffer too small");
    k++;
    if (type_is_binary(t)) {
      if (!b.buf_addrs[k] || b.buf_sizes[k] < (n + 1) * 4) return fail(GZ_INVALID, "offsets buffer too small");
      k++;
    }
    const int64_t sz = b.buf_sizes[k];
    if (t == GZ_TYPE_BOOL ? sz < (n + 7) / 8 : (!type_is_binary(t) && t != GZ_TYPE_NA && sz < n * type_width(t)))
      return fail(GZ_INVALID, "values buffer too small");
    if (sz > 0 && !b.buf_addrs[k]) return fail(GZ_INVALID, "missing values buffer");
    k++;
  }
  return GZ_OK;
}

int64_t stage_bytes(const gz_batch& b) {
  int64_t s = 0;
  if (!b.buf_addrs || !b.buf_sizes) return 0;
  for (int k = 0; k < b.num_bufs; k++)
    if (b.buf_addrs[k] && b.buf_sizes[k] > 0) s += align_up(b.buf_sizes[k]);
  return s;
}

int stage_add(gz_context* ctx, BatchStage* s, const gz_batch& b, bool* staged) {
  *staged = false;
  const int64_t need = stage_bytes(b);
  if (s->cap == 0 || (s->empty() && need * 4 > s->cap && s->cap < kArenaMax)) {
    // sized for kArenaBatches batches like this one (an empty arena grows for a larger batch)
    const int64_t cap = std::min(kArenaMax, std::max(kArenaMin, align_up(need * kArenaBatches)));
    if (need * 4 > cap) return GZ_OK;  // a large batch: the caller runs it directly
    ctx_free(ctx, &s->arena);  // stream-ordered: a later owner of the block runs after the queued copies
    s->cap = 0;
    if (ctx_alloc(ctx, (size_t)cap, &s->arena) != GZ_OK) {
      (void)hipGetLastError();
      return GZ_OK;  // no room for an arena (HBM budget): run directly
    }
    s->cap = cap;
  }
  if (need * 4 > s->cap || s->used + need > s->cap) return GZ_OK;
  const size_t first = s->addrs.size();
  CopyList L;
  L.njobs = 0;
  int64_t maxn = 0;
  auto launch = [&]() {
    if (L.njobs == 0) return;
    const unsigned gx = (unsigned)std::max<int64_t>(1, std::min<int64_t>((maxn + 4095) / 4096, 256));
    hipLaunchKernelGGL(stage_copy_kernel, dim3(gx, (unsigned)L.njobs), dim3(256), 0, ctx->stream, L);
    L.njobs = 0;
    maxn = 0;
  };
  for (int k = 0; k < b.num_bufs; k++) {
    // a batch without rows keeps its (small) buffers too: the batched path
    // checks their sizes as for any batch
    const int64_t sz = b.buf_sizes ? b.buf_sizes[k] : 0;
    const int64_t src = b.buf_addrs ? b.buf_addrs[k] : 0;
    if (!src || sz <= 0) {
      s->addrs.push_back(0);
      s->sizes.push_back(0);
      continue;
    }
    uint8_t* dst = (uint8_t*)s->arena.p + s->used;
    s->addrs.push_back((int64_t)dst);
    s->sizes.push_back(sz);
    s->used += align_up(sz);
    L.job[L.njobs++] = CopyJob{(const uint8_t*)src, dst, sz};
    maxn = std::max(maxn, sz);
    if (L.njobs == kCopyJobs) launch();
  }
  launch();
  GZ_HIP(hipGetLastError());
  s->first.push_back(first);
  s->rows.push_back(b.num_rows);
  s->nbufs.push_back(b.num_bufs);
  *staged = true;
  return GZ_OK;
}

void stage_view(const BatchStage& s, std::vector<gz_batch>* out) {
  out->resize(s.rows.size());
  for (size_t i = 0; i < s.rows.size(); i++)
    (*out)[i] = gz_batch{s.rows[i], s.addrs.data() + s.first[i], s.sizes.data() + s.first[i], s.nbufs[i]};
}

void stage_clear(BatchStage* s) {
  s->used = 0;
  s->addrs.clear();
  s->sizes.clear();
  s->first.clear();
  s->rows.clear();
  s->nbufs.clear();
}

void stage_release(gz_context* ctx, BatchStage* s) {
  stage_clear(s);
  ctx_free(ctx, &s->arena);
  s->cap = 0;
}

}  // namespace gz
