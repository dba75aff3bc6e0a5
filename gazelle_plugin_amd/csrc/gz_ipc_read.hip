// Shuffle read for MI355X (gfx950): Arrow IPC streams of a shuffle partition
// back into device batches.
//
// The reading side of the shuffle: what the JVM's
// SchemaAwareArrowCompressedStreamReader + ShuffleDecompressionJniWrapper
// (jni_wrapper.cc:1236-1322, DecompressBuffers jni_common.h:471-560) produce,
// as device batches.  The host parses the message metadata; the buffers of an
// uncompressed record batch are addressed inside the device copy of the stream
// (zero copy: IPC bodies keep buffers 8-byte aligned).  A record batch with
// LZ4_FRAME body compression (the Spark "lz4" codec, jni_common.h:453-468)
// stores each non-empty buffer as an int64 uncompressed length then one LZ4
// frame; those buffers are decoded on the device into a caller-provided arena.
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gz_device.h"
#include "gz_internal.h"

namespace gz {

// Bounds-checked reads of a flatbuffer (the stream is untrusted input).
struct FlatView {
  const uint8_t* b;
  size_t n;
  bool ok = true;
  template <typename T>
  T get(size_t pos) {
    if (pos + sizeof(T) > n) {
      ok = false;
      return T(0);
    }
    T v;
    memcpy(&v, b + pos, sizeof(T));
    return v;
  }
  // field `id` of the table at `tpos`: its position, or 0 when absent
  size_t field(size_t tpos, int id) {
    const int32_t soff = get<int32_t>(tpos);
    const int64_t vt = (int64_t)tpos - soff;
    if (vt < 0 || (size_t)vt + 4 > n) {
      ok = false;
      return 0;
    }
    const uint16_t vsize = get<uint16_t>((size_t)vt);
    if (4 + 2 * (size_t)id + 2 > vsize) return 0;
    const uint16_t fo = get<uint16_t>((size_t)vt + 4 + 2 * id);
    return fo ? tpos + fo : 0;
  }
  size_t deref(size_t fpos) { return fpos ? fpos + get<uint32_t>(fpos) : 0; }
};

// Where one JNI-order buffer of a parsed batch lives.
enum BufWhere : int8_t { kNowhere = 0, kInStream = 1, kInArena = 2 };
struct ParsedBuf {
  int8_t where;
  int64_t off;   // stream offset (kInStream) or arena offset (kInArena)
  int64_t size;  // bytes the buffer holds (decoded size for kInArena)
  int64_t src;   // kInArena: stream offset of the LZ4 data (after the length prefix)
  int64_t src_len;
};
struct ParsedStream {
  std::vector<int64_t> rows;
  std::vector<ParsedBuf> bufs;  // rows.size() * nbufs
  int64_t arena = 0;            // arena bytes the compressed buffers decode into
};

constexpr int64_t kArenaAlign = 64;

int ipc_parse(const uint8_t* stream, int64_t len, int num_cols, const int32_t* col_types, ParsedStream* out) {
  int nbufs = 0;
  for (int c = 0; c < num_cols; c++) {
    if (!type_supported(col_types[c])) return fail(GZ_NOT_IMPLEMENTED, "column type " + std::to_string(col_types[c]));
    nbufs += type_num_bufs(col_types[c]);
  }
  int expect = 0;
  for (int c = 0; c < num_cols; c++) expect += col_types[c] == GZ_TYPE_NA ? 0 : type_is_binary(col_types[c]) ? 3 : 2;
  int64_t pos = 0;
  while (pos + 4 <= len) {
    int32_t w0;
    memcpy(&w0, stream + pos, 4);
    int64_t mlen, mstart;
    if (w0 == -1) {  // continuation token, then the metadata length
      if (pos + 8 > len) return fail(GZ_IO_ERROR, "truncated IPC message header");
      int32_t l;
      memcpy(&l, stream + pos + 4, 4);
      mlen = l;
      mstart = pos + 8;
    } else {  // the pre-0.15 framing: the length alone
      mlen = w0;
      mstart = pos + 4;
    }
    if (mlen == 0) break;  // end of stream
    if (mlen < 0 || mstart + mlen > len) return fail(GZ_IO_ERROR, "IPC message runs past the stream");
    FlatView fv{stream + mstart, (size_t)mlen};
    const size_t root = fv.get<uint32_t>(0);
    const size_t htype_f = fv.field(root, 1), header_f = fv.field(root, 2), blen_f = fv.field(root, 3);
    const uint8_t htype = htype_f ? fv.get<uint8_t>(htype_f) : 0;
    const int64_t body_len = blen_f ? fv.get<int64_t>(blen_f) : 0;
    const int64_t body = mstart + mlen;
    if (!fv.ok || body_len < 0 || body + body_len > len) return fail(GZ_IO_ERROR, "malformed IPC message");
    if (htype == 2) return fail(GZ_NOT_IMPLEMENTED, "dictionary batches");
    if (htype == 3) {  // RecordBatch
      const size_t rb = fv.deref(header_f);
      const size_t len_f = fv.field(rb, 0), nodes_f = fv.field(rb, 1), bufs_f = fv.field(rb, 2);
      const size_t comp = fv.deref(fv.field(rb, 3));
      bool lz4 = false;
      if (comp) {  // BodyCompression { codec: byte = LZ4_FRAME (0), method: byte = BUFFER (0) }
        const size_t codec_f = fv.field(comp, 0), method_f = fv.field(comp, 1);
        const uint8_t codec = codec_f ? fv.get<uint8_t>(codec_f) : 0;
        const uint8_t method = method_f ? fv.get<uint8_t>(method_f) : 0;
        if (codec != 0) return fail(GZ_NOT_IMPLEMENTED, "body compression codec " + std::to_string(codec) + " (LZ4_FRAME only)");
        if (method != 0) return fail(GZ_NOT_IMPLEMENTED, "body compression method other than BUFFER");
        lz4 = true;
      }
      const int64_t rows = len_f ? fv.get<int64_t>(len_f) : 0;
      const size_t nodes = fv.deref(nodes_f), bufs = fv.deref(bufs_f);
      const uint32_t nn = nodes ? fv.get<uint32_t>(nodes) : 0, nbuf = bufs ? fv.get<uint32_t>(bufs) : 0;
      if (!fv.ok || rows < 0) return fail(GZ_IO_ERROR, "malformed record batch");
      if ((int)nn != num_cols) return fail(GZ_INVALID, "record batch field count does not match the types");
      if ((int)nbuf != expect) return fail(GZ_INVALID, "record batch buffer count does not match the types");
      out->rows.push_back(rows);
      uint32_t bi = 0;
      for (int c = 0; c < num_cols; c++) {
        const int t = col_types[c];
        if (t == GZ_TYPE_NA) {
          out->bufs.push_back({kNowhere, 0, 0, 0, 0});
          out->bufs.push_back({kNowhere, 0, 0, 0, 0});
          continue;
        }
        const int m = type_is_binary(t) ? 3 : 2;
        for (int j = 0; j < m; j++, bi++) {
          // vector elements start 4 bytes after its length; each Buffer is 16 bytes
          const int64_t off = fv.get<int64_t>(bufs + 4 + 16 * bi), blen = fv.get<int64_t>(bufs + 4 + 16 * bi + 8);
          if (!fv.ok || off < 0 || blen < 0 || off > body_len || blen > body_len - off)
            return fail(GZ_IO_ERROR, "buffer outside the body");
          if (blen == 0) {
            out->bufs.push_back({kNowhere, 0, 0, 0, 0});
          } else if (!lz4) {
            out->bufs.push_back({kInStream, body + off, blen, 0, 0});
          } else {
            if (blen < 8) return fail(GZ_INVALID, "Likely corrupted message, compressed buffers are larger than 8 bytes by construction");
            int64_t ulen;
            memcpy(&ulen, stream + body + off, 8);
            if (ulen == -1) {  // stored uncompressed after the prefix
              out->bufs.push_back({blen > 8 ? kInStream : kNowhere, body + off + 8, blen - 8, 0, 0});
            } else if (ulen < 0) {
              return fail(GZ_IO_ERROR, "negative uncompressed length");
            } else if (ulen == 0) {
              out->bufs.push_back({kNowhere, 0, 0, 0, 0});
            } else {
              out->bufs.push_back({kInArena, out->arena, ulen, body + off + 8, blen - 8});
              out->arena += (ulen + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
            }
          }
        }
      }
    }
    pos = body + body_len;
  }
  (void)nbufs;
  return GZ_OK;
}

// ---- LZ4 frame decoding ------------------------------------------------------
// The published LZ4 frame and block formats (the lz4 library behind
// arrow::Compression::LZ4_FRAME).  The host walks the frame and block headers
// (metadata, a few bytes per 64 KiB block); one wave decodes one buffer, its
// blocks in order, since linked blocks match into earlier output of the frame.
struct Lz4Block {
  const uint8_t* src;
  uint32_t len;
  uint32_t flags;  // bit 0: stored (uncompressed) block; bit 1: first block of a frame
};
struct Lz4Task {
  uint8_t* dst;
  int64_t cap;   // decoded bytes expected (the length prefix)
  int32_t b0, nb;
};

// Frame and block headers of the LZ4 frames in stream[src, src + n).
int lz4_walk(const uint8_t* stream, int64_t src, int64_t n, int64_t dev_base, std::vector<Lz4Block>* blocks) {
  const uint8_t* s = stream + src;
  int64_t ip = 0;
  bool any = false;
  while (ip < n) {
    if (n - ip < 4) return fail(GZ_IO_ERROR, "LZ4 frame truncated");
    uint32_t magic;
    memcpy(&magic, s + ip, 4);
    if ((magic & 0xFFFFFFF0u) == 0x184D2A50u) {  // skippable frame
      if (n - ip < 8) return fail(GZ_IO_ERROR, "LZ4 skippable frame truncated");
      uint32_t sk;
      memcpy(&sk, s + ip + 4, 4);
      if ((int64_t)sk > n - ip - 8) return fail(GZ_IO_ERROR, "LZ4 skippable frame truncated");
      ip += 8 + sk;
      continue;
    }
    if (magic != 0x184D2204u) return fail(GZ_IO_ERROR, "LZ4 decompress failed: unknown frame magic");
    if (n - ip < 7) return fail(GZ_IO_ERROR, "LZ4 frame header truncated");
    const uint8_t flg = s[ip + 4];
    if ((flg >> 6) != 1 || (flg & 0x02)) return fail(GZ_IO_ERROR, "LZ4 frame header: bad version or reserved bit");
    if (flg & 1) return fail(GZ_NOT_IMPLEMENTED, "LZ4 frames with a dictionary id");
    const int block_checksum = (flg >> 4) & 1, content_size = (flg >> 3) & 1, content_checksum = (flg >> 2) & 1;
    const uint8_t bd = s[ip + 5];
    const int bsid = (bd >> 4) & 7;
    if (bsid < 4 || (bd & 0x8F)) return fail(GZ_IO_ERROR, "LZ4 frame header: bad block size id");
    const int64_t max_block = 1ll << (8 + 2 * bsid);  // 64 KiB, 256 KiB, 1 MiB, 4 MiB
    ip += 6 + (content_size ? 8 : 0) + 1;
    if (ip > n) return fail(GZ_IO_ERROR, "LZ4 frame header truncated");
    bool first = true;
    for (;;) {
      if (n - ip < 4) return fail(GZ_IO_ERROR, "LZ4 block header truncated");
      uint32_t bs;
      memcpy(&bs, s + ip, 4);
      ip += 4;
      if (bs == 0) break;  // end mark
      const int64_t blen = bs & 0x7FFFFFFFu;
      if (blen > n - ip || blen > max_block) return fail(GZ_IO_ERROR, "LZ4 block runs past its buffer");
      blocks->push_back({(const uint8_t*)(dev_base + src + ip), (uint32_t)blen,
                         (uint32_t)((bs >> 31) | (first ? 2u : 0u))});
      first = false;
      ip += blen + (block_checksum ? 4 : 0);
    }
    ip += content_checksum ? 4 : 0;
    if (ip > n) return fail(GZ_IO_ERROR, "LZ4 frame truncated");
    any = true;
  }
  if (!any) return fail(GZ_IO_ERROR, "LZ4 buffer without a frame");
  return GZ_OK;
}


// dst[k] = src[k] for k < n by one wave (the ranges never overlap where this is
// used).  Short copies go byte per lane.  Longer ones copy a byte head up to a
// 16-byte aligned dst, then 16 bytes per lane: the source is read as aligned
// dwords (each holds at least one byte of the range, so it stays inside the
// allocation holding the range) and realigned with v_alignbyte; then a byte tail.
__device__ __forceinline__ void wave_copy(uint8_t* dst, const uint8_t* src, int64_t n, int lane) {
  if (n <= kWave) {
    if (lane < n) gst(dst + lane, gld(src + lane));
    return;
  }
  const int64_t h = (int64_t)((16 - ((uint64_t)dst & 15)) & 15);
  if (lane < h) gst(dst + lane, gld(src + lane));
  dst += h;
  src += h;
  n -= h;
  const uint32_t m = (uint32_t)((uint64_t)src & 3);
  const uint32_t* sa = (const uint32_t*)(src - m);
  const int64_t units = n >> 4;
  auto unit = [&](int64_t u) {
    const uint32_t* w = sa + 4 * u;
    const uint32_t w0 = gld(w), w1 = gld(w + 1), w2 = gld(w + 2), w3 = gld(w + 3);
    const uint32_t w4 = m ? gld(w + 4) : w3;  // needed (and in bounds) only when misaligned
    u32x4 o;
    o.x = __builtin_amdgcn_alignbyte(w1, w0, m);
    o.y = __builtin_amdgcn_alignbyte(w2, w1, m);
    o.z = __builtin_amdgcn_alignbyte(w3, w2, m);
    o.w = __builtin_amdgcn_alignbyte(w4, w3, m);
    return o;
  };
  int64_t u = lane;
  for (; u + kWave < units; u += 2 * kWave) {
    const u32x4 a = unit(u), b = unit(u + kWave);
    gst((u32x4*)(dst + 16 * u), a);
    gst((u32x4*)(dst + 16 * (u + kWave)), b);
  }
  if (u < units) gst((u32x4*)(dst + 16 * u), unit(u));
  const int64_t t = units << 4;
  if (lane < n - t) gst(dst + t + lane, gld(src + t + lane));
}

// dst[k] = from[k mod period] for k < n: an overlapping LZ4 match, from + period
// == dst, every byte read precedes dst.  Periods up to a wave: the pattern sits
// one byte per lane and each lane assembles dwords from it with lane shuffles.
// Longer periods: copies of whole periods, each from the one before.
__device__ __forceinline__ void wave_repeat(uint8_t* dst, const uint8_t* from, uint32_t period, int64_t n, int lane) {
  if (period > (uint32_t)kWave) {
    for (int64_t c = 0; c < n; c += period) {
      wave_copy(dst + c, from + c, n - c < (int64_t)period ? n - c : (int64_t)period, lane);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
    return;
  }
  const int pb = lane < (int)period ? gld(from + lane) : 0;
  auto pat = [&](uint32_t k) { return (uint32_t)__shfl(pb, (int)(k % period), kWave) & 0xFFu; };
  const int64_t h = (int64_t)((4 - ((uint64_t)dst & 3)) & 3) < n ? (int64_t)((4 - ((uint64_t)dst & 3)) & 3) : n;
  {  // byte head up to a 4-byte aligned dst (every lane takes part in the shuffle)
    const uint32_t v = pat((uint32_t)lane);
    if (lane < h) gst(dst + lane, (uint8_t)v);
  }
  const int64_t words = (n - h) >> 2;
  uint32_t* dw = (uint32_t*)(dst + h);
  // lane's first byte index (relative to dst) and its residue mod period
  uint32_t r = (uint32_t)((h + 4 * lane) % period);
  const uint32_t step = (uint32_t)((4 * kWave) % period);
  const int64_t iters = (words + kWave - 1) / kWave;
  for (int64_t i = 0; i < iters; i++) {
    const int64_t wi = i * kWave + lane;
    uint32_t x = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      uint32_t rj = r + j;
      rj = rj >= period ? rj - period : rj;
      rj = rj >= period ? rj - period : rj;
      rj = rj >= period ? rj - period : rj;
      rj = rj >= period ? rj - period : rj;
      x |= ((uint32_t)__shfl(pb, (int)rj, kWave) & 0xFFu) << (8 * j);
    }
    if (wi < words) gst(dw + wi, x);
    r += step;
    r = r >= period ? r - period : r;
  }
  const int64_t t = h + (words << 2);
  const uint32_t v = pat((uint32_t)(t + lane));
  if (lane < n - t) gst(dst + t + lane, (uint8_t)v);
}

// Per-lane copies of short runs: four loads in flight before their stores.
__device__ __forceinline__ void lane_copy(uint8_t* d, const uint8_t* s, uint32_t n) {
  uint32_t k = 0;
  for (; k + 4 <= n; k += 4) {
    const uint8_t a = gld(s + k), b = gld(s + k + 1), c = gld(s + k + 2), e = gld(s + k + 3);
    gst(d + k, a);
    gst(d + k + 1, b);
    gst(d + k + 2, c);
    gst(d + k + 3, e);
  }
  for (; k < n; k++) gst(d + k, gld(s + k));
}
// d[k] = from[k mod period] (from + period == d): reads only the pattern.
__device__ __forceinline__ void lane_repeat(uint8_t* d, const uint8_t* from, uint32_t period, uint32_t n) {
  uint32_t r = 0;
  for (uint32_t k = 0; k < n; k++) {
    gst(d + k, gld(from + r));
    r = r + 1 == period ? 0 : r + 1;
  }
}

__device__ __forceinline__ int64_t wave_min_i64(int64_t v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const int64_t x = __shfl_xor((long long)v, o, kWave);
    v = x < v ? x : v;
  }
  return v;
}

// One LZ4 block into dst[op, cap): sequences of (token, literals, offset,
// match) until the block ends after a literal run.  Matches reach back no
// further than fbase (the frame's first output byte).  Returns the new op, or
// -1 on malformed input.
//
// The wave takes the block 64 sequences at a time.  (1) It parses them
// serially (wave-uniform, input bytes through the scalar cache); lane i keeps
// sequence i.  (2) Literal runs only read the input, so all 64 are copied at
// once.  (3) Matches read earlier output: in rounds, every match whose source
// bytes lie below the earliest output still pending is copied; the earliest
// pending match is always among them, and matches reaching far back (the usual
// case) all go in the first round.  Short runs are copied by their lane, long
// ones by the whole wave.
// v_writelane_b32: value (wave-uniform) into lane l of old, the other lanes kept.
__device__ int gz_writelane(int value, int l, int old) __asm("llvm.amdgcn.writelane.i32");

__device__ __forceinline__ int64_t readlane64(int64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

constexpr uint32_t kLz4Cand = 512;  // input positions decoded as candidate sequence starts per window
constexpr uint32_t kLz4Bad = 0xFFFFFFFFu;

// LDS of one wave for the parse: the window's input bytes (plus a lookahead)
// and, for every position of the window, the sequence that would start there.
struct Lz4Lds {
  uint8_t in[kLz4Cand + 64];
  u32x4 seq[kLz4Cand];  // next start (kLz4Bad: malformed), literal length, match length, header | offset << 16
};

// Sequence header at block position p, read from the staged window where it
// holds the bytes (from HBM past it): the position of the next sequence, the
// literal run (length, header bytes before it), the match (length, offset).
// Any position may be tried; only the chain from the block's start is used.
__device__ __forceinline__ u32x4 lz4_candidate(const uint8_t* blk, uint32_t blen, const Lz4Lds& L, uint32_t wbase,
                                               uint32_t p) {
  auto byte = [&](uint32_t x) -> uint32_t {
    const uint32_t o = x - wbase;
    return o < kLz4Cand + 64 ? (uint32_t)L.in[o] : (uint32_t)gld(blk + x);
  };
  const u32x4 bad{kLz4Bad, 0u, 0u, 0u};
  if (p >= blen) return bad;
  const uint32_t t = byte(p);
  uint32_t lit = t >> 4, q = p + 1;
  if (lit == 15) {
    uint32_t b;
    do {
      if (q >= blen) return bad;
      b = byte(q++);
      lit += b;
    } while (b == 255);
  }
  if (lit > blen - q) return bad;
  const uint32_t hdr = q - p;
  uint32_t e = q + lit;
  if (e == blen) return u32x4{blen, lit, 0u, hdr};  // the last sequence: literals only
  if (blen - e < 2) return bad;
  const uint32_t off = byte(e) | (byte(e + 1) << 8);
  e += 2;
  uint32_t ml = t & 15;
  if (ml == 15) {
    uint32_t b;
    do {
      if (e >= blen) return bad;
      b = byte(e++);
      ml += b;
    } while (b == 255 && ml <= 0x7FFFFFFFu);
    if (ml > 0x7FFFFFFFu) return bad;
  }
  if (hdr > 0xFFFFu) return bad;
  return u32x4{e, lit, ml + 4, hdr | (off << 16)};
}

__device__ int64_t lz4_block(const uint8_t* blk, uint32_t blen, uint8_t* dst, int64_t op, int64_t fbase,
                             int64_t cap, int lane, Lz4Lds& L) {
  constexpr uint32_t kLong = 32;
  // The sequence chain is serial, but a sequence's header can be decoded from
  // any position: every lane decodes the would-be sequences at its positions
  // of a 512-byte window (staged in LDS), then the wave walks the real chain
  // through them, one LDS read per sequence, placing each sequence's fields in
  // lane n with v_writelane.
  uint32_t wbase = kLz4Bad;
  auto stage = [&](uint32_t base) {
    wbase = base;
    for (uint32_t k = lane; k < kLz4Cand + 64; k += kWave)
      L.in[k] = base + k < blen ? gld(blk + base + k) : (uint8_t)0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (uint32_t k = lane; k < kLz4Cand; k += kWave) L.seq[k] = lz4_candidate(blk, blen, L, base, base + k);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  };
  uint32_t ip = 0;  // input offset within the block (blocks are at most 4 MiB)
  while (ip < blen) {
    int n = 0;
    uint32_t my_lsrc = 0, my_lit_len = 0, my_off = 0, my_mlen = 0, my_dlo = 0, my_dhi = 0;
    const int64_t op0 = op;
    uint32_t room = cap - op > 0xFFFFFFFFll ? 0xFFFFFFFFu : (uint32_t)(cap - op);
    uint32_t back = op - fbase > 65536 ? 65536u : (uint32_t)(op - fbase);
    uint32_t rel = 0;  // output produced by this window's sequences
    while (n < kWave && ip < blen) {
      if (wbase == kLz4Bad || ip - wbase >= kLz4Cand) stage(ip);
      const u32x4 f = L.seq[ip - wbase];
      if (f.x == kLz4Bad) return -1;
      const uint32_t lit = f.y, mlen = f.z, lsrc = ip + (f.w & 0xFFFFu), off = f.w >> 16;
      if (lit > room) return -1;
      room -= lit;
      rel += lit;
      back = back + lit > 65536u ? 65536u : back + lit;
      if (mlen && (off == 0 || off > back || mlen > room)) return -1;
      const int64_t mdst = op0 + rel;
      my_lsrc = (uint32_t)gz_writelane((int)lsrc, n, (int)my_lsrc);
      my_lit_len = (uint32_t)gz_writelane((int)lit, n, (int)my_lit_len);
      my_off = (uint32_t)gz_writelane((int)off, n, (int)my_off);
      my_mlen = (uint32_t)gz_writelane((int)mlen, n, (int)my_mlen);
      my_dlo = (uint32_t)gz_writelane((int)(uint32_t)mdst, n, (int)my_dlo);
      my_dhi = (uint32_t)gz_writelane((int)(uint32_t)((uint64_t)mdst >> 32), n, (int)my_dhi);
      room -= mlen;
      rel += mlen;
      back = back + mlen > 65536u ? 65536u : back + mlen;
      ip = f.x;
      n++;
    }
    op = op0 + rel;
    const uint8_t* my_lit = blk + my_lsrc;
    const int64_t my_mdst = (int64_t)(((uint64_t)my_dhi << 32) | my_dlo);
    const bool has = lane < n;
    // (2) literals
    if (has && my_lit_len <= kLong) lane_copy(dst + my_mdst - my_lit_len, my_lit, my_lit_len);
    for (uint64_t lg = __ballot(has && my_lit_len > kLong); lg; lg &= lg - 1) {
      const int l = __builtin_ctzll(lg);
      const uint8_t* s = (const uint8_t*)readlane64((int64_t)my_lit, l);
      const uint32_t len = (uint32_t)__builtin_amdgcn_readlane((int)my_lit_len, l);
      const int64_t d = readlane64(my_mdst, l) - len;
      wave_copy(dst + d, s, len, lane);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // (3) matches
    bool pend = has && my_mlen > 0;
    const int64_t msrc = my_mdst - my_off;
    const int64_t need = msrc + my_mlen < my_mdst ? msrc + my_mlen : my_mdst;  // bytes read below my_mdst
    while (__ballot(pend)) {
      const int64_t m = wave_min_i64(pend ? my_mdst : INT64_MAX);
      const bool ready = pend && need <= m;
      if (ready && my_mlen <= kLong) {
        if (my_off >= my_mlen)
          lane_copy(dst + my_mdst, dst + msrc, my_mlen);
        else
          lane_repeat(dst + my_mdst, dst + msrc, my_off, my_mlen);
      }
      for (uint64_t lg = __ballot(ready && my_mlen > kLong); lg; lg &= lg - 1) {
        const int l = __builtin_ctzll(lg);
        const int64_t d = readlane64(my_mdst, l);
        const uint32_t off = (uint32_t)__builtin_amdgcn_readlane((int)my_off, l),
                       len = (uint32_t)__builtin_amdgcn_readlane((int)my_mlen, l);
        if (off >= len)
          wave_copy(dst + d, dst + d - off, len, lane);
        else
          wave_repeat(dst + d, dst + d - off, off, len, lane);
      }
      pend = pend && !ready;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
  }
  return op;
}

// One wave per buffer (4 per block); produced[t] = decoded bytes or -1.
__global__ __launch_bounds__(256) void lz4_decode_kernel(const Lz4Task* __restrict__ tasks, int ntasks,
                                                         const Lz4Block* __restrict__ blocks,
                                                         int64_t* __restrict__ produced) {
  __shared__ Lz4Lds s_lz4[4];
  const int t = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
  const int lane = threadIdx.x & 63;
  if (t >= ntasks) return;
  const Lz4Task tk = tasks[t];
  int64_t op = 0, fbase = 0;
  for (int b = 0; b < tk.nb; b++) {
    const Lz4Block bl = blocks[tk.b0 + b];
    if (bl.flags & 2) fbase = op;
    if (bl.flags & 1) {  // stored block
      if ((int64_t)bl.len > tk.cap - op) {
        op = -1;
        break;
      }
      wave_copy(tk.dst + op, bl.src, bl.len, lane);
      op += bl.len;
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    } else {
      op = lz4_block(bl.src, bl.len, tk.dst, op, fbase, tk.cap, lane, s_lz4[threadIdx.x >> 6]);
      if (op < 0) break;
    }
  }
  if (lane == 0) produced[t] = op;
}

}  // namespace gz

using namespace gz;

// The decode jobs of a stream: one task per LZ4_FRAME buffer (largest first: a
// wave decodes one buffer serially, so the tail is the longest buffer started
// last) and the frame's blocks.
struct Lz4Jobs {
  std::vector<Lz4Task> tasks;
  std::vector<Lz4Block> blocks;
};

static int lz4_jobs(gz_context* ctx, const uint8_t* stream, int64_t len, int64_t device_base, int32_t num_cols,
                    const int32_t* col_types, void* arena, int64_t arena_capacity, Lz4Jobs* J) {
  if (!ctx || !stream || len < 0 || !device_base || num_cols < 0 || (num_cols > 0 && !col_types))
    return fail(GZ_INVALID, "gz_ipc_decompress: null argument");
  ParsedStream ps;
  GZ_TRY(ipc_parse(stream, len, num_cols, col_types, &ps));
  if (ps.arena == 0) return GZ_OK;
  if (!arena || arena_capacity < ps.arena) return fail(GZ_INVALID, "decompression arena smaller than the stream needs");
  std::vector<Lz4Task> tasks;
  for (const ParsedBuf& pb : ps.bufs) {
    if (pb.where != kInArena) continue;
    const int32_t b0 = (int32_t)J->blocks.size();
    GZ_TRY(lz4_walk(stream, pb.src, pb.src_len, device_base, &J->blocks));
    tasks.push_back({(uint8_t*)arena + pb.off, pb.size, b0, (int32_t)J->blocks.size() - b0});
  }
  std::vector<int32_t> order(tasks.size());
  for (size_t i = 0; i < order.size(); i++) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return tasks[a].cap > tasks[b].cap; });
  J->tasks.resize(tasks.size());
  for (size_t i = 0; i < order.size(); i++) J->tasks[i] = tasks[order[i]];
  return GZ_OK;
}

// DecompressBuffer's statuses (jni_common.h:471-492) from the decoded lengths.
static int lz4_check(const Lz4Jobs& J, const std::vector<int64_t>& prod) {
  for (size_t i = 0; i < prod.size(); i++) {
    if (prod[i] < 0) return fail(GZ_IO_ERROR, "LZ4 decompress failed: corrupt block data");
    if (prod[i] != J.tasks[i].cap)
      return fail(GZ_INVALID, "Failed to fully decompress buffer, expected " + std::to_string(J.tasks[i].cap) +
                                  " bytes but decompressed " + std::to_string(prod[i]));
  }
  return GZ_OK;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---- allocation-free decode (gz_ipc_decompress_noalloc)
// Its task table, block table and decoded lengths live in this file-scope
// device array instead of memory allocated per call.  Each GPU holds its own
// copy until the process ends; in host code `scratch` names a host-side
// stand-in, so the device address is taken once per device with
// hipGetSymbolAddress (that device current) and only that address is used.
constexpr int64_t kNoAllocScratchBytes = 8ll << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static uint8_t* g_scratch_ptr[kMaxScratchDevices];

static int scratch_address(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxScratchDevices) return fail(GZ_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    uint8_t* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(GZ_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return GZ_OK;
}

extern "C" {

int gz_ipc_read_batches(const uint8_t* stream, int64_t len, int64_t device_base, int64_t arena_base, int32_t num_cols,
                        const int32_t* col_types, int32_t max_batches, int64_t* batch_rows, int64_t* buf_addrs,
                        int64_t* buf_sizes, int32_t* num_batches, int64_t* arena_bytes) {
  if (!stream || len < 0 || !num_batches || num_cols < 0 || (num_cols > 0 && !col_types))
    return fail(GZ_INVALID, "gz_ipc_read_batches: null argument");
  ParsedStream ps;
  GZ_TRY(ipc_parse(stream, len, num_cols, col_types, &ps));
  const int64_t nb = (int64_t)ps.rows.size();
  const int64_t nbufs = nb ? (int64_t)ps.bufs.size() / nb : 0;
  if (batch_rows && buf_addrs && buf_sizes) {
    for (int64_t b = 0; b < std::min<int64_t>(nb, max_batches); b++) {
      batch_rows[b] = ps.rows[b];
      for (int64_t k = 0; k < nbufs; k++) {
        const ParsedBuf& pb = ps.bufs[b * nbufs + k];
        buf_addrs[b * nbufs + k] = pb.where == kInStream ? device_base + pb.off
                                   : pb.where == kInArena ? arena_base + pb.off : 0;
        buf_sizes[b * nbufs + k] = pb.size;
      }
    }
  }
  *num_batches = (int32_t)nb;
  if (arena_bytes) *arena_bytes = ps.arena;
  return GZ_OK;
}

int gz_ipc_decompress(gz_context* ctx, const uint8_t* stream, int64_t len, int64_t device_base, int32_t num_cols,
                      const int32_t* col_types, void* arena, int64_t arena_capacity) {
  Lz4Jobs J;
  GZ_TRY(lz4_jobs(ctx, stream, len, device_base, num_cols, col_types, arena, arena_capacity, &J));
  if (J.tasks.empty()) return GZ_OK;
  GZ_HIP(hipSetDevice(ctx->device));
  DBuf d_tasks, d_blocks, d_prod;
  int rc = GZ_OK;
  if ((rc = upload(ctx, &d_tasks, J.tasks.data(), sizeof(Lz4Task) * J.tasks.size())) ||
      (rc = upload(ctx, &d_blocks, J.blocks.data(), sizeof(Lz4Block) * std::max<size_t>(J.blocks.size(), 1))) ||
      (rc = ctx_alloc(ctx, sizeof(int64_t) * J.tasks.size(), &d_prod))) {
    ctx_free(ctx, &d_tasks);
    ctx_free(ctx, &d_blocks);
    return rc;
  }
  const int nt = (int)J.tasks.size();
  {
    KernelTimer kt(ctx, "ipc_lz4_decode");
    hipLaunchKernelGGL(lz4_decode_kernel, dim3((nt + 3) / 4), dim3(256), 0, ctx->stream,
                       (const Lz4Task*)d_tasks.p, nt, (const Lz4Block*)d_blocks.p, (int64_t*)d_prod.p);
  }
  std::vector<int64_t> prod(J.tasks.size());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(prod.data(), d_prod.p, sizeof(int64_t) * prod.size(), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx_free(ctx, &d_tasks);
  ctx_free(ctx, &d_blocks);
  ctx_free(ctx, &d_prod);
  if (e != hipSuccess) return fail(GZ_HIP_ERROR, std::string("gz_ipc_decompress: ") + hipGetErrorString(e));
  return lz4_check(J, prod);
}

int gz_ipc_decompress_noalloc(gz_context* ctx, const uint8_t* stream, int64_t len, int64_t device_base,
                              int32_t num_cols, const int32_t* col_types, void* arena, int64_t arena_capacity) {
  Lz4Jobs J;
  GZ_TRY(lz4_jobs(ctx, stream, len, device_base, num_cols, col_types, arena, arena_capacity, &J));
  if (J.tasks.empty()) return GZ_OK;
  const size_t tb = align_up(sizeof(Lz4Task) * J.tasks.size(), 256);
  const size_t bb = align_up(sizeof(Lz4Block) * J.blocks.size(), 256);
  const size_t pb = sizeof(int64_t) * J.tasks.size();
  if (tb + bb + pb > (size_t)kNoAllocScratchBytes)
    return fail(GZ_CAPACITY_ERROR, "gz_ipc_decompress_noalloc: " + std::to_string(J.tasks.size()) + " buffers and " +
                                       std::to_string(J.blocks.size()) + " LZ4 blocks need " +
                                       std::to_string(tb + bb + pb) + " bytes of scratch, more than its " +
                                       std::to_string(kNoAllocScratchBytes));
  GZ_HIP(hipSetDevice(ctx->device));
  uint8_t* sp = nullptr;
  GZ_TRY(scratch_address(ctx->device, &sp));
  Lz4Task* d_tasks = (Lz4Task*)sp;
  Lz4Block* d_blocks = (Lz4Block*)(sp + tb);
  int64_t* d_prod = (int64_t*)(sp + tb + bb);  // every task's word is written by the kernel
  GZ_HIP(hipMemcpyAsync(d_tasks, J.tasks.data(), sizeof(Lz4Task) * J.tasks.size(), hipMemcpyHostToDevice, ctx->stream));
  if (!J.blocks.empty())
    GZ_HIP(hipMemcpyAsync(d_blocks, J.blocks.data(), sizeof(Lz4Block) * J.blocks.size(), hipMemcpyHostToDevice,
                          ctx->stream));
  const int nt = (int)J.tasks.size();
  {
    KernelTimer kt(ctx, "ipc_lz4_decode");
    hipLaunchKernelGGL(lz4_decode_kernel, dim3((nt + 3) / 4), dim3(256), 0, ctx->stream, (const Lz4Task*)d_tasks, nt,
                       (const Lz4Block*)d_blocks, d_prod);
  }
  std::vector<int64_t> prod(J.tasks.size());
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(prod.data(), d_prod, pb, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return fail(GZ_HIP_ERROR, std::string("gz_ipc_decompress_noalloc: ") + hipGetErrorString(e));
  return lz4_check(J, prod);
}

}  // extern "C"
