// LZ4 frame decoding on the CPU (test infrastructure: the checker of the device
// decoder, never part of the product).
//
// The reference compresses shuffle buffers with arrow::Compression::LZ4_FRAME
// (splitter.cc:315-317; the Spark default codec "lz4" maps to it,
// jni_common.h:453-468) and reads them back with DecompressBuffer
// (jni_common.h:471-492): an int64 little-endian uncompressed length, then one
// LZ4 frame.  The codec lives in the lz4 library that arrow-4.0.0 bundles
// (lz4 1.9.x; absent from the reference tree), so this restates the published
// formats it implements: "LZ4 Frame Format Description" 1.6.x and "LZ4 Block
// Format Description".  Pinned by pyarrow's LZ4_FRAME codec (the same lz4
// library) in tests/test_oracle_lz4.py.
#include <stdint.h>
#include <string.h>

static uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}

// One LZ4 block: sequences of (token, literals, offset, match) until the input
// ends after a literal run.  Matches may reach back into earlier output of the
// same frame (linked blocks), down to dst_frame.  Returns bytes written or -1.
static int64_t block_decode(const uint8_t* ip, int64_t n, uint8_t* dst_frame, int64_t op, int64_t cap) {
  const uint8_t* iend = ip + n;
  const int64_t op0 = op;
  while (ip < iend) {
    const uint32_t token = *ip++;
    int64_t lit = token >> 4;
    if (lit == 15) {
      uint32_t b;
      do {
        if (ip >= iend) return -1;
        b = *ip++;
        lit += b;
      } while (b == 255);
    }
    if (lit > iend - ip || lit > cap - op) return -1;
    memcpy(dst_frame + op, ip, (size_t)lit);
    ip += lit;
    op += lit;
    if (ip == iend) break;  // the last sequence carries literals only
    if (iend - ip < 2) return -1;
    const uint32_t off = (uint32_t)ip[0] | ((uint32_t)ip[1] << 8);
    ip += 2;
    int64_t mlen = (token & 15);
    if (mlen == 15) {
      uint32_t b;
      do {
        if (ip >= iend) return -1;
        b = *ip++;
        mlen += b;
      } while (b == 255);
    }
    mlen += 4;
    if (off == 0 || off > op || mlen > cap - op) return -1;
    for (int64_t k = 0; k < mlen; k++) dst_frame[op + k] = dst_frame[op - off + k];  // overlap copies forward
    op += mlen;
  }
  return op - op0;
}

// Every frame in src (LZ4 frames and skippable frames back to back) into dst.
// Returns the decompressed length, or -1 on malformed input or overflow.
int64_t gzo_lz4f_decompress(const uint8_t* src, int64_t n, uint8_t* dst, int64_t cap) {
  int64_t ip = 0, op = 0;
  while (ip < n) {
    if (n - ip < 4) return -1;
    const uint32_t magic = rd32(src + ip);
    if ((magic & 0xFFFFFFF0u) == 0x184D2A50u) {  // skippable frame
      if (n - ip < 8) return -1;
      ip += 8 + (int64_t)rd32(src + ip + 4);
      continue;
    }
    if (magic != 0x184D2204u || n - ip < 7) return -1;
    const uint8_t flg = src[ip + 4];
    if ((flg >> 6) != 1 || (flg & 0x02)) return -1;  // version 01, reserved bit
    const int block_checksum = (flg >> 4) & 1, content_size = (flg >> 3) & 1;
    const int content_checksum = (flg >> 2) & 1, dict_id = flg & 1;
    if (dict_id) return -1;  // dictionaries are not used by the codec
    int64_t hp = ip + 6 + (content_size ? 8 : 0);
    int64_t declared = -1;
    if (content_size) memcpy(&declared, src + ip + 6, 8);
    if (hp + 1 > n) return -1;
    ip = hp + 1;  // header checksum byte
    const int64_t frame_start = op;
    for (;;) {
      if (n - ip < 4) return -1;
      const uint32_t bs = rd32(src + ip);
      ip += 4;
      if (bs == 0) break;  // end mark
      const int64_t len = bs & 0x7FFFFFFFu;
      if (len > n - ip) return -1;
      if (bs & 0x80000000u) {  // stored block
        if (len > cap - op) return -1;
        memcpy(dst + op, src + ip, (size_t)len);
        op += len;
      } else {
        const int64_t w = block_decode(src + ip, len, dst + frame_start, op - frame_start, cap - frame_start);
        if (w < 0) return -1;
        op += w;
      }
      ip += len + (block_checksum ? 4 : 0);
    }
    if (content_checksum) ip += 4;
    if (ip > n) return -1;
    if (declared >= 0 && declared != op - frame_start) return -1;
  }
  return op;
}
