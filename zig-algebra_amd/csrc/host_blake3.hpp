// BLAKE3 (32-byte output, unkeyed) on the host, for the Fiat-Shamir
// transcript.  The transcript hashes a length-prefixed byte stream that can
// span many 1024-byte chunks, so this is the whole hash: 64-byte blocks
// compressed by 7 rounds of the quarter-round G, chunks chained by the chunk
// counter, chunk chaining values merged pairwise into a binary tree (PARENT
// nodes), the last compression flagged ROOT.  That is BLAKE3 as its
// specification defines it, which is what std.crypto.hash.Blake3 -- the hash
// zig-transcript uses (libs/transcript/src/root.zig:31) -- implements.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace zntt {
namespace host {

class Blake3 {
 public:
  // Hash of `len` bytes at `in`.
  static void hash(const uint8_t *in, size_t len, uint8_t out[32]) {
    std::vector<Cv> stack;
    uint64_t chunk = 0;
    size_t pos = 0;
    for (;;) {
      const size_t rest = len - pos;
      const bool last = rest <= kChunk;
      const size_t clen = last ? rest : kChunk;
      const size_t blocks = clen == 0 ? 1 : (clen + 63) / 64;
      Cv cv = iv();
      for (size_t b = 0; b < blocks; b++) {
        const size_t blen = clen - b * 64 < 64 ? clen - b * 64 : 64;
        uint32_t flags = (b == 0 ? kStart : 0) | (b + 1 == blocks ? kEnd : 0);
        if (last && stack.empty() && b + 1 == blocks) flags |= kRoot;
        cv = compress(cv, words(in + pos + b * 64, blen), chunk, (uint32_t)blen, flags);
      }
      if (last) {
        while (!stack.empty()) {
          const Cv left = stack.back();
          stack.pop_back();
          cv = parent(left, cv, stack.empty() ? kRoot : 0);
        }
        for (int i = 0; i < 8; i++)
          for (int k = 0; k < 4; k++) out[4 * i + k] = (uint8_t)(cv.w[i] >> (8 * k));
        return;
      }
      chunk++;
      for (uint64_t total = chunk; (total & 1) == 0; total >>= 1) {
        cv = parent(stack.back(), cv, 0);
        stack.pop_back();
      }
      stack.push_back(cv);
      pos += kChunk;
    }
  }

 private:
  struct Cv {
    uint32_t w[8];
  };
  struct Block {
    uint32_t w[16];
  };
  static constexpr size_t kChunk = 1024;
  static constexpr uint32_t kStart = 1, kEnd = 2, kParent = 4, kRoot = 8;

  static Cv iv() {
    return Cv{{0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au, 0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu,
               0x5BE0CD19u}};
  }
  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

  static Block words(const uint8_t *p, size_t len) {
    uint8_t buf[64] = {0};
    if (len) memcpy(buf, p, len);
    Block m;
    for (int i = 0; i < 16; i++)
      m.w[i] = (uint32_t)buf[4 * i] | (uint32_t)buf[4 * i + 1] << 8 | (uint32_t)buf[4 * i + 2] << 16 |
               (uint32_t)buf[4 * i + 3] << 24;
    return m;
  }

  static void g(uint32_t *v, int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] += v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] += v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] += v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] += v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
  }

  static Cv compress(const Cv &cv, Block m, uint64_t counter, uint32_t len, uint32_t flags) {
    static const uint8_t perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
    const Cv k = iv();
    uint32_t v[16] = {cv.w[0], cv.w[1], cv.w[2], cv.w[3], cv.w[4], cv.w[5], cv.w[6], cv.w[7],
                      k.w[0],  k.w[1],  k.w[2],  k.w[3],  (uint32_t)counter, (uint32_t)(counter >> 32), len,
                      flags};
    for (int r = 0; r < 7; r++) {
      g(v, 0, 4, 8, 12, m.w[0], m.w[1]);
      g(v, 1, 5, 9, 13, m.w[2], m.w[3]);
      g(v, 2, 6, 10, 14, m.w[4], m.w[5]);
      g(v, 3, 7, 11, 15, m.w[6], m.w[7]);
      g(v, 0, 5, 10, 15, m.w[8], m.w[9]);
      g(v, 1, 6, 11, 12, m.w[10], m.w[11]);
      g(v, 2, 7, 8, 13, m.w[12], m.w[13]);
      g(v, 3, 4, 9, 14, m.w[14], m.w[15]);
      Block t;
      for (int i = 0; i < 16; i++) t.w[i] = m.w[perm[i]];
      m = t;
    }
    Cv o;
    for (int i = 0; i < 8; i++) o.w[i] = v[i] ^ v[8 + i];
    return o;
  }

  static Cv parent(const Cv &l, const Cv &r, uint32_t flags) {
    Block m;
    memcpy(m.w, l.w, 32);
    memcpy(m.w + 8, r.w, 32);
    return compress(iv(), m, 0, 64, kParent | flags);
  }
};

}  // namespace host
}  // namespace zntt
