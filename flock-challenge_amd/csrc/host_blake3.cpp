#include "host_blake3.h"

namespace flock {
namespace hb3 {

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// Message word order of each round: the BLAKE3 permutation
// (2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8) applied r times, so
// the rounds index the block directly instead of permuting it.
static const uint8_t SCHEDULE[7][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15},
    {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8},
    {3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1},
    {10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6},
    {12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4},
    {9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7},
    {11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13}};

void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t len, uint32_t flags,
              uint32_t out[16]) {
  uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7], IV[0], IV[1], IV[2], IV[3],
                    (uint32_t)counter, (uint32_t)(counter >> 32), len, flags};
  auto g = [&](int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] = v[a] + v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] = v[a] + v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
  };
  for (int r = 0; r < 7; ++r) {
    const uint8_t *s = SCHEDULE[r];
    g(0, 4, 8, 12, m[s[0]], m[s[1]]);
    g(1, 5, 9, 13, m[s[2]], m[s[3]]);
    g(2, 6, 10, 14, m[s[4]], m[s[5]]);
    g(3, 7, 11, 15, m[s[6]], m[s[7]]);
    g(0, 5, 10, 15, m[s[8]], m[s[9]]);
    g(1, 6, 11, 12, m[s[10]], m[s[11]]);
    g(2, 7, 8, 13, m[s[12]], m[s[13]]);
    g(3, 4, 9, 14, m[s[14]], m[s[15]]);
  }
  for (int i = 0; i < 8; ++i) {
    out[i] = v[i] ^ v[i + 8];
    out[i + 8] = v[i + 8] ^ cv[i];
  }
}

static void le_words(const uint8_t *b, uint32_t *w) {
  for (int i = 0; i < 16; ++i) memcpy(&w[i], b + 4 * i, 4);
}

void Hasher::reset() {
  memcpy(cv, IV, 32);
  chunk = 0;
  memset(block, 0, 64);
  block_len = 0;
  blocks_done = 0;
  depth = 0;
}

void Hasher::update(const void *data, size_t n) {
  const uint8_t *p = (const uint8_t *)data;
  while (n) {
    if (block_len == 64) {
      // A full block is only compressed once more input arrives, so the
      // final block of the stream is always available for finalization.
      uint32_t w[16], o[16];
      le_words(block, w);
      if (blocks_done == 15) {
        // Block 16 of a full chunk: close the chunk and push its CV.
        compress(cv, w, chunk, 64, (blocks_done == 0 ? CHUNK_START : 0) | CHUNK_END, o);
        uint32_t c[8];
        memcpy(c, o, 32);
        uint64_t total = chunk + 1;
        while ((total & 1) == 0) {
          uint32_t pm[16], po[16];
          memcpy(pm, stack[depth - 1], 32);
          memcpy(pm + 8, c, 32);
          compress(IV, pm, 0, 64, PARENT, po);
          memcpy(c, po, 32);
          --depth;
          total >>= 1;
        }
        memcpy(stack[depth++], c, 32);
        ++chunk;
        memcpy(cv, IV, 32);
        blocks_done = 0;
      } else {
        compress(cv, w, chunk, 64, blocks_done == 0 ? CHUNK_START : 0, o);
        memcpy(cv, o, 32);
        ++blocks_done;
      }
      memset(block, 0, 64);
      block_len = 0;
    }
    size_t take = 64 - block_len;
    if (take > n) take = n;
    memcpy(block + block_len, p, take);
    block_len += (uint32_t)take;
    p += take;
    n -= take;
  }
}

struct Node {
  uint32_t cv[8], m[16];
  uint64_t counter;
  uint32_t len, flags;
};

static Node final_node(const Hasher &h) {
  Node nd;
  memcpy(nd.cv, h.cv, 32);
  le_words(h.block, nd.m);
  nd.counter = h.chunk;
  nd.len = h.block_len;
  nd.flags = (h.blocks_done == 0 ? CHUNK_START : 0) | CHUNK_END;
  for (int i = (int)h.depth - 1; i >= 0; --i) {
    uint32_t o[16];
    compress(nd.cv, nd.m, nd.counter, nd.len, nd.flags, o);
    memcpy(nd.m, h.stack[i], 32);
    memcpy(nd.m + 8, o, 32);
    memcpy(nd.cv, IV, 32);
    nd.counter = 0;
    nd.len = 64;
    nd.flags = PARENT;
  }
  return nd;
}

void Hasher::finalize_xof(uint8_t *out, size_t n) const {
  Node nd = final_node(*this);
  uint64_t k = 0;
  while (n) {
    uint32_t o[16];
    compress(nd.cv, nd.m, k++, nd.len, nd.flags | ROOT, o);
    size_t take = n < 64 ? n : 64;
    memcpy(out, o, take);
    out += take;
    n -= take;
  }
}

void Hasher::finalize_non_root(uint8_t out[32]) const {
  Node nd = final_node(*this);
  uint32_t o[16];
  compress(nd.cv, nd.m, nd.counter, nd.len, nd.flags, o);
  memcpy(out, o, 32);
}

}  // namespace hb3
}  // namespace flock
