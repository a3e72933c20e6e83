// Host BLAKE3 (public BLAKE3 specification): compression, incremental hasher
// with the chunk-CV stack, root XOF, and the two Merkle node functions the
// reference uses (crates/flock-core/src/merkle.rs:200-210).  Used for the
// Fiat-Shamir transcript, the statement digest and the PoW state digest.
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <string.h>

namespace flock {
namespace hb3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };
static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t len, uint32_t flags,
              uint32_t out[16]);

struct Hasher {
  uint32_t cv[8];
  uint64_t chunk;
  uint8_t block[64];
  uint32_t block_len = 0, blocks_done = 0;
  uint32_t stack[54][8];
  uint32_t depth = 0;
  Hasher() { reset(); }
  void reset();
  void update(const void *data, size_t n);
  void finalize_xof(uint8_t *out, size_t n) const;  // root output from offset 0
  void finalize_non_root(uint8_t out[32]) const;
};

inline void hash(const void *d, size_t n, uint8_t out[32]) {
  Hasher h;
  h.update(d, n);
  h.finalize_xof(out, 32);
}

}  // namespace hb3
}  // namespace flock
