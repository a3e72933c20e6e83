// The library's host BLAKE3 (flock-challenge_amd/csrc/host_blake3.cpp, used by
// the transcript, the verifier and the host grinds) against the oracle's:
// the compression function on random inputs, and the incremental hasher's
// XOF and non-root outputs on random data fed in random pieces.  Prints "ok".
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "host_blake3.h"
extern "C" {
#include "oracle.h"
}

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint64_t next() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return rng;
}

int main() {
  for (int it = 0; it < 200000; ++it) {
    uint32_t cv[8], m[16], a[16], b[16];
    for (auto &w : cv) w = (uint32_t)next();
    for (auto &w : m) w = (uint32_t)next();
    const uint64_t counter = next() >> (next() & 63);
    const uint32_t len = (uint32_t)(next() % 65), flags = (uint32_t)(next() & 15);
    flock::hb3::compress(cv, m, counter, len, flags, a);
    ob3_compress(cv, m, counter, len, flags, b);
    if (memcmp(a, b, sizeof a)) {
      printf("compress differs at iteration %d\n", it);
      return 1;
    }
  }
  for (int it = 0; it < 300; ++it) {
    const size_t n = (size_t)(next() % 20000);
    std::vector<uint8_t> d(n);
    for (auto &x : d) x = (uint8_t)next();
    flock::hb3::Hasher h;
    ob3_hasher o;
    ob3_init(&o);
    for (size_t off = 0; off < n;) {
      size_t piece = (size_t)(next() % 300);
      if (piece > n - off) piece = n - off;
      h.update(d.data() + off, piece);
      ob3_update(&o, d.data() + off, piece);
      off += piece;
    }
    uint8_t x1[200], x2[200], r1[32], r2[32];
    h.finalize_xof(x1, sizeof x1);
    ob3_finalize_xof(&o, x2, sizeof x2);
    h.finalize_non_root(r1);
    ob3_finalize_non_root(&o, r2);
    if (memcmp(x1, x2, sizeof x1) || memcmp(r1, r2, sizeof r1)) {
      printf("hasher differs for %zu bytes\n", n);
      return 1;
    }
  }
  printf("ok\n");
  return 0;
}
