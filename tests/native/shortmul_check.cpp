// The table-free product by a wavefront-uniform constant (flock-challenge_amd/
// csrc/gf128_short.h, used by the NTT's deepest layers) against the oracle's
// GF(2^128) product: random v against constants of every bit length up to
// 128 and of every popcount, single bits at each position, the twiddles of
// the two deepest layers of the dimension-20 standard-basis NTT (which must
// fit in 64 bits, the reason the path exists), and zero.  Prints "ok".
#include <stdint.h>
#include <stdio.h>

#include "gf128_short.h"
extern "C" {
#include "oracle.h"
}

static uint64_t rng = 0x2545F4914F6CDD1Dull;
static uint64_t next() {
  rng ^= rng << 13;
  rng ^= rng >> 7;
  rng ^= rng << 17;
  return rng;
}

static int check(of128 v, of128 t) {
  const uint32_t vw[4] = {(uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi, (uint32_t)(v.hi >> 32)};
  uint32_t o[4];
  flock::shortmul::mul(vw, (uint32_t)t.lo, (uint32_t)(t.lo >> 32), (uint32_t)t.hi, (uint32_t)(t.hi >> 32), o);
  const of128 want = of128_mul(v, t);
  const uint64_t lo = o[0] | ((uint64_t)o[1] << 32), hi = o[2] | ((uint64_t)o[3] << 32);
  if (lo != want.lo || hi != want.hi) {
    printf("v=%016llx%016llx t=%016llx%016llx: got %016llx%016llx want %016llx%016llx\n",
           (unsigned long long)v.hi, (unsigned long long)v.lo, (unsigned long long)t.hi, (unsigned long long)t.lo,
           (unsigned long long)hi, (unsigned long long)lo, (unsigned long long)want.hi,
           (unsigned long long)want.lo);
    return 1;
  }
  // the seeded form (a butterfly's u0 = s + v t) against s XOR the product
  const uint64_t slo = next(), shi = next();
  const uint32_t sw[4] = {(uint32_t)slo, (uint32_t)(slo >> 32), (uint32_t)shi, (uint32_t)(shi >> 32)};
  flock::shortmul::mul_add(vw, (uint32_t)t.lo, (uint32_t)(t.lo >> 32), (uint32_t)t.hi, (uint32_t)(t.hi >> 32), sw, o);
  const uint64_t alo = o[0] | ((uint64_t)o[1] << 32), ahi = o[2] | ((uint64_t)o[3] << 32);
  if (alo != (want.lo ^ slo) || ahi != (want.hi ^ shi)) {
    printf("mul_add v=%016llx%016llx t=%016llx%016llx: wrong\n", (unsigned long long)v.hi,
           (unsigned long long)v.lo, (unsigned long long)t.hi, (unsigned long long)t.lo);
    return 1;
  }
  return 0;
}

static of128 masked(int len) {
  of128 t = {next(), next()};
  if (len < 128) {
    if (len <= 64) {
      t.hi = 0;
      t.lo = len == 64 ? t.lo : len == 0 ? 0 : t.lo & ((1ull << len) - 1);
    } else {
      t.hi &= (1ull << (len - 64)) - 1;
    }
  }
  return t;
}

int main() {
  for (int it = 0; it < 200000; ++it) {
    of128 v = {next(), next()};
    if (it % 7 == 0) v.hi |= 0xFF00000000000000ull;  // top bits set: both folds
    if (check(v, masked(it % 129))) return 1;
  }
  for (int j = 0; j < 128; ++j)
    for (int it = 0; it < 50; ++it) {
      of128 v = {next(), next() | (1ull << 63)};
      of128 t = {j < 64 ? 1ull << j : 0, j >= 64 ? 1ull << (j - 64) : 0};
      if (check(v, t)) return 1;
    }
  // every popcount, bits spread over all 128 positions
  for (int pc = 0; pc <= 128; ++pc)
    for (int it = 0; it < 100; ++it) {
      of128 t = {0, 0};
      while (__builtin_popcountll(t.lo) + __builtin_popcountll(t.hi) < pc) {
        const int b = (int)(next() & 127);
        if (b < 64) t.lo |= 1ull << b; else t.hi |= 1ull << (b - 64);
      }
      of128 v = {next(), next()};
      if (check(v, t)) return 1;
    }
  // the twiddles of the two deepest layers of the dimension-20 standard NTT:
  // at most 20 and 37 bits
  ontt nt;
  if (ontt_init(&nt, 20)) {
    printf("ntt init failed\n");
    return 1;
  }
  for (int layer = 18; layer < 20; ++layer)
    for (size_t b = 0; b < ((size_t)1 << layer); b += 97) {
      const of128 t = ontt_twiddle(&nt, layer, b);
      const int bound = layer == 19 ? 20 : 37;
      if (t.hi || (t.lo >> bound)) {
        printf("layer %d block %zu: twiddle above %d bits\n", layer, b, bound);
        return 1;
      }
      of128 v = {next(), next()};
      if (check(v, t)) return 1;
    }
  ontt_free(&nt);
  printf("ok\n");
  return 0;
}
