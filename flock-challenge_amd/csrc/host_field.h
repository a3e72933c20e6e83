// Host-side GF(2^128) / GF(2^8) arithmetic for the prover's host logic
// (twiddle generation, transcript-side algebra, verifier replay).
// Same representation as the reference F128 (crates/flock-core/src/field/
// gf2_128.rs:22-26): lo = x^0..x^63, hi = x^64..x^127, modulus 0x87.
#pragma once
#include <stdint.h>
#include <wmmintrin.h>
#include <smmintrin.h>

namespace flock {

struct F128 {
  uint64_t lo = 0, hi = 0;
  constexpr F128() = default;
  constexpr F128(uint64_t l, uint64_t h) : lo(l), hi(h) {}
  bool operator==(const F128 &o) const { return lo == o.lo && hi == o.hi; }
  bool operator!=(const F128 &o) const { return !(*this == o); }
  F128 operator+(const F128 &o) const { return F128(lo ^ o.lo, hi ^ o.hi); }
  F128 &operator+=(const F128 &o) { lo ^= o.lo; hi ^= o.hi; return *this; }
  bool is_zero() const { return (lo | hi) == 0; }
  static constexpr F128 zero() { return F128(0, 0); }
  static constexpr F128 one() { return F128(1, 0); }
};

struct F256 {
  uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
  F256 &operator^=(const F256 &o) { r0 ^= o.r0; r1 ^= o.r1; r2 ^= o.r2; r3 ^= o.r3; return *this; }
};

__attribute__((target("pclmul,sse4.1"))) inline F256 mul_unreduced(F128 a, F128 b) {
  __m128i x = _mm_set_epi64x((long long)a.hi, (long long)a.lo);
  __m128i y = _mm_set_epi64x((long long)b.hi, (long long)b.lo);
  __m128i ll = _mm_clmulepi64_si128(x, y, 0x00);
  __m128i hh = _mm_clmulepi64_si128(x, y, 0x11);
  __m128i lh = _mm_xor_si128(_mm_clmulepi64_si128(x, y, 0x10), _mm_clmulepi64_si128(x, y, 0x01));
  F256 r;
  r.r0 = (uint64_t)_mm_cvtsi128_si64(ll);
  r.r1 = (uint64_t)_mm_extract_epi64(ll, 1) ^ (uint64_t)_mm_cvtsi128_si64(lh);
  r.r2 = (uint64_t)_mm_cvtsi128_si64(hh) ^ (uint64_t)_mm_extract_epi64(lh, 1);
  r.r3 = (uint64_t)_mm_extract_epi64(hh, 1);
  return r;
}

// Fold bits 128..255 with x^128 = x^7 + x^2 + x + 1.
inline F128 reduce(F256 u) {
  // Two-step fold: the top word's overflow past bit 127 is folded again.
  uint64_t r2 = u.r2, r3 = u.r3;
  uint64_t lo = u.r0 ^ r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
  uint64_t hi = u.r1 ^ r3 ^ (r3 << 1) ^ (r2 >> 63) ^ (r3 << 2) ^ (r2 >> 62) ^ (r3 << 7) ^ (r2 >> 57);
  uint64_t over = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
  lo ^= over ^ (over << 1) ^ (over << 2) ^ (over << 7);
  return F128(lo, hi);
}

inline F128 operator*(F128 a, F128 b) { return reduce(mul_unreduced(a, b)); }
inline F128 &operator*=(F128 &a, F128 b) { a = a * b; return a; }

inline F128 mul_x(F128 v) {
  uint64_t c = v.hi >> 63;
  return F128((v.lo << 1) ^ (0x87 & (0 - c)), (v.hi << 1) | (v.lo >> 63));
}

inline F128 inv(F128 a) {
  // a^(2^128 - 2) = prod_{i=1..127} a^(2^i)
  F128 r = F128::one();
  F128 sq = a * a;
  for (int i = 1; i < 128; ++i) {
    r = r * sq;
    sq = sq * sq;
  }
  return r;
}

// GF(2^8), AES polynomial x^8 + x^4 + x^3 + x + 1.
inline uint8_t f8_mul(uint8_t a, uint8_t b) {
  uint32_t p = 0;
  for (int i = 0; i < 8; ++i)
    if ((b >> i) & 1) p ^= (uint32_t)a << i;
  for (int i = 14; i >= 8; --i)
    if ((p >> i) & 1) p ^= 0x11Bu << (i - 8);
  return (uint8_t)p;
}

// phi8: the embedding GF(2^8) -> GF(2^128) fixed by the image of 0x02
// (crates/flock-core/src/field/phi8.rs:27-30), extended GF(2)-linearly.
struct Phi8 {
  F128 table[256];
  Phi8() {
    F128 g(0x6b8330483c2e9849ULL, 0x0dcb364640a222feULL);
    F128 pw[8];
    pw[0] = F128::one();
    for (int i = 1; i < 8; ++i) pw[i] = pw[i - 1] * g;
    for (int a = 0; a < 256; ++a) {
      F128 acc;
      for (int i = 0; i < 8; ++i)
        if ((a >> i) & 1) acc += pw[i];
      table[a] = acc;
    }
  }
};
const Phi8 &phi8_table();
inline F128 phi8(uint8_t a) { return phi8_table().table[a]; }

}  // namespace flock
