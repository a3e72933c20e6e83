// Products v * t in GF(2^128) by a wavefront-uniform constant t through
// shifted XORs: no tables, no LDS, a cost proportional to the set bits of t.
//
// The deepest layers of the additive NTT over the standard basis
// {1, x, .., x^(d-1)} have short twiddles: layer d-1's are in the span of
// x .. x^(d-1) (at most d bits) and layer d-2's in the span of
// x^(j-1) (1 + x + .. + x^(j-1)), j = 2 .. d-1 (at most 2d - 3 bits)
// (ntt/additive_ntt_f128.rs:57-127: evals[0] is the basis itself, evals[1]
// its normalized W_1 images).  There a twiddle meets only 64 .. 128
// products, so building its nibble tables costs as much as the lookups;
// the shifted XORs below cost about 7.5 VALU per set bit of t instead, and
// a zero twiddle (block 0 of every layer) costs nothing.
//
// v * t = sum over set bits j of t of v x^j: each v x^j is v shifted into a
// 256-bit accumulator, whose bits 128 .. 255 (H) then fold back as
// H (x^7 + x^2 + x + 1) (x^128 = x^7 + x^2 + x + 1 mod P).
//
// Host and device compile the same code (the CPU test checks it against
// an independent GF(2^128)); on the device t is wavefront-uniform, so the
// loop over its set bits runs on the scalar unit and the shift counts are
// scalar operands.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define FLOCK_SHORT_HD __host__ __device__ __forceinline__
#else
#define FLOCK_SHORT_HD inline
#endif

namespace flock {
namespace shortmul {

// bits s .. s + 31 of the 64-bit value hi:lo, 0 < s < 32 (one v_alignbit)
FLOCK_SHORT_HD uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, s);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}

FLOCK_SHORT_HD uint32_t x3(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
#else
  return a ^ b ^ c;
#endif
}

// acc[OFF .. OFF + 4] ^= v x^j for every set bit j of m (0 < j < 32): the
// lowest bit alone when the count is odd, then two bits per step, so each
// accumulator word takes one three-input XOR and the loop has one body.
template <int OFF>
FLOCK_SHORT_HD void add_shifts(uint32_t (&acc)[8], const uint32_t (&v)[4], uint32_t m) {
  if (__builtin_popcount(m) & 1) {
    const uint32_t j1 = (uint32_t)__builtin_ctz(m), s1 = 32 - j1;
    m &= m - 1;
    acc[OFF + 0] ^= v[0] << j1;
    acc[OFF + 1] ^= funnel(v[1], v[0], s1);
    acc[OFF + 2] ^= funnel(v[2], v[1], s1);
    acc[OFF + 3] ^= funnel(v[3], v[2], s1);
    acc[OFF + 4] ^= v[3] >> s1;
  }
  while (m) {
    const uint32_t j1 = (uint32_t)__builtin_ctz(m);
    m &= m - 1;
    const uint32_t j2 = (uint32_t)__builtin_ctz(m);
    m &= m - 1;
    const uint32_t s1 = 32 - j1, s2 = 32 - j2;
    acc[OFF + 0] = x3(acc[OFF + 0], v[0] << j1, v[0] << j2);
    acc[OFF + 1] = x3(acc[OFF + 1], funnel(v[1], v[0], s1), funnel(v[1], v[0], s2));
    acc[OFF + 2] = x3(acc[OFF + 2], funnel(v[2], v[1], s1), funnel(v[2], v[1], s2));
    acc[OFF + 3] = x3(acc[OFF + 3], funnel(v[3], v[2], s1), funnel(v[3], v[2], s2));
    acc[OFF + 4] = x3(acc[OFF + 4], v[3] >> s1, v[3] >> s2);
  }
}

// acc[OFF .. OFF + 3] ^= v when bit 0 of the word is set, then its other
// bits through add_shifts.
template <int OFF>
FLOCK_SHORT_HD void add_word(uint32_t (&acc)[8], const uint32_t (&v)[4], uint32_t t) {
  if (t & 1u) {
    acc[OFF + 0] ^= v[0];
    acc[OFF + 1] ^= v[1];
    acc[OFF + 2] ^= v[2];
    acc[OFF + 3] ^= v[3];
  }
  add_shifts<OFF>(acc, v, t & ~1u);
}

// out = s + v * t mod P for t = t0 + 2^32 t1 + 2^64 t2 + 2^96 t3 (any t;
// the cost is proportional to its set bits, and a zero word costs a scalar
// test).  v, s and out are the four little-endian words of an F128 (bits
// 0..31 first); s enters the low half of the unreduced sum, which the
// reduction leaves as it is.
FLOCK_SHORT_HD void mul_add(const uint32_t (&v)[4], uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3,
                            const uint32_t (&s)[4], uint32_t (&out)[4]) {
  uint32_t acc[8] = {s[0], s[1], s[2], s[3], 0u, 0u, 0u, 0u};
  if (t0) add_word<0>(acc, v, t0);
  if (t1) add_word<1>(acc, v, t1);
  if (t2) add_word<2>(acc, v, t2);
  if (t3) add_word<3>(acc, v, t3);
  // H = acc[4..7] (bits 128 .. 255): H (1 + x + x^2 + x^7) spans bits 0 ..
  // 262 - 128 = 134, whose top 7 bits (g) fold once more
  const uint32_t h0 = acc[4], h1 = acc[5], h2 = acc[6], h3 = acc[7];
  const uint32_t r0 = x3(h0 ^ (h0 << 1), h0 << 2, h0 << 7);
  const uint32_t r1 = x3(h1 ^ funnel(h1, h0, 31), funnel(h1, h0, 30), funnel(h1, h0, 25));
  const uint32_t r2 = x3(h2 ^ funnel(h2, h1, 31), funnel(h2, h1, 30), funnel(h2, h1, 25));
  const uint32_t r3 = x3(h3 ^ funnel(h3, h2, 31), funnel(h3, h2, 30), funnel(h3, h2, 25));
  const uint32_t g = x3(h3 >> 31, h3 >> 30, h3 >> 25);  // bits 128 .. 134
  out[0] = x3(acc[0], r0, x3(g, g << 1, (g << 2) ^ (g << 7)));
  out[1] = acc[1] ^ r1;
  out[2] = acc[2] ^ r2;
  out[3] = acc[3] ^ r3;
}

// out = v * t mod P.
FLOCK_SHORT_HD void mul(const uint32_t (&v)[4], uint32_t t0, uint32_t t1, uint32_t t2, uint32_t t3,
                        uint32_t (&out)[4]) {
  const uint32_t zero[4] = {0u, 0u, 0u, 0u};
  mul_add(v, t0, t1, t2, t3, zero, out);
}

}  // namespace shortmul
}  // namespace flock
