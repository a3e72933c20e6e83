// GF(2^128) arithmetic on gfx950 wavefronts.
//
// Representation: the reference's F128 {lo, hi} (crates/flock-core/src/field/
// gf2_128.rs:22-26) stored little-endian as four dwords
//   x = bits 0..31, y = 32..63, z = 64..95, w = 96..127,
// modulus x^128 + x^7 + x^2 + x + 1.  CDNA4 has no carry-less multiply, so
// products by a wavefront-uniform constant go through per-constant lookup
// tables in LDS (fulltab.h, nibtab.h) or, for short constants, shifted XORs
// (gf128_short.h), and general products through 4-bit windowed tables built
// per operand (f128_mul below).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef uint4 f128;

__device__ __forceinline__ f128 f128_zero() { return make_uint4(0u, 0u, 0u, 0u); }
__device__ __forceinline__ f128 f128_xor(f128 a, f128 b) {
  return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}
__device__ __forceinline__ void f128_xor_into(f128 &a, f128 b) {
  a.x ^= b.x; a.y ^= b.y; a.z ^= b.z; a.w ^= b.w;
}

// Streaming accesses (the nt cache policy): for tables read or written once
// per sweep, far larger than the L2.
typedef uint32_t f128_v4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f128 f128_load_nt(const f128 *p) {
  const f128_v4 v = __builtin_nontemporal_load((const f128_v4 *)p);
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void f128_store_nt(f128 *p, f128 a) {
  f128_v4 v;
  v.x = a.x;
  v.y = a.y;
  v.z = a.z;
  v.w = a.w;
  __builtin_nontemporal_store(v, (f128_v4 *)p);
}

// a ^ b ^ c in one v_bitop3_b32 per dword (truth table 0x96).
__device__ __forceinline__ uint32_t xor3_u32(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ f128 f128_xor3(f128 a, f128 b, f128 c) {
  return make_uint4(xor3_u32(a.x, b.x, c.x), xor3_u32(a.y, b.y, c.y), xor3_u32(a.z, b.z, c.z),
                    xor3_u32(a.w, b.w, c.w));
}

// a * x mod P
__device__ __forceinline__ f128 f128_mulx(f128 v) {
  uint32_t c = v.w >> 31;
  f128 r;
  r.w = (v.w << 1) | (v.z >> 31);
  r.z = (v.z << 1) | (v.y >> 31);
  r.y = (v.y << 1) | (v.x >> 31);
  r.x = (v.x << 1) ^ (0x87u & (0u - c));
  return r;
}

// h * (x^7 + x^2 + x + 1) for h < 2^25 fits in one dword.
__device__ __forceinline__ uint32_t fold87(uint32_t h) { return h ^ (h << 1) ^ (h << 2) ^ (h << 7); }

// a * x^4 mod P: the 4 bits shifted out fold back as h * 0x87 (< 2^11).
__device__ __forceinline__ f128 f128_shl4(f128 a) {
  uint32_t h = a.w >> 28;
  f128 r;
  r.w = __builtin_amdgcn_alignbit(a.w, a.z, 28);
  r.z = __builtin_amdgcn_alignbit(a.z, a.y, 28);
  r.y = __builtin_amdgcn_alignbit(a.y, a.x, 28);
  r.x = (a.x << 4) ^ fold87(h);
  return r;
}

// Reduce a 256-bit carry-less product (d0 = bits 0..31 ... d7) mod P.
__device__ __forceinline__ f128 f128_reduce256(uint32_t l0, uint32_t l1, uint32_t l2, uint32_t l3, uint32_t h0,
                                               uint32_t h1, uint32_t h2, uint32_t h3) {
  // H x^128 = H (x^7 + x^2 + x + 1); bits pushed past 127 fold once more.
  const uint32_t o = xor3_u32(h3 >> 31, h3 >> 30, h3 >> 25);
  f128 r;
  r.x = xor3_u32(xor3_u32(l0, h0, h0 << 1), xor3_u32(h0 << 2, h0 << 7, o),
                 xor3_u32(o << 1, o << 2, o << 7));
  r.y = xor3_u32(xor3_u32(l1, h1, __builtin_amdgcn_alignbit(h1, h0, 31)), __builtin_amdgcn_alignbit(h1, h0, 30),
                 __builtin_amdgcn_alignbit(h1, h0, 25));
  r.z = xor3_u32(xor3_u32(l2, h2, __builtin_amdgcn_alignbit(h2, h1, 31)), __builtin_amdgcn_alignbit(h2, h1, 30),
                 __builtin_amdgcn_alignbit(h2, h1, 25));
  r.w = xor3_u32(xor3_u32(l3, h3, __builtin_amdgcn_alignbit(h3, h2, 31)), __builtin_amdgcn_alignbit(h3, h2, 30),
                 __builtin_amdgcn_alignbit(h3, h2, 25));
  return r;
}

// General product a * b with both operands varying per lane (comb method).
// Each lane owns a 16-entry table T[v] = v(x) * a mod P in LDS; then for
// nibble position j = 7..0 of the four 32-bit words b_i of b, the entries
// T[nibble_j(b_i)] are XORed into a 256-bit accumulator at dword offset i,
// and the accumulator shifts by 4 between positions: 32 lookups, 7 wide
// shifts and one final reduction.  No cross-lane traffic: a lane only reads
// back what it wrote, so no barrier is needed.
//
// Tables of a workgroup's wavefronts share one area mt (16 KiB each).
// Entry v of lane L in wavefront w sits at byte
//   w * 16384 + (L >> 4) * 4096 + v * 256 + (L & 15) * 16,
// so the 16 lanes the LDS serves together hit 16 distinct bank groups, and
// a lookup address is one v_perm: byte 0 = (L & 15) * 16, byte 1 = the
// nibble with (L >> 4) and (w & 3) in its high bits, byte 2 = w >> 2.
struct CombLane {
  uint32_t lo;    // bytes [(L & 15) * 16, 0, w >> 2, 0]
  uint32_t hb;    // (L >> 4) << 4 | (w & 3) << 6 in every byte
  uint32_t base;  // byte offset of the lane's entry 0
  __device__ __forceinline__ explicit CombLane(uint32_t tid) {
    const uint32_t L = tid & 63u, w = tid >> 6;
    lo = ((L & 15u) << 4) | ((w >> 2) << 16);
    hb = ((((L >> 4) & 3u) << 4) | ((w & 3u) << 6)) * 0x01010101u;
    base = (w << 14) | ((L >> 4) << 12) | ((L & 15u) << 4);
  }
};

// The lane's table of a (entries v * 256 bytes apart from its base).
__device__ __forceinline__ void f128_mul_build(f128 a, f128 *__restrict__ mt, const CombLane &C) {
  f128 t1 = a, t2 = f128_mulx(a), t4 = f128_mulx(t2), t8 = f128_mulx(t4);
  f128 t3 = f128_xor(t1, t2), t5 = f128_xor(t4, t1), t6 = f128_xor(t4, t2), t7 = f128_xor(t4, t3);
  char *T = (char *)mt + C.base;
  const f128 v[16] = {f128_zero(),       t1,
                      t2,                t3,
                      t4,                t5,
                      t6,                t7,
                      t8,                f128_xor(t8, t1),
                      f128_xor(t8, t2),  f128_xor(t8, t3),
                      f128_xor(t8, t4),  f128_xor(t8, t5),
                      f128_xor(t8, t6),  f128_xor(t8, t7)};
#pragma unroll
  for (int e = 0; e < 16; ++e) *(f128 *)(T + e * 256) = v[e];
}

// The comb's 256-bit product of b by the lane's table (f128_mul_build), in
// d[0..7] (its reduction is the product).
__device__ __forceinline__ void f128_comb256(f128 b, const f128 *__restrict__ mt, const CombLane &C, uint32_t (&d)[8]) {
  const char *T = (const char *)mt;
  // even / odd nibbles of each word, one per byte, with the lane's block bits
  const uint32_t bw[4] = {b.x, b.y, b.z, b.w};
  uint32_t ev[4], od[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    ev[i] = (bw[i] & 0x0F0F0F0Fu) | C.hb;
    od[i] = ((bw[i] >> 4) & 0x0F0F0F0Fu) | C.hb;
  }
  uint32_t d0 = 0, d1 = 0, d2 = 0, d3 = 0, d4 = 0, d5 = 0, d6 = 0, d7 = 0;
#pragma unroll
  for (int j = 7; j >= 0; --j) {
    // nibble j of word i: byte j / 2 of od (j odd) or ev (j even)
    const uint32_t sel = 4u | ((uint32_t)(j >> 1) << 8) | (6u << 16) | (12u << 24);
    const f128 e0 = *(const f128 *)(T + __builtin_amdgcn_perm(C.lo, (j & 1) ? od[0] : ev[0], sel));
    const f128 e1 = *(const f128 *)(T + __builtin_amdgcn_perm(C.lo, (j & 1) ? od[1] : ev[1], sel));
    const f128 e2 = *(const f128 *)(T + __builtin_amdgcn_perm(C.lo, (j & 1) ? od[2] : ev[2], sel));
    const f128 e3 = *(const f128 *)(T + __builtin_amdgcn_perm(C.lo, (j & 1) ? od[3] : ev[3], sel));
    d0 ^= e0.x;
    d1 = xor3_u32(d1, e0.y, e1.x);
    d2 = xor3_u32(xor3_u32(d2, e0.z, e1.y), e2.x, 0u);
    d3 = xor3_u32(xor3_u32(d3, e0.w, e1.z), e2.y, e3.x);
    d4 = xor3_u32(xor3_u32(d4, e1.w, e2.z), e3.y, 0u);
    d5 = xor3_u32(d5, e2.w, e3.z);
    d6 ^= e3.w;
    if (j) {
      d7 = __builtin_amdgcn_alignbit(d7, d6, 28);
      d6 = __builtin_amdgcn_alignbit(d6, d5, 28);
      d5 = __builtin_amdgcn_alignbit(d5, d4, 28);
      d4 = __builtin_amdgcn_alignbit(d4, d3, 28);
      d3 = __builtin_amdgcn_alignbit(d3, d2, 28);
      d2 = __builtin_amdgcn_alignbit(d2, d1, 28);
      d1 = __builtin_amdgcn_alignbit(d1, d0, 28);
      d0 <<= 4;
    }
  }
  d[0] = d0;
  d[1] = d1;
  d[2] = d2;
  d[3] = d3;
  d[4] = d4;
  d[5] = d5;
  d[6] = d6;
  d[7] = d7;
}

// a * b with the lane's table of a already built (f128_mul_build).
__device__ __forceinline__ f128 f128_mul_comb(f128 b, const f128 *__restrict__ mt, const CombLane &C) {
  uint32_t d[8];
  f128_comb256(b, mt, C, d);
  return f128_reduce256(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
}

// A sum of products kept before reduction (reduction is linear, so the sum
// reduces once).
struct F256w {
  uint32_t d[8];
};

__device__ __forceinline__ f128 f128_reduce256w(const F256w &a) {
  return f128_reduce256(a.d[0], a.d[1], a.d[2], a.d[3], a.d[4], a.d[5], a.d[6], a.d[7]);
}

// acc ^= a * b before reduction, with the lane's table of a already built.
__device__ __forceinline__ void f128_mul_comb_acc(f128 b, const f128 *__restrict__ mt, const CombLane &C, F256w &acc) {
  uint32_t d[8];
  f128_comb256(b, mt, C, d);
#pragma unroll
  for (int i = 0; i < 8; ++i) acc.d[i] ^= d[i];
}

__device__ __forceinline__ f128 f128_mul(f128 a, f128 b, f128 *__restrict__ mt, const CombLane &C) {
  f128_mul_build(a, mt, C);
  return f128_mul_comb(b, mt, C);
}

// LDS bytes a wavefront needs for f128_mul scratch.
#define F128_MUL_LDS_PER_WAVE (16 * 64 * 16)
