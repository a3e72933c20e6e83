// Products by a workgroup-uniform constant t through full byte tables in LDS.
//
// T_k[b] = (b x^(8k)) t for the 16 byte positions k of the other operand, so
// v t = sum_k T_k[byte_k(v)]: 16 independent lookups, no shifts.  The tables
// (4096 entries, 64 KiB) are built by the whole workgroup from the 128 basis
// values t x^i and shared by every lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gf128_dev.h"

namespace flock {

// Conflict-free lookups.  The 16 byte tables are interleaved: entry (b, k)
// = T_k[b] sits at byte offset b * 256 + k * 16, so its 16-byte bank group is
// k mod 16.  At step (q, i) lane L reads table k = (4 q + i) ^ (L mod 16), so
// the 16 lanes the LDS serves together always touch distinct bank groups
// whatever the data.  The lane first permutes the bytes of v the same way
// (byte 4 q + i of x = byte (4 q + i) ^ rot of v).
//
// Every lookup address is one v_perm: byte 0 = k * 16 from a per-lane offset
// register, byte 1 = the data byte, byte 2 = 0 or 1 (a second table 64 KiB
// above the first, for double buffering).
struct FulltabLane {
  uint32_t sel_hi;  // per-lane v_perm selector: dword q ^ 2 or q (rot bit 3)
  uint32_t sel_lo;  // dword q ^ 1 or q (rot bit 2), bytes i ^ (rot & 3)
  uint32_t o01, o23;  // offsets of q = 0: bytes [off(i), off(i + 1), 1, 0]
  __device__ __forceinline__ explicit FulltabLane(uint32_t rot) {
    sel_hi = (rot & 8u) ? 0x07060504u : 0x03020100u;
    const uint32_t src = (rot & 4u) ? 4u : 0u, r3 = rot & 3u;
    sel_lo = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) sel_lo |= (src + (j ^ r3)) << (8 * j);
    const uint32_t r = rot & 15u;
    o01 = ((0u ^ r) << 4) | (((1u ^ r) << 4) << 8) | (1u << 16);
    o23 = ((2u ^ r) << 4) | (((3u ^ r) << 4) << 8) | (1u << 16);
  }
  // byte 4 q + i of the result = byte (4 q + i) ^ rot of v
  __device__ __forceinline__ f128 xperm(f128 v) const {
    const uint32_t a0 = __builtin_amdgcn_perm(v.z, v.x, sel_hi), a1 = __builtin_amdgcn_perm(v.w, v.y, sel_hi);
    const uint32_t a2 = __builtin_amdgcn_perm(v.x, v.z, sel_hi), a3 = __builtin_amdgcn_perm(v.y, v.w, sel_hi);
    f128 o;
    o.x = __builtin_amdgcn_perm(a1, a0, sel_lo);
    o.y = __builtin_amdgcn_perm(a0, a1, sel_lo);
    o.z = __builtin_amdgcn_perm(a3, a2, sel_lo);
    o.w = __builtin_amdgcn_perm(a2, a3, sel_lo);
    return o;
  }
};

// v_perm selector of lookup i: byte 0 = offset byte (i & 1) of the pair
// register, byte 1 = data byte i, byte 2 = the pair register's 1 (HI) or 0.
__device__ __forceinline__ constexpr uint32_t fulltab_sel(int i, bool hi) {
  return (4u + (uint32_t)(i & 1)) | ((uint32_t)i << 8) | ((hi ? 6u : 12u) << 16) | (12u << 24);
}

__device__ __forceinline__ uint32_t f128_word(const f128 &x, int q) {
  return q == 0 ? x.x : q == 1 ? x.y : q == 2 ? x.z : x.w;
}

// s + v * t from the interleaved tables of t at base (+ 64 KiB when HI).
// The four steps are unrolled with scheduling barriers between them, so one
// step's four lookups are in flight at a time and nothing is hoisted.  The
// sum starts at s (a butterfly's XOR costs nothing).
template <bool HI = false>
__device__ __forceinline__ f128 f128_mul_fulltab_add(f128 v, const char *__restrict__ base, const FulltabLane &L,
                                                     f128 s) {
  const f128 x = L.xperm(v);
  f128 acc = s;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t w = f128_word(x, q);
    const uint32_t oa = L.o01 ^ (0x4040u * q), ob = L.o23 ^ (0x4040u * q);
    f128 t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      t[i] = *(const f128 *)(base + __builtin_amdgcn_perm(i < 2 ? oa : ob, w, fulltab_sel(i, HI)));
    acc = f128_xor3(acc, t[0], t[1]);
    acc = f128_xor3(acc, t[2], t[3]);
    __builtin_amdgcn_sched_barrier(0);
  }
  return acc;
}

// v * t.
template <bool HI = false>
__device__ __forceinline__ f128 f128_mul_fulltab(f128 v, const char *__restrict__ base, const FulltabLane &L) {
  return f128_mul_fulltab_add<HI>(v, base, L, f128_zero());
}

// Two independent products added into p0 and p1, interleaved for latency:
// v0 by the table at base (+ 64 KiB when HI), v1 by the one HI1 selects (the
// same by default).
template <bool HI = false, bool HI1 = HI>
__device__ __forceinline__ void f128_mul2_fulltab_add(f128 v0, f128 v1, const char *__restrict__ base,
                                                      const FulltabLane &L, f128 &p0, f128 &p1) {
  const f128 x0 = L.xperm(v0), x1 = L.xperm(v1);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t w0 = f128_word(x0, q), w1 = f128_word(x1, q);
    const uint32_t oa = L.o01 ^ (0x4040u * q), ob = L.o23 ^ (0x4040u * q);
    f128 t0[4], t1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t o = i < 2 ? oa : ob;
      t0[i] = *(const f128 *)(base + __builtin_amdgcn_perm(o, w0, fulltab_sel(i, HI)));
      t1[i] = *(const f128 *)(base + __builtin_amdgcn_perm(o, w1, fulltab_sel(i, HI1)));
    }
    p0 = f128_xor3(p0, t0[0], t0[1]);
    p1 = f128_xor3(p1, t1[0], t1[1]);
    p0 = f128_xor3(p0, t0[2], t0[3]);
    p1 = f128_xor3(p1, t1[2], t1[3]);
    __builtin_amdgcn_sched_barrier(0);
  }
}

// p0 = v0 * t, p1 = v1 * t' (as f128_mul2_fulltab_add from zero).
template <bool HI = false, bool HI1 = HI>
__device__ __forceinline__ void f128_mul2_fulltab(f128 v0, f128 v1, const char *__restrict__ base,
                                                  const FulltabLane &L, f128 &p0, f128 &p1) {
  p0 = f128_zero();
  p1 = f128_zero();
  f128_mul2_fulltab_add<HI, HI1>(v0, v1, base, L, p0, p1);
}

// t * x^i mod P for 0 <= i < 128 (i may differ per lane): the dword shift
// by i / 32 through lane masks and the bit shift by i % 32 through 64-bit
// shifts, so no lane-dependent branch.
__device__ __forceinline__ f128 f128_mul_xpow(f128 t, uint32_t i) {
  const uint32_t q = i >> 5, r = i & 31;
  uint32_t m[4];
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) m[j] = 0u - (uint32_t)(q == j);
  const uint32_t tw[4] = {t.x, t.y, t.z, t.w};
  uint32_t d[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t acc = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (k - j >= 0 && k - j <= 3) acc |= tw[k - j] & m[j];
    d[k] = acc;
  }
  uint32_t e[8];
  e[0] = d[0] << r;
#pragma unroll
  for (int k = 1; k < 8; ++k) e[k] = (uint32_t)(((((uint64_t)d[k]) << 32 | d[k - 1]) << r) >> 32);
  return f128_reduce256(e[0], e[1], e[2], e[3], e[4], e[5], e[6], e[7]);
}

// Full byte tables of one twiddle from its 128 basis values, E = 4096 /
// THREADS entries per thread.  basis[i * 16 + k] = t x^(8k + i); thread
// (k = tid & 15, u = tid >> 4) writes entries b = E u .. E u + E - 1 of
// table k (byte offset b * 256 + k * 16 from tab), so the lanes of one store
// hit distinct bank groups.  The E entries differ in their low bits only and
// are formed in registers in Gray-code order.
template <int THREADS>
__device__ __forceinline__ void build_fulltab(const f128 *__restrict__ basis, f128 *__restrict__ tab, uint32_t tid) {
  constexpr int E = 4096 / THREADS;
  constexpr int LE = E == 4 ? 2 : E == 8 ? 3 : E == 16 ? 4 : -1;
  static_assert(LE > 0, "4, 8 or 16 entries per thread");
  const uint32_t kk = tid & 15, u = tid >> 4;
  f128 low[LE];
#pragma unroll
  for (int i = 0; i < LE; ++i) low[i] = basis[i * 16 + kk];
  f128 h = f128_zero();
#pragma unroll
  for (int i = LE; i < 8; ++i) {
    const f128 bi = basis[i * 16 + kk];
    const uint32_t m = 0u - ((u >> (i - LE)) & 1u);
    h.x ^= bi.x & m;
    h.y ^= bi.y & m;
    h.z ^= bi.z & m;
    h.w ^= bi.w & m;
  }
  f128 *dst = tab + (E * u) * 16 + kk;  // entry (b, kk) at b * 16 + kk
  dst[0] = h;
#pragma unroll
  for (int g = 1; g < E; ++g) {
    // Gray code: entry g ^ (g >> 1) differs from the previous in one bit
    const int bit = __builtin_ctz(g);
    h = f128_xor(h, low[bit]);
    dst[(g ^ (g >> 1)) * 16] = h;
  }
}


// basis[i * 16 + kb] = t x^(8 kb + i) for one constant (threads 0..127).
__device__ __forceinline__ void fulltab_basis(f128 t, f128 *__restrict__ basis, uint32_t tid) {
  if (tid < 128) {
    const uint32_t i = (tid >> 4) & 7, kb = tid & 15;
    basis[tid] = f128_mul_xpow(t, 8 * kb + i);
  }
}

}  // namespace flock
