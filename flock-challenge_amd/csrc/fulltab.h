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
// = T_k[b] sits at f128 index b * 16 + k, so its 16-byte bank group is k mod
// 16 (k mod 8 with 32 banks).  At step q lane L reads table k = q ^ (L mod
// 16), so the 16 (or 8) lanes the LDS serves together always touch distinct
// bank groups whatever the data.  The lane first permutes the bytes of v the
// same way (byte q of the result = byte q ^ rot of v).
__device__ __forceinline__ f128 f128_xperm_bytes(f128 v, uint32_t r) {
  uint32_t w0 = v.x, w1 = v.y, w2 = v.z, w3 = v.w;
  if (r & 8) {
    uint32_t t = w0; w0 = w2; w2 = t;
    t = w1; w1 = w3; w3 = t;
  }
  if (r & 4) {
    uint32_t t = w0; w0 = w1; w1 = t;
    t = w2; w2 = w3; w3 = t;
  }
  // within a dword: bit 1 swaps the halves, bit 0 swaps bytes inside halves
  const uint32_t sel = (r & 3) == 0 ? 0x03020100u : (r & 3) == 1 ? 0x02030001u : (r & 3) == 2 ? 0x01000302u
                                                                                            : 0x00010203u;
  f128 o;
  o.x = __builtin_amdgcn_perm(0, w0, sel);
  o.y = __builtin_amdgcn_perm(0, w1, sel);
  o.z = __builtin_amdgcn_perm(0, w2, sel);
  o.w = __builtin_amdgcn_perm(0, w3, sel);
  return o;
}

// Per-lane byte offsets of table k = (4 q + i) ^ rot inside one b row:
// off(q, i) = ((4 q) ^ (rot & 12)) * 16 + (i ^ (rot & 3)) * 16.
struct FulltabLane {
  uint32_t lo[4];  // (i ^ (rot & 3)) * 16
  uint32_t hi;     // (rot & 12) * 16
  __device__ __forceinline__ explicit FulltabLane(uint32_t rot) {
#pragma unroll
    for (uint32_t i = 0; i < 4; ++i) lo[i] = (i ^ (rot & 3u)) << 4;
    hi = (rot & 12u) << 4;
  }
};

// v * t from the interleaved tables of t (byte addresses: tab + (b << 8) + k * 16).
__device__ __forceinline__ f128 f128_mul_fulltab(f128 v, const char *__restrict__ tab, const FulltabLane &L,
                                                 uint32_t rot) {
  const f128 x = f128_xperm_bytes(v, rot);
  f128 acc = f128_zero();
  uint32_t a = x.x, b = x.y, c = x.z, d = x.w;
  // A real loop (not unrolled) keeps one word's four lookups in flight at a
  // time; unrolled, the compiler hoists every lookup of the pass and spills.
#pragma unroll 1
  for (int q = 0; q < 4; ++q) {
    const uint32_t qoff = L.hi ^ (64u * q);
    f128 t[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) t[i] = *(const f128 *)(tab + ((((a >> (8 * i)) & 255u) << 8) | qoff | L.lo[i]));
    acc = f128_xor3(acc, t[0], t[1]);
    acc = f128_xor3(acc, t[2], t[3]);
    a = b;
    b = c;
    c = d;
  }
  return acc;
}

// Two independent products by the same table, interleaved for latency.
__device__ __forceinline__ void f128_mul2_fulltab(f128 v0, f128 v1, const char *__restrict__ tab,
                                                  const FulltabLane &L, uint32_t rot, f128 &p0, f128 &p1) {
  const f128 x0 = f128_xperm_bytes(v0, rot), x1 = f128_xperm_bytes(v1, rot);
  p0 = f128_zero();
  p1 = f128_zero();
  uint32_t a0 = x0.x, b0 = x0.y, c0 = x0.z, d0 = x0.w;
  uint32_t a1 = x1.x, b1 = x1.y, c1 = x1.z, d1 = x1.w;
#pragma unroll 1
  for (int q = 0; q < 4; ++q) {
    const uint32_t qoff = L.hi ^ (64u * q);
    f128 t0[4], t1[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint32_t o = qoff | L.lo[i];
      t0[i] = *(const f128 *)(tab + ((((a0 >> (8 * i)) & 255u) << 8) | o));
      t1[i] = *(const f128 *)(tab + ((((a1 >> (8 * i)) & 255u) << 8) | o));
    }
    p0 = f128_xor3(p0, t0[0], t0[1]);
    p1 = f128_xor3(p1, t1[0], t1[1]);
    p0 = f128_xor3(p0, t0[2], t0[3]);
    p1 = f128_xor3(p1, t1[2], t1[3]);
    a0 = b0;
    b0 = c0;
    c0 = d0;
    a1 = b1;
    b1 = c1;
    c1 = d1;
  }
}

// t * x^i mod P for 0 <= i < 128.
__device__ __forceinline__ uint32_t f128_dword_or_zero(f128 t, int idx) {
  return idx == 0 ? t.x : idx == 1 ? t.y : idx == 2 ? t.z : idx == 3 ? t.w : 0u;
}
__device__ __forceinline__ f128 f128_mul_xpow(f128 t, uint32_t i) {
  const int q = (int)(i >> 5);
  const uint32_t r = i & 31;
  uint32_t d[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    // dword k of (t << i) takes bits from dwords k - q and k - q - 1 of t
    const uint32_t hi = f128_dword_or_zero(t, k - q), lo = f128_dword_or_zero(t, k - q - 1);
    d[k] = r ? __builtin_amdgcn_alignbit(hi, lo, 32 - r) : hi;
  }
  return f128_reduce256(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7]);
}

// Full byte tables of one twiddle from its 128 basis values, E = 4096 /
// THREADS entries per thread.  basis[i * 16 + k] = t x^(8k + i); thread
// (k = tid & 15, u = tid >> 4) writes entries b = E u .. E u + E - 1 of
// table k, so the lanes of one store hit distinct bank groups.
template <int THREADS>
__device__ __forceinline__ void build_fulltab(const f128 *__restrict__ basis, f128 *__restrict__ tab, uint32_t tid) {
  constexpr int E = 4096 / THREADS;
  constexpr int LE = E == 4 ? 2 : E == 8 ? 3 : E == 16 ? 4 : -1;
  static_assert(LE > 0, "4, 8 or 16 entries per thread");
  const uint32_t kk = tid & 15, u = tid >> 4;
  f128 h = f128_zero();
#pragma unroll 1
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
  for (int lo = 1; lo < E; ++lo) {
    // entry lo = entry (lo with its lowest set bit cleared) + that bit's basis
    const int bit = __builtin_ctz(lo);
    h = f128_xor(dst[(lo & (lo - 1)) * 16], basis[bit * 16 + kk]);
    dst[lo * 16] = h;
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
