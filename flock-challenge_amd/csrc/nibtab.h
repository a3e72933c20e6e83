// Products by a wavefront-uniform constant t through nibble tables in LDS.
//
// T_k[n] = (n x^(4k)) t for the 32 nibble positions k of the other operand and
// n = 0..15, so v t = sum_k T_k[nibble_k(v)]: 32 independent lookups, no
// shifts.  A table is 512 entries (8 KiB), small enough that one wavefront
// builds its own from t in about a hundred instructions per lane: the right
// size for a constant that meets only 64..1024 products.
//
// Layout: entry (n, k) at byte (k >> 4) * 4096 + n * 256 + (k & 15) * 16, so
// its 16-byte bank group is k mod 16.  At step s (0..31) lane L reads nibble
// position k = s ^ (L mod 16): the 16 lanes the LDS serves together touch
// distinct bank groups whatever the data.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fulltab.h"
#include "gf128_dev.h"

namespace flock {

constexpr uint32_t NIBTAB_BYTES = 8192;

struct NibtabLane {
  uint32_t dsel;        // dword d ^ 1 or d (rot bit 3)
  uint32_t sel_e, sel_o;  // from (odd, even) nibble words: half rot&1, byte j ^ ((rot >> 1) & 3)
  uint32_t off[4];      // byte i of off[q] = ((4 q + i) ^ rot) * 16
  __device__ __forceinline__ explicit NibtabLane(uint32_t rot) {
    const uint32_t r = rot & 15u, r12 = (r >> 1) & 3u;
    dsel = (r & 8u) ? 0x07060504u : 0x03020100u;
    const uint32_t se = (r & 1u) ? 4u : 0u, so = 4u - se;
    sel_e = 0;
    sel_o = 0;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      sel_e |= (se + (j ^ r12)) << (8 * j);
      sel_o |= (so + (j ^ r12)) << (8 * j);
    }
#pragma unroll
    for (uint32_t q = 0; q < 4; ++q) {
      off[q] = 0;
#pragma unroll
      for (uint32_t i = 0; i < 4; ++i) off[q] |= (((4u * q + i) ^ r) << 4) << (8 * i);
    }
  }
};

// v_perm selector: byte 0 = offset byte i of off[q], byte 1 = data byte j.
__device__ __forceinline__ constexpr uint32_t nibtab_sel(int i, int j) {
  return (4u + (uint32_t)i) | ((uint32_t)j << 8) | (12u << 16) | (12u << 24);
}

// s + v * t from the nibble tables of t at byte offset TAB of lds (the sum
// starts at s: a butterfly's XOR costs nothing).
template <uint32_t TAB>
__device__ __forceinline__ f128 f128_mul_nibtab_add(f128 v, const char *__restrict__ lds, const NibtabLane &L,
                                                    f128 s) {
  // dword d of w = dword d ^ (rot bit 3) of v
  const uint32_t w0 = __builtin_amdgcn_perm(v.y, v.x, L.dsel), w1 = __builtin_amdgcn_perm(v.x, v.y, L.dsel);
  const uint32_t w2 = __builtin_amdgcn_perm(v.w, v.z, L.dsel), w3 = __builtin_amdgcn_perm(v.z, v.w, L.dsel);
  const uint32_t w[4] = {w0, w1, w2, w3};
  // E[d] byte j = nibble (half rot&1) of byte j ^ r12 of w[d]; O[d] the other half
  uint32_t E[4], O[4];
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const uint32_t ev = w[d] & 0x0F0F0F0Fu, od = (w[d] >> 4) & 0x0F0F0F0Fu;
    E[d] = __builtin_amdgcn_perm(od, ev, L.sel_e);
    O[d] = __builtin_amdgcn_perm(od, ev, L.sel_o);
  }
  f128 acc = s;
#pragma unroll
  for (int g = 0; g < 8; ++g) {
    f128 t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int s = 4 * g + u;  // nibble k = s ^ rot: half s&1, byte (s>>1)&3, dword s>>3
      const uint32_t word = (s & 1) ? O[s >> 3] : E[s >> 3];
      const uint32_t a = __builtin_amdgcn_perm(L.off[(s & 15) >> 2], word, nibtab_sel(s & 3, (s >> 1) & 3));
      t[u] = *(const f128 *)(lds + TAB + (uint32_t)(s >> 4) * 4096u + a);
    }
    acc = f128_xor3(acc, t[0], t[1]);
    acc = f128_xor3(acc, t[2], t[3]);
    __builtin_amdgcn_sched_barrier(0);
  }
  return acc;
}

// v * t from the nibble tables of t at byte offset TAB of lds.
template <uint32_t TAB>
__device__ __forceinline__ f128 f128_mul_nibtab(f128 v, const char *__restrict__ lds, const NibtabLane &L) {
  return f128_mul_nibtab_add<TAB>(v, lds, L, f128_zero());
}

// The 512 entries of t's nibble tables at byte offset TAB, built by one
// wavefront (lane = k + 32 h: position k, entries n = 8 h .. 8 h + 7).
template <uint32_t TAB>
__device__ __forceinline__ void build_nibtab(f128 t, char *__restrict__ lds, uint32_t lane) {
  const uint32_t k = lane & 31u, h = lane >> 5;
  const f128 b0 = f128_mul_xpow(t, 4u * k), b1 = f128_mulx(b0), b2 = f128_mulx(b1);
  const f128 b3 = f128_mulx(b2);
  const uint32_t hm = 0u - h;
  f128 e = make_uint4(b3.x & hm, b3.y & hm, b3.z & hm, b3.w & hm);  // entry n = 8 h
  char *dst = lds + TAB + (k >> 4) * 4096u + (k & 15u) * 16u + h * 8u * 256u;
  *(f128 *)dst = e;
  const f128 low[3] = {b0, b1, b2};
#pragma unroll
  for (int g = 1; g < 8; ++g) {
    e = f128_xor(e, low[__builtin_ctz(g)]);
    *(f128 *)(dst + (uint32_t)(g ^ (g >> 1)) * 256u) = e;
  }
}

// The same 512 entries built by a workgroup of 256 threads: thread
// (k = tid mod 32, g = tid / 32) writes entries n = 2g and 2g + 1 of
// position k (the 16 lanes of a store hit distinct bank groups).
template <uint32_t TAB>
__device__ __forceinline__ void build_nibtab256(f128 t, char *__restrict__ lds, uint32_t tid) {
  const uint32_t k = tid & 31u, g = tid >> 5;
  const f128 b0 = f128_mul_xpow(t, 4u * k), b1 = f128_mulx(b0), b2 = f128_mulx(b1), b3 = f128_mulx(b2);
  const uint32_t m1 = 0u - (g & 1u), m2 = 0u - ((g >> 1) & 1u), m3 = 0u - ((g >> 2) & 1u);
  f128 e;  // entry 2g = g0 b1 + g1 b2 + g2 b3
  e.x = xor3_u32(b1.x & m1, b2.x & m2, b3.x & m3);
  e.y = xor3_u32(b1.y & m1, b2.y & m2, b3.y & m3);
  e.z = xor3_u32(b1.z & m1, b2.z & m2, b3.z & m3);
  e.w = xor3_u32(b1.w & m1, b2.w & m2, b3.w & m3);
  char *dst = lds + TAB + (k >> 4) * 4096u + (k & 15u) * 16u + 2u * g * 256u;
  *(f128 *)dst = e;
  *(f128 *)(dst + 256) = f128_xor(e, b0);
}

}  // namespace flock
