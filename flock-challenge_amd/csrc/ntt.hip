// Interleaved additive NTT of the PCS commit on gfx950 (64 RS lanes).
//
// Layout in HBM (the reference's, crates/flock-core/src/pcs/commit.rs:235-245):
// codeword[pos * 64 + lane] F128.
//
// A pass applies F consecutive layers [l, l+F) of the neighbors-last forward
// transform (ntt/additive_ntt_f128.rs:1281-1320).  The 2^F positions that
// exchange data in those layers share every position bit except F contiguous
// ones, so they form a "group"
//     pos(e) = (B << (k-l)) | (e << s) | r,   s = k - l - F,  e in [0, 2^F).
// One wavefront owns one group: lane = RS lane, and each lane keeps its 2^F
// elements in VGPRs, so every butterfly is lane-local and the global loads
// and stores are 1 KiB-contiguous per position (fully coalesced).  The
// twiddle of a butterfly depends only on (layer, block), i.e. it is
// wavefront-uniform: its products run through lookup tables of that twiddle
// in LDS, full byte tables shared by a workgroup (fulltab.h) where the
// twiddle meets many products, per-wavefront nibble tables (nibtab.h) in the
// last layers.
//
// Several kernels can run a pass (NttKernel); which ones, and how the layers
// are split into passes, is a plan per transform shape (k, start layer).
// ntt_autotune times every feasible (layer, F, kernel) on the device and
// picks the fastest partition by dynamic programming; tuned plans are kept
// in a process-wide table, loaded at first use from a JSON file beside the
// library (tune/ntt_plans_gfx950.json), and any valid plan gives the same
// codeword.
#include <hip/hip_runtime.h>
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/flock_gpu.h"
#include "commit.h"
#include "ctx.h"
#include "fulltab.h"
#include "fused_job.h"
#include "gf128_dev.h"
#include "gf128_short.h"
#include "nibtab.h"

namespace flock {

// The workgroup a pass launch's hardware index w stands for.  A first pass
// (reading the message, src_mask != ~0) starts at layer l = log_inv_rate, so
// the 2^l blocks B of its first layer are the 2^l replicas of every message
// position, and B is the top l bits of the workgroup index.  The mapping
// below gives the replicas of the same positions to consecutive workgroups
// of one XCD (hardware indices 8 apart: dispatch round-robins over the 8
// XCDs, each with its own L2), so all but the first read of a message row
// hit that L2.  A bijection on [0, n_wgs) (a power of two); any other pass
// keeps w.
__device__ __forceinline__ uint64_t pass_wg(uint64_t w, uint64_t n_wgs, int l, uint64_t src_mask) {
  if (src_mask == ~0ull || l == 0 || n_wgs < (8ull << l)) return w;
  const uint64_t xcd = w & 7, t = w >> 3;
  const uint64_t b = t & ((1ull << l) - 1), u = t >> l;
  return b * (n_wgs >> l) + (u << 3) + xcd;
}

// ---------------------------------------------------------------- tail passes
// Layers whose twiddles are shared by too few groups to pay for full byte
// tables (the last layers of the transform, where a twiddle meets 64 .. 512
// products).  One wavefront per group; it builds the 8 KiB nibble tables of
// its own twiddles, NB at a time, in NB * 8 KiB of LDS (NB = 2: two products
// by different twiddles in flight; NB = 1: twice the wavefronts per CU).
// With SJ = 0 (no SJ < F is instantiated: mixed with table layers in one
// kernel, the compiler hoists all of a table product's lookups ahead of its
// XORs and runs out of registers) every layer takes shifted XORs instead
// (gf128_short.h): no table, no LDS and no barrier, at a cost proportional
// to the twiddle's set bits, which pays in the transform's deepest layers
// (twiddles of at most d and 2d - 3 bits in the last two layers of a
// transform of 2^d positions).  Any twiddle gives the right product either
// way; the plan decides where each variant runs.
// Step (J, C) handles twiddles C .. C + NT - 1 of layer offset J; template
// recursion keeps every register index constant.

// s + v t for twiddle t = n of a group's pass, in scalar registers, from
// the vector in which lane n holds twiddle n (one load for all of them, 4
// VGPRs).
__device__ __forceinline__ f128 f128_mul_shifts_add(f128 v, f128 lanes, int n, f128 s) {
  const uint32_t vw[4] = {v.x, v.y, v.z, v.w}, sw[4] = {s.x, s.y, s.z, s.w};
  uint32_t o[4];
  shortmul::mul_add(vw, __builtin_amdgcn_readlane(lanes.x, n), __builtin_amdgcn_readlane(lanes.y, n),
                    __builtin_amdgcn_readlane(lanes.z, n), __builtin_amdgcn_readlane(lanes.w, n), sw, o);
  return make_uint4(o[0], o[1], o[2], o[3]);
}

template <int F, int NB, int SJ, int J, int C>
__device__ __forceinline__ void tail_pass_step(f128 (&v)[1 << F], const f128 *__restrict__ tw, int l, uint64_t B,
                                               f128 twl, char *lds, const NibtabLane &NL, uint32_t lane) {
  constexpr int nb = 1 << (F - 1 - J);
  constexpr int NT = (J < SJ && NB == 2 && (1 << J) - C >= 2) ? 2 : 1;
  if constexpr (J >= SJ) {
#pragma unroll
    for (int el = 0; el < nb; ++el) {
      const int e0 = (C << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
      const f128 u0 = f128_mul_shifts_add(v[e1], twl, (1 << J) - 1 + C, v[e0]);
      v[e0] = u0;
      f128_xor_into(v[e1], u0);
      __builtin_amdgcn_sched_barrier(0);
    }
  } else {
    const uint64_t tb = (1ull << (l + J)) - 1ull + (B << J);
    __syncthreads();  // the previous step's tables are read
    build_nibtab<0>(tw[tb + C], lds, lane);
    if constexpr (NT == 2) build_nibtab<NIBTAB_BYTES>(tw[tb + C + 1], lds, lane);
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int el = 0; el < nb; ++el) {
      const int e0 = (C << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
      const f128 u0 = f128_mul_nibtab_add<0>(v[e1], lds, NL, v[e0]);
      v[e0] = u0;
      f128_xor_into(v[e1], u0);
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (NT == 2) {
#pragma unroll
      for (int el = 0; el < nb; ++el) {
        const int e0 = ((C + 1) << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
        const f128 u0 = f128_mul_nibtab_add<NIBTAB_BYTES>(v[e1], lds, NL, v[e0]);
        v[e0] = u0;
        f128_xor_into(v[e1], u0);
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  if constexpr (C + NT < (1 << J))
    tail_pass_step<F, NB, SJ, J, C + NT>(v, tw, l, B, twl, lds, NL, lane);
  else if constexpr (J + 1 < F)
    tail_pass_step<F, NB, SJ, J + 1, 0>(v, tw, l, B, twl, lds, NL, lane);
}

// src is the codeword itself, or (first pass) the message, read with
// positions masked by src_mask: position p of the replicated codeword holds
// message position p & src_mask, the exact state after the replication layers
// (pcs/commit.rs:451-467).
template <int F, int NB, int SJ>
__device__ __forceinline__ void ntt_tail_body(f128 *data, const f128 *__restrict__ tw, int k, int l, const f128 *src,
                                              uint64_t src_mask, uint64_t g, char *lds) {
  const uint32_t lane = threadIdx.x;
  const NibtabLane NL(lane & 15u);
  const int s = k - l - F;
  const uint64_t r = g & ((1ull << s) - 1ull);
  const uint64_t B = g >> s;
  const uint64_t base = (B << (k - l)) | r;
  f128 v[1 << F];
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) v[e] = src[(((base | ((uint64_t)e << s)) & src_mask) << 6) + lane];
  // lane n = 2^j - 1 + c < 2^F - 1 with j >= SJ loads twiddle n (layer l + j,
  // block (B << j) + c) for the shifted-XOR layers
  f128 twl = f128_zero();
  if constexpr (SJ < F) {
    if (lane >= (1u << SJ) - 1u && lane < (1u << F) - 1u) {
      const uint32_t j = 31 - __builtin_clz(lane + 1), c = lane + 1 - (1u << j);
      twl = tw[(1ull << (l + j)) - 1ull + (B << j) + c];
    }
  }
  tail_pass_step<F, NB, SJ, 0, 0>(v, tw, l, B, twl, lds, NL, lane);
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) data[((base | ((uint64_t)e << s)) << 6) + lane] = v[e];
}

template <int F, int NB, int SJ>
__global__ __launch_bounds__(64) void ntt_tail_kernel(f128 *data, const f128 *__restrict__ tw, int k, int l,
                                                      const f128 *src, uint64_t src_mask) {
  __shared__ __align__(16) char lds[SJ > 0 ? NB * NIBTAB_BYTES : 16];
  ntt_tail_body<F, NB, SJ>(data, tw, k, l, src, src_mask, pass_wg(blockIdx.x, gridDim.x, l, src_mask), lds);
}

template <int F, int NB, int SJ>
__global__ __launch_bounds__(64) void ntt_tail_batch_kernel(const NttJob *__restrict__ jobs, uint32_t n_jobs) {
  __shared__ __align__(16) char lds[SJ > 0 ? NB * NIBTAB_BYTES : 16];
  const NttJob &J = job_of(jobs, n_jobs);
  ntt_tail_body<F, NB, SJ>(J.data, J.tw, J.k, J.l, J.src ? J.src : J.data, J.src ? J.src_pos_mask : ~0ull,
                           blockIdx.x - J.wg_begin, lds);
}

// ---------------------------------------------------------------- shared-table passes
// When 2^s >= WAVES * G groups share the block B of the pass's first layer,
// they share all 2^F - 1 twiddles of the pass.  A workgroup of WAVES
// wavefronts takes WAVES * G such groups (G per wavefront).  Up front its
// threads compute the 128 values t x^i of every twiddle t of the pass; then
// for each twiddle they build the full byte tables T_k[b] = (b x^(8k)) t,
// k = 0..15 (64 KiB), and a product v t is 16 independent lookups, one per
// byte of v, with no shifts.  Two table buffers alternate, so one barrier per
// twiddle suffices: the buffer being rebuilt was last read before the
// previous barrier.
constexpr int SHARED_TWIDDLES = 15;  // 2^F - 1 for F <= 4

// The butterflies of twiddle n = 2^J - 1 + C of a pass of F layers in each
// of the wavefront's G groups, the products through full byte tables at
// tbase (+ 64 KiB when HI), two in flight per lane when PAIR.
template <int F, int G, bool PAIR, bool HI, int J, int C>
__device__ __forceinline__ void fulltab_butterflies(f128 (&v)[G][1 << F], const char *tbase, const FulltabLane &FL) {
  constexpr int nb = 1 << (F - 1 - J);
  constexpr int SLOTS = G * nb;
  constexpr int STEP = PAIR ? 2 : 1;
  // product slots (group, butterfly) in pairs
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl += STEP) {
    const int ga = sl / nb, la = sl % nb, gb = (sl + 1) / nb, lb = (sl + 1) % nb;
    const int a0 = (C << (F - J)) | la, a1 = a0 | (1 << (F - 1 - J));
    if (STEP == 2 && sl + 1 < SLOTS) {
      const int b0 = (C << (F - J)) | lb, b1 = b0 | (1 << (F - 1 - J));
      f128 u0 = v[ga][a0], u1 = v[gb][b0];
      f128_mul2_fulltab_add<HI>(v[ga][a1], v[gb][b1], tbase, FL, u0, u1);
      v[ga][a0] = u0;
      f128_xor_into(v[ga][a1], u0);
      v[gb][b0] = u1;
      f128_xor_into(v[gb][b1], u1);
    } else {
      const f128 u0 = f128_mul_fulltab_add<HI>(v[ga][a1], tbase, FL, v[ga][a0]);
      v[ga][a0] = u0;
      f128_xor_into(v[ga][a1], u0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

// The same butterflies for a zero twiddle (block 0 of a layer): u0 = v0,
// u1 = v1 + v0.
template <int F, int G, int J, int C>
__device__ __forceinline__ void zero_butterflies(f128 (&v)[G][1 << F]) {
  constexpr int nb = 1 << (F - 1 - J);
#pragma unroll
  for (int sl = 0; sl < G * nb; ++sl) {
    const int ga = sl / nb, la = sl % nb;
    const int a0 = (C << (F - J)) | la, a1 = a0 | (1 << (F - 1 - J));
    f128_xor_into(v[ga][a1], v[ga][a0]);
  }
}

// Twiddle n = 2^J - 1 + C of a shared-table pass (layer offset J, block C):
// build its tables into buffer n & 1, then every butterfly of it in each of
// the wave's G groups, two products in flight per lane (one with 16
// wavefronts per workgroup, whose 128-VGPR budget holds no more).
template <int F, int G, int THREADS, int J, int C>
__device__ __forceinline__ void shared_pass_step(f128 (&v)[G][1 << F], f128 *tabs, const f128 *basis,
                                                 const FulltabLane &FL, uint32_t tid, uint32_t zmask) {
  constexpr int N = (1 << J) - 1 + C;
  constexpr bool HI = N & 1;
  const char *tbase = (const char *)tabs;  // buffer 0 at +0, buffer 1 at +64 KiB
  if ((zmask >> N) & 1u) {
    // The barrier stays, so buffer N & 1 is not rebuilt (for twiddle N + 2)
    // while a wave still reads it for twiddle N - 1 ... N + 1.
    __syncthreads();
    zero_butterflies<F, G, J, C>(v);
  } else {
    build_fulltab<THREADS>(basis + N * 128, tabs + (HI ? 4096 : 0), tid);
    __syncthreads();
    fulltab_butterflies<F, G, (THREADS <= 512), HI, J, C>(v, tbase, FL);
  }
  if constexpr (C + 1 < (1 << J))
    shared_pass_step<F, G, THREADS, J, C + 1>(v, tabs, basis, FL, tid, zmask);
  else if constexpr (J + 1 < F)
    shared_pass_step<F, G, THREADS, J + 1, 0>(v, tabs, basis, FL, tid, zmask);
}

// Bit n set when twiddle n = 2^j - 1 + c of the pass (layer l + j, block
// (B << j) + c) is zero: block 0 of a layer, which only the workgroups of
// block B = 0 meet.  Wavefront-uniform.
__device__ __forceinline__ uint32_t pass_zero_twiddles(const f128 *__restrict__ tw, int l, int F, uint64_t B) {
  uint32_t zmask = 0;
  for (uint32_t n = 0; n < (1u << F) - 1u; ++n) {
    const uint32_t j = 31 - __builtin_clz(n + 1), c = n + 1 - (1u << j);
    const f128 t = tw[(1ull << (l + j)) - 1ull + (B << j) + c];
    zmask |= (uint32_t)(__builtin_amdgcn_readfirstlane(t.x | t.y | t.z | t.w) == 0u) << n;
  }
  return zmask;
}

constexpr int SHARED_LDS_F128 = 2 * 4096 + SHARED_TWIDDLES * 128;

template <int F, int WAVES, int G>
__device__ __forceinline__ void ntt_pass_shared_body(f128 *data, const f128 *__restrict__ tw, int k, int l,
                                                     const f128 *src, uint64_t src_mask, uint64_t wg, f128 *lds) {
  constexpr uint32_t THREADS = 64 * WAVES;
  static_assert((1 << F) - 1 <= SHARED_TWIDDLES, "twiddle basis area");
  f128 *tabs = lds;                // 2 x 16 x 256
  f128 *basis = lds + 2 * 4096;    // SHARED_TWIDDLES x 128
  const uint32_t tid = threadIdx.x, lane = tid & 63, rot = lane & 15;
  const FulltabLane FL(rot);
  const int s = k - l - F;
  // wavefront-uniform group indices (scalar registers for the addresses)
  const uint64_t g0 = (wg * WAVES + __builtin_amdgcn_readfirstlane(tid >> 6)) * G;
  const uint64_t B = g0 >> s;  // the same for the whole workgroup
  // basis of every twiddle: entry n * 128 + i * 16 + kb = t_n x^(8 kb + i),
  // t_n = twiddle n of the pass in (layer offset j, block c) order,
  // n = 2^j - 1 + c
  for (uint32_t e = tid; e < ((1u << F) - 1u) * 128u; e += THREADS) {
    const uint32_t n = e >> 7, i = (e >> 4) & 7, kb = e & 15;
    const uint32_t j = 31 - __builtin_clz(n + 1);
    const uint32_t c = n + 1 - (1u << j);
    basis[e] = f128_mul_xpow(tw[(1ull << (l + j)) - 1ull + (B << j) + c], 8 * kb + i);
  }
  f128 v[G][1 << F];
  uint64_t base[G];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const uint64_t g = g0 + gi;
    base[gi] = (B << (k - l)) | (g & ((1ull << s) - 1ull));
#pragma unroll
    for (int e = 0; e < (1 << F); ++e) v[gi][e] = src[(((base[gi] | ((uint64_t)e << s)) & src_mask) << 6) + lane];
  }
  const uint32_t zmask = B == 0 ? pass_zero_twiddles(tw, l, F, B) : 0u;
  __syncthreads();  // basis visible
  shared_pass_step<F, G, THREADS, 0, 0>(v, tabs, basis, FL, tid, zmask);
#pragma unroll
  for (int gi = 0; gi < G; ++gi)
#pragma unroll
    for (int e = 0; e < (1 << F); ++e) data[((base[gi] | ((uint64_t)e << s)) << 6) + lane] = v[gi][e];
}

template <int F, int WAVES, int G>
__global__ __launch_bounds__(64 * WAVES) void ntt_pass_shared_kernel(f128 *data, const f128 *__restrict__ tw, int k,
                                                                     int l, const f128 *src, uint64_t src_mask) {
  __shared__ f128 lds[SHARED_LDS_F128];
  ntt_pass_shared_body<F, WAVES, G>(data, tw, k, l, src, src_mask, pass_wg(blockIdx.x, gridDim.x, l, src_mask),
                                    lds);
}

template <int F, int WAVES, int G>
__global__ __launch_bounds__(64 * WAVES) void ntt_pass_shared_batch_kernel(const NttJob *__restrict__ jobs,
                                                                           uint32_t n_jobs) {
  __shared__ f128 lds[SHARED_LDS_F128];
  const NttJob &J = job_of(jobs, n_jobs);
  ntt_pass_shared_body<F, WAVES, G>(J.data, J.tw, J.k, J.l, J.src ? J.src : J.data, J.src ? J.src_pos_mask : ~0ull,
                                    blockIdx.x - J.wg_begin, lds);
}

// ---------------------------------------------------------------- single-table shared passes
// The shared-table pass with one 64 KiB table buffer and the basis of the
// current and the next twiddle only (68 KiB of LDS), so two workgroups share
// a CU: one's loads, table builds and barriers overlap the other's products.
// Two barriers per twiddle: the table is rebuilt once every wavefront is
// done with the previous one, and read once it is built.  PAIR: two
// products in flight per lane.
constexpr int SHARED1_LDS_F128 = 4096 + 2 * 128 + 16;

// Step of twiddle n = 2^J - 1 + C: its basis is in basis buffer n & 1 (made
// visible by the first barrier), twiddle n + 1's goes into the other one
// while the table is built; that buffer was last read by the previous
// step's build, before the previous step's second barrier.
template <int F, int G, int THREADS, bool PAIR, int J, int C>
__device__ __forceinline__ void shared1_pass_step(f128 (&v)[G][1 << F], f128 *tab, f128 *basis, const f128 *twv,
                                                  const FulltabLane &FL, uint32_t tid, uint32_t zmask) {
  constexpr int N = (1 << J) - 1 + C, NT = (1 << F) - 1;
  const bool zero = (zmask >> N) & 1u;
  __syncthreads();  // every wavefront done with table N - 1; basis N visible
  if (!zero) build_fulltab<THREADS>(basis + (N & 1) * 128, tab, tid);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (N + 1 < NT) fulltab_basis(twv[N + 1], basis + ((N + 1) & 1) * 128, tid);
  __builtin_amdgcn_sched_barrier(0);
  __syncthreads();  // table N and basis N + 1 visible
  if (zero)
    zero_butterflies<F, G, J, C>(v);
  else
    fulltab_butterflies<F, G, PAIR, false, J, C>(v, (const char *)tab, FL);
  if constexpr (C + 1 < (1 << J))
    shared1_pass_step<F, G, THREADS, PAIR, J, C + 1>(v, tab, basis, twv, FL, tid, zmask);
  else if constexpr (J + 1 < F)
    shared1_pass_step<F, G, THREADS, PAIR, J + 1, 0>(v, tab, basis, twv, FL, tid, zmask);
}

template <int F, int WAVES, int G, bool PAIR>
__device__ __forceinline__ void ntt_pass_shared1_body(f128 *data, const f128 *__restrict__ tw, int k, int l,
                                                      const f128 *src, uint64_t src_mask, uint64_t wg, f128 *lds) {
  constexpr uint32_t THREADS = 64 * WAVES;
  constexpr uint32_t NT = (1u << F) - 1u;
  f128 *tab = lds;                   // 16 x 256 entries
  f128 *basis = lds + 4096;          // 2 x 128
  f128 *twv = lds + 4096 + 2 * 128;  // the pass's twiddles, n = 2^j - 1 + c
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const FulltabLane FL(lane & 15u);
  const int s = k - l - F;
  const uint64_t g0 = (wg * WAVES + __builtin_amdgcn_readfirstlane(tid >> 6)) * G;
  const uint64_t B = g0 >> s;  // the same for the whole workgroup
  if (tid < NT) {
    const uint32_t j = 31 - __builtin_clz(tid + 1), c = tid + 1 - (1u << j);
    twv[tid] = tw[(1ull << (l + j)) - 1ull + (B << j) + c];
  }
  fulltab_basis(tw[(1ull << l) - 1ull + B], basis, tid);  // twiddle 0
  __builtin_amdgcn_sched_barrier(0);  // its registers free before the loads
  f128 v[G][1 << F];
  uint64_t base[G];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const uint64_t g = g0 + gi;
    base[gi] = (B << (k - l)) | (g & ((1ull << s) - 1ull));
#pragma unroll
    for (int e = 0; e < (1 << F); ++e) v[gi][e] = src[(((base[gi] | ((uint64_t)e << s)) & src_mask) << 6) + lane];
  }
  const uint32_t zmask = B == 0 ? pass_zero_twiddles(tw, l, F, B) : 0u;
  shared1_pass_step<F, G, THREADS, PAIR, 0, 0>(v, tab, basis, twv, FL, tid, zmask);
#pragma unroll
  for (int gi = 0; gi < G; ++gi)
#pragma unroll
    for (int e = 0; e < (1 << F); ++e) data[((base[gi] | ((uint64_t)e << s)) << 6) + lane] = v[gi][e];
}

// Two workgroups per CU (WAVES / 2 wavefronts per SIMD): at 8 wavefronts
// each a lane may hold 128 VGPRs, at 4 wavefronts 256.
template <int F, int WAVES, int G, bool PAIR>
__global__ __launch_bounds__(64 * WAVES, WAVES / 2) void ntt_pass_shared1_kernel(f128 *data,
                                                                                  const f128 *__restrict__ tw, int k,
                                                                                  int l, const f128 *src,
                                                                                  uint64_t src_mask) {
  __shared__ f128 lds[SHARED1_LDS_F128];
  ntt_pass_shared1_body<F, WAVES, G, PAIR>(data, tw, k, l, src, src_mask,
                                           pass_wg(blockIdx.x, gridDim.x, l, src_mask), lds);
}

template <int F, int WAVES, int G, bool PAIR>
__global__ __launch_bounds__(64 * WAVES, WAVES / 2) void ntt_pass_shared1_batch_kernel(
    const NttJob *__restrict__ jobs, uint32_t n_jobs) {
  __shared__ f128 lds[SHARED1_LDS_F128];
  const NttJob &J = job_of(jobs, n_jobs);
  ntt_pass_shared1_body<F, WAVES, G, PAIR>(J.data, J.tw, J.k, J.l, J.src ? J.src : J.data,
                                           J.src ? J.src_pos_mask : ~0ull, blockIdx.x - J.wg_begin, lds);
}

// ---------------------------------------------------------------- nibble-table shared passes
// The 2^F - 1 twiddles of the pass are shared by the workgroup's WAVES
// groups (2^s >= WAVES); their nibble tables (8 KiB each, 120 KiB for F = 4)
// are built once, up front, wavefront n building twiddle n's, and then every
// butterfly of the pass runs with no further barrier.  A product is 32
// lookups (16 with full byte tables), but nothing is rebuilt per twiddle, and
// a lane's 128-VGPR budget lets 16 wavefronts share a CU.
template <int F, int J, int C>
__device__ __forceinline__ void nibshared_step(f128 (&v)[1 << F], const char *lds, const NibtabLane &NL,
                                               uint32_t zmask) {
  constexpr int N = (1 << J) - 1 + C;
  constexpr int nb = 1 << (F - 1 - J);
  if ((zmask >> N) & 1u) {
    // a zero twiddle: u0 = v0, u1 = v1 + v0
#pragma unroll
    for (int el = 0; el < nb; ++el) {
      const int e0 = (C << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
      f128_xor_into(v[e1], v[e0]);
    }
  } else {
#pragma unroll
    for (int el = 0; el < nb; ++el) {
      const int e0 = (C << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
      const f128 u0 = f128_mul_nibtab_add<(uint32_t)N * NIBTAB_BYTES>(v[e1], lds, NL, v[e0]);
      v[e0] = u0;
      f128_xor_into(v[e1], u0);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if constexpr (C + 1 < (1 << J))
    nibshared_step<F, J, C + 1>(v, lds, NL, zmask);
  else if constexpr (J + 1 < F)
    nibshared_step<F, J + 1, 0>(v, lds, NL, zmask);
}

template <int F, int WAVES>
__device__ __forceinline__ void ntt_pass_nibshared_body(f128 *data, const f128 *__restrict__ tw, int k, int l,
                                                        const f128 *src, uint64_t src_mask, uint64_t wg, char *lds) {
  const uint32_t tid = threadIdx.x, lane = tid & 63;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const NibtabLane NL(lane & 15u);
  const int s = k - l - F;
  const uint64_t g = wg * WAVES + wave;
  const uint64_t B = g >> s;  // the same for the whole workgroup
  const uint64_t base = (B << (k - l)) | (g & ((1ull << s) - 1ull));
  f128 v[1 << F];
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) v[e] = src[(((base | ((uint64_t)e << s)) & src_mask) << 6) + lane];
  // twiddle n = 2^j - 1 + c of the pass (layer offset j, block c); a zero
  // twiddle gets no table
  const uint32_t zmask = B == 0 ? pass_zero_twiddles(tw, l, F, B) : 0u;
  for (uint32_t n = wave; n < (1u << F) - 1u; n += WAVES) {
    const uint32_t j = 31 - __builtin_clz(n + 1), c = n + 1 - (1u << j);
    if (!((zmask >> n) & 1u))
      build_nibtab<0>(tw[(1ull << (l + j)) - 1ull + (B << j) + c], lds + n * NIBTAB_BYTES, lane);
  }
  __syncthreads();
  nibshared_step<F, 0, 0>(v, lds, NL, zmask);
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) data[((base | ((uint64_t)e << s)) << 6) + lane] = v[e];
}

template <int F, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ntt_pass_nibshared_kernel(f128 *data, const f128 *__restrict__ tw,
                                                                        int k, int l, const f128 *src,
                                                                        uint64_t src_mask) {
  __shared__ __align__(16) char lds[((1 << F) - 1) * NIBTAB_BYTES];
  ntt_pass_nibshared_body<F, WAVES>(data, tw, k, l, src, src_mask, pass_wg(blockIdx.x, gridDim.x, l, src_mask),
                                    lds);
}

template <int F, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void ntt_pass_nibshared_batch_kernel(const NttJob *__restrict__ jobs,
                                                                              uint32_t n_jobs) {
  __shared__ __align__(16) char lds[((1 << F) - 1) * NIBTAB_BYTES];
  const NttJob &J = job_of(jobs, n_jobs);
  ntt_pass_nibshared_body<F, WAVES>(J.data, J.tw, J.k, J.l, J.src ? J.src : J.data,
                                    J.src ? J.src_pos_mask : ~0ull, blockIdx.x - J.wg_begin, lds);
}

// ---------------------------------------------------------------- launch
namespace {

struct PassArgs {
  f128 *data;
  const f128 *tw;
  int k, l;
  const f128 *src;
  uint64_t mask;
};

template <int F, int W, int G>
void launch_shared(dim3 grid, hipStream_t st, const PassArgs &a, const NttJob *jobs, uint32_t n_jobs) {
  if (jobs)
    hipLaunchKernelGGL((ntt_pass_shared_batch_kernel<F, W, G>), grid, dim3(64 * W), 0, st, jobs, n_jobs);
  else
    hipLaunchKernelGGL((ntt_pass_shared_kernel<F, W, G>), grid, dim3(64 * W), 0, st, a.data, a.tw, a.k, a.l, a.src,
                       a.mask);
}

template <int F, int W, int G, bool PAIR>
void launch_shared1(dim3 grid, hipStream_t st, const PassArgs &a, const NttJob *jobs, uint32_t n_jobs) {
  if (jobs)
    hipLaunchKernelGGL((ntt_pass_shared1_batch_kernel<F, W, G, PAIR>), grid, dim3(64 * W), 0, st, jobs, n_jobs);
  else
    hipLaunchKernelGGL((ntt_pass_shared1_kernel<F, W, G, PAIR>), grid, dim3(64 * W), 0, st, a.data, a.tw, a.k, a.l,
                       a.src, a.mask);
}

template <int F, int W>
void launch_nibshared(dim3 grid, hipStream_t st, const PassArgs &a, const NttJob *jobs, uint32_t n_jobs) {
  if (jobs)
    hipLaunchKernelGGL((ntt_pass_nibshared_batch_kernel<F, W>), grid, dim3(64 * W), 0, st, jobs, n_jobs);
  else
    hipLaunchKernelGGL((ntt_pass_nibshared_kernel<F, W>), grid, dim3(64 * W), 0, st, a.data, a.tw, a.k, a.l, a.src,
                       a.mask);
}

template <int F, int NB, int SJ>
void launch_tail(dim3 grid, hipStream_t st, const PassArgs &a, const NttJob *jobs, uint32_t n_jobs) {
  if (jobs)
    hipLaunchKernelGGL((ntt_tail_batch_kernel<F, NB, SJ>), grid, dim3(64), 0, st, jobs, n_jobs);
  else
    hipLaunchKernelGGL((ntt_tail_kernel<F, NB, SJ>), grid, dim3(64), 0, st, a.data, a.tw, a.k, a.l, a.src, a.mask);
}

bool is_tail(int kn) { return kn == NTT_TAIL_NB2 || kn == NTT_TAIL_NB1 || kn == NTT_TAIL_SHIFT; }

const char *kernel_phase(NttKernel kn) { return is_tail(kn) ? "ntt_tail_kernel" : "ntt_pass_shared_kernel"; }

// One pass, directly (jobs null) or as a fused launch over jobs; timed as
// `phase` (default: by kernel kind).
int launch_pass(Ctx *ctx, hipStream_t st, int F, NttKernel kn, uint64_t wgs, const PassArgs &a, const NttJob *jobs,
                uint32_t n_jobs, const char *phase = nullptr) {
  ScopedPhase ph(ctx, phase ? phase : kernel_phase(kn), st);
  const dim3 grid((unsigned)wgs);
  switch ((int)kn * 8 + F) {
#define FLOCK_SHARED(KN, F_, W, G) \
  case (int)KN * 8 + F_: launch_shared<F_, W, G>(grid, st, a, jobs, n_jobs); break;
#define FLOCK_TAIL(KN, F_, NB, SJ) \
  case (int)KN * 8 + F_: launch_tail<F_, NB, SJ>(grid, st, a, jobs, n_jobs); break;
    FLOCK_SHARED(NTT_SHARED_W8G2, 3, 8, 2)
    FLOCK_SHARED(NTT_SHARED_W8G2, 4, 8, 2)
    FLOCK_SHARED(NTT_SHARED_W8G1, 3, 8, 1)
    FLOCK_SHARED(NTT_SHARED_W8G1, 4, 8, 1)
    FLOCK_SHARED(NTT_SHARED_W16G1, 3, 16, 1)
    FLOCK_SHARED(NTT_SHARED_W16G1, 4, 16, 1)
    FLOCK_TAIL(NTT_TAIL_NB2, 1, 2, 1)
    FLOCK_TAIL(NTT_TAIL_NB2, 2, 2, 2)
    FLOCK_TAIL(NTT_TAIL_NB2, 3, 2, 3)
    FLOCK_TAIL(NTT_TAIL_NB2, 4, 2, 4)
    FLOCK_TAIL(NTT_TAIL_NB1, 1, 1, 1)
    FLOCK_TAIL(NTT_TAIL_NB1, 2, 1, 2)
    FLOCK_TAIL(NTT_TAIL_NB1, 3, 1, 3)
    FLOCK_TAIL(NTT_TAIL_NB1, 4, 1, 4)
    FLOCK_TAIL(NTT_TAIL_SHIFT, 1, 1, 0)
    FLOCK_TAIL(NTT_TAIL_SHIFT, 2, 1, 0)
    FLOCK_TAIL(NTT_TAIL_SHIFT, 3, 1, 0)
    FLOCK_TAIL(NTT_TAIL_SHIFT, 4, 1, 0)
#define FLOCK_NIBSHARED(KN, F_, W) \
  case (int)KN * 8 + F_: launch_nibshared<F_, W>(grid, st, a, jobs, n_jobs); break;
    FLOCK_NIBSHARED(NTT_NIBSHARED_W16, 3, 16)
    FLOCK_NIBSHARED(NTT_NIBSHARED_W16, 4, 16)
    FLOCK_NIBSHARED(NTT_NIBSHARED_W8, 3, 8)
    FLOCK_NIBSHARED(NTT_NIBSHARED_W8, 4, 8)
#define FLOCK_SHARED1(KN, F_, W, G, PAIR) \
  case (int)KN * 8 + F_: launch_shared1<F_, W, G, PAIR>(grid, st, a, jobs, n_jobs); break;
    FLOCK_SHARED1(NTT_SHARED1_W8G1, 3, 8, 1, false)
    FLOCK_SHARED1(NTT_SHARED1_W8G1, 4, 8, 1, false)
    FLOCK_SHARED1(NTT_SHARED1_W4G2, 3, 4, 2, true)
    FLOCK_SHARED1(NTT_SHARED1_W4G2, 4, 4, 2, true)
#undef FLOCK_SHARED
#undef FLOCK_SHARED1
#undef FLOCK_TAIL
#undef FLOCK_NIBSHARED
    default: ctx->last_error = "ntt: no kernel for this pass"; return FLOCK_E_ARG;
  }
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace

// ---------------------------------------------------------------- plans
int ntt_kernel_groups_log(NttKernel kn) {
  switch (kn) {
    case NTT_SHARED_W8G2: return 4;
    case NTT_SHARED_W8G1: return 3;
    case NTT_SHARED_W16G1: return 4;
    case NTT_NIBSHARED_W16: return 4;
    case NTT_NIBSHARED_W8: return 3;
    case NTT_SHARED1_W8G1: return 3;
    case NTT_SHARED1_W4G2: return 3;
    default: return 0;
  }
}

bool ntt_pass_feasible(int k, int l, int F, int kn) {
  if (l < 0 || F < 1 || F > 4 || l + F > k || kn < 0 || kn >= NTT_KERNELS) return false;
  if (is_tail(kn)) return true;
  return F >= 3 && k - l - F >= ntt_kernel_groups_log((NttKernel)kn);
}

uint64_t ntt_pass_wgs(int k, int F, NttKernel kn) { return (1ull << (k - F)) >> ntt_kernel_groups_log(kn); }

std::vector<NttPass> ntt_default_plan(int k, int start) {
  std::vector<NttPass> p;
  for (int l = start; l < k;) {
    const int F = std::min(4, k - l), s = k - l - F;
    const NttKernel kn = F == 4 && s >= 4 ? NTT_SHARED_W8G2 : F == 4 && s == 3 ? NTT_SHARED_W8G1 : NTT_TAIL_NB2;
    p.push_back({l, F, kn});
    l += F;
  }
  return p;
}

bool ntt_plan_valid(int k, int start, const std::vector<NttPass> &p) {
  if (start < 0 || k > 40 || start >= k) return p.empty();
  int l = start;
  for (const NttPass &q : p) {
    if (q.l != l || !ntt_pass_feasible(k, q.l, q.F, q.kernel)) return false;
    l += q.F;
  }
  return l == k;
}

namespace {

std::mutex g_plan_mu;
std::map<std::pair<int, int>, std::vector<NttPass>> g_plans;
bool g_plans_loaded = false;

std::string default_plan_path() {
  if (const char *e = getenv("FLOCK_NTT_PLANS")) return e;  // empty: no file
  Dl_info info;
  if (!dladdr((void *)&ntt_default_plan, &info) || !info.dli_fname) return "";
  std::string so = info.dli_fname;
  const size_t slash = so.rfind('/');
  const std::string dir = slash == std::string::npos ? "." : so.substr(0, slash);
  return dir + "/../tune/ntt_plans_gfx950.json";
}

bool parse_int(const std::string &t, size_t &i, long &out) {
  while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\t' || t[i] == '\r' || t[i] == ':' || t[i] == ','))
    ++i;
  size_t j = i;
  if (j < t.size() && t[j] == '-') ++j;
  if (j >= t.size() || t[j] < '0' || t[j] > '9') return false;
  char *end = nullptr;
  out = strtol(t.c_str() + i, &end, 10);
  i = (size_t)(end - t.c_str());
  return true;
}

// The plans of a file {"plans": [{"k": K, "start": S, "passes": [[l, F,
// kernel], ...], ...}, ...]} (other keys ignored).
bool parse_plans(const std::string &t, std::map<std::pair<int, int>, std::vector<NttPass>> &out, std::string &err) {
  size_t at = t.find("\"plans\"");
  if (at == std::string::npos) {
    err = "no \"plans\" key";
    return false;
  }
  for (size_t ob = t.find('{', at); ob != std::string::npos; ob = t.find('{', ob + 1)) {
    const size_t cb = t.find('}', ob);
    if (cb == std::string::npos) {
      err = "unterminated plan object";
      return false;
    }
    const std::string o = t.substr(ob, cb - ob + 1);
    long k = -1, start = -1;
    size_t i = o.find("\"k\"");
    if (i == std::string::npos || !parse_int(o, i += 3, k)) {
      err = "plan without \"k\"";
      return false;
    }
    i = o.find("\"start\"");
    if (i == std::string::npos || !parse_int(o, i += 7, start)) {
      err = "plan without \"start\"";
      return false;
    }
    i = o.find("\"passes\"");
    if (i == std::string::npos || (i = o.find('[', i)) == std::string::npos) {
      err = "plan without \"passes\"";
      return false;
    }
    ++i;
    std::vector<NttPass> p;
    while (true) {
      while (i < o.size() && (o[i] == ' ' || o[i] == ',' || o[i] == '\n' || o[i] == '\t' || o[i] == '\r')) ++i;
      if (i >= o.size()) {
        err = "unterminated passes";
        return false;
      }
      if (o[i] == ']') break;
      if (o[i] != '[') {
        err = "a pass is not [l, F, kernel]";
        return false;
      }
      ++i;
      long l, F, kn;
      if (!parse_int(o, i, l) || !parse_int(o, i, F) || !parse_int(o, i, kn)) {
        err = "a pass is not [l, F, kernel]";
        return false;
      }
      while (i < o.size() && (o[i] == ' ' || o[i] == '\n' || o[i] == '\t' || o[i] == '\r')) ++i;
      if (i >= o.size() || o[i] != ']') {
        err = "a pass is not [l, F, kernel]";
        return false;
      }
      ++i;
      p.push_back({(int)l, (int)F, (NttKernel)kn});
    }
    if (!ntt_plan_valid((int)k, (int)start, p)) {
      err = "invalid plan for k=" + std::to_string(k) + " start=" + std::to_string(start);
      return false;
    }
    out[{(int)k, (int)start}] = p;
    ob = cb;
  }
  return true;
}

int load_plans_locked(const std::string &path, std::string *err) {
  if (path.empty()) return FLOCK_OK;
  std::ifstream f(path);
  if (!f) {
    if (err) *err = "cannot open " + path;
    return err ? FLOCK_E_ARG : FLOCK_OK;  // the default file is optional
  }
  std::stringstream ss;
  ss << f.rdbuf();
  std::map<std::pair<int, int>, std::vector<NttPass>> plans;
  std::string e;
  if (!parse_plans(ss.str(), plans, e)) {
    if (err) *err = path + ": " + e;
    return FLOCK_E_ARG;
  }
  for (auto &kv : plans) g_plans[kv.first] = kv.second;
  return FLOCK_OK;
}

void ensure_loaded_locked() {
  if (!g_plans_loaded) {
    g_plans_loaded = true;
    load_plans_locked(default_plan_path(), nullptr);
  }
}

}  // namespace

std::vector<NttPass> ntt_plan(int k, int start) {
  std::lock_guard<std::mutex> g(g_plan_mu);
  ensure_loaded_locked();
  auto it = g_plans.find({k, start});
  return it != g_plans.end() ? it->second : ntt_default_plan(k, start);
}

int ntt_plan_install(int k, int start, const std::vector<NttPass> &p, std::string &err) {
  std::lock_guard<std::mutex> g(g_plan_mu);
  ensure_loaded_locked();
  if (p.empty()) {
    g_plans.erase({k, start});
    return FLOCK_OK;
  }
  if (!ntt_plan_valid(k, start, p)) {
    err = "ntt plan: passes must cover layers start..k-1 in order, each with a kernel that can run it";
    return FLOCK_E_ARG;
  }
  g_plans[{k, start}] = p;
  return FLOCK_OK;
}

int ntt_plans_load(const char *path, std::string &err) {
  std::lock_guard<std::mutex> g(g_plan_mu);
  g_plans_loaded = true;
  std::string e;
  const int rc = load_plans_locked(path ? std::string(path) : default_plan_path(), &e);
  if (rc) err = "ntt plans: " + e;
  return rc;
}

// ---------------------------------------------------------------- transform
int ntt_forward_interleaved_device(Ctx *ctx, hipStream_t st, void *d_data, int k, int num_lanes, int start_layer,
                                   const void *d_msg, const void *d_twiddles) {
  if (num_lanes != 64) {
    ctx->last_error = "ntt: only 64 interleaved lanes are implemented";
    return FLOCK_E_SHAPE;
  }
  const F128 *tw = d_twiddles ? (const F128 *)d_twiddles : ctx->device_twiddles(k);
  if (!tw) return FLOCK_E_DEVICE;
  const std::vector<NttPass> plan = ntt_plan(k, start_layer);
  bool first = true;
  for (const NttPass &q : plan) {
    const bool from_msg = first && d_msg;
    const PassArgs a{(f128 *)d_data, (const f128 *)tw, k, q.l, from_msg ? (const f128 *)d_msg : (const f128 *)d_data,
                     from_msg ? (1ull << (k - start_layer)) - 1 : ~0ull};
    // the pass that reads the message moves fewer bytes: its own phase
    if (int rc = launch_pass(ctx, st, q.F, q.kernel, ntt_pass_wgs(k, q.F, q.kernel), a, nullptr, 0,
                             from_msg ? "ntt_pass_from_msg" : nullptr))
      return rc;
    first = false;
  }
  return FLOCK_OK;
}

int ntt_batch_device(Ctx *ctx, hipStream_t st, int F, NttKernel kn, const NttJob *d_jobs, uint32_t n_jobs,
                     uint64_t wgs) {
  const PassArgs none{nullptr, nullptr, 0, 0, nullptr, 0};
  return launch_pass(ctx, st, F, kn, wgs, none, d_jobs, n_jobs);
}

// ---------------------------------------------------------------- autotuning
// Pseudo-random contents for the timed buffers (SplitMix64 of the word
// index): table lookups address by data bytes, so constant data would time
// a pattern no codeword has.
__global__ void fill_random_kernel(uint64_t *p, uint64_t n_words, uint64_t seed) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    p[i] = z ^ (z >> 31);
  }
}

int ntt_autotune(Ctx *ctx, int k, int start, int reps, std::string &json) {
  if (k < 2 || k > 30 || start < 1 || start >= k || reps < 1) {
    ctx->last_error = "ntt autotune: needs 1 <= start < k <= 30 and reps >= 1";
    return FLOCK_E_ARG;
  }
  const F128 *tw = ctx->device_twiddles(k);
  if (!tw) return FLOCK_E_DEVICE;
  hipStream_t st = ctx->stream;
  const size_t cw_bytes = (size_t)1024 << k, msg_bytes = cw_bytes >> start;
  void *cw = nullptr, *msg = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FLOCK_OK;
  // t[l][F][kernel] in ms (< 0: infeasible)
  std::vector<double> t((size_t)(k + 1) * 5 * NTT_KERNELS, -1.0);
  auto T = [&](int l, int F, int kn) -> double & { return t[((size_t)l * 5 + F) * NTT_KERNELS + kn]; };
  std::ostringstream timings;
  bool first_timing = true;
  const bool was_timing = ctx->timing;
  ctx->timing = false;
  do {
    if (hipMalloc(&cw, cw_bytes) != hipSuccess || hipMalloc(&msg, msg_bytes) != hipSuccess) {
      (void)hipGetLastError();
      ctx->last_error = "ntt autotune: out of device memory";
      rc = FLOCK_E_OOM;
      break;
    }
    hipLaunchKernelGGL(fill_random_kernel, dim3(4096), dim3(256), 0, st, (uint64_t *)cw, (uint64_t)(cw_bytes / 8),
                       0x5eed0001ull);
    hipLaunchKernelGGL(fill_random_kernel, dim3(4096), dim3(256), 0, st, (uint64_t *)msg, (uint64_t)(msg_bytes / 8),
                       0x5eed0002ull);
    if (hipGetLastError() != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
      ctx->last_error = "ntt autotune: setup failed";
      rc = FLOCK_E_DEVICE;
      break;
    }
    for (int l = start; l < k && !rc; ++l)
      for (int F = 1; F <= 4 && !rc; ++F)
        for (int kn = 0; kn < NTT_KERNELS && !rc; ++kn) {
          if (!ntt_pass_feasible(k, l, F, kn)) continue;
          const bool from_msg = l == start;
          const PassArgs a{(f128 *)cw, (const f128 *)tw, k, l, from_msg ? (const f128 *)msg : (const f128 *)cw,
                           from_msg ? (1ull << (k - start)) - 1 : ~0ull};
          const uint64_t wgs = ntt_pass_wgs(k, F, (NttKernel)kn);
          if ((rc = launch_pass(ctx, st, F, (NttKernel)kn, wgs, a, nullptr, 0))) break;  // warm-up
          if (hipEventRecord(e0, st) != hipSuccess) rc = FLOCK_E_DEVICE;
          for (int r = 0; r < reps && !rc; ++r) rc = launch_pass(ctx, st, F, (NttKernel)kn, wgs, a, nullptr, 0);
          if (rc || hipEventRecord(e1, st) != hipSuccess) {
            rc = rc ? rc : FLOCK_E_DEVICE;
            break;
          }
          if ((rc = ctx->wait_event(e1))) break;
          float ms = 0.f;
          if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = FLOCK_E_DEVICE;
            break;
          }
          T(l, F, kn) = ms / reps;
          timings << (first_timing ? "" : ", ") << "[" << l << ", " << F << ", " << kn << ", " << ms / reps << "]";
          first_timing = false;
        }
  } while (false);
  ctx->timing = was_timing;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (cw || msg) (void)hipStreamSynchronize(st);
  if (cw) (void)hipFree(cw);
  if (msg) (void)hipFree(msg);
  if (rc) return rc;
  // fastest partition: best[l] = min over passes (l, F, kernel) of its time
  // plus best[l + F]
  std::vector<double> best(k + 1, 1e30);
  std::vector<NttPass> choice(k + 1);
  best[k] = 0.0;
  for (int l = k - 1; l >= start; --l)
    for (int F = 1; F <= 4 && l + F <= k; ++F)
      for (int kn = 0; kn < NTT_KERNELS; ++kn) {
        const double x = T(l, F, kn);
        if (x < 0 || best[l + F] >= 1e30) continue;
        if (x + best[l + F] < best[l]) {
          best[l] = x + best[l + F];
          choice[l] = {l, F, (NttKernel)kn};
        }
      }
  std::vector<NttPass> plan;
  for (int l = start; l < k; l += choice[l].F) plan.push_back(choice[l]);
  double def_ms = 0.0;
  for (const NttPass &q : ntt_default_plan(k, start)) def_ms += T(q.l, q.F, q.kernel);
  if ((rc = ntt_plan_install(k, start, plan, ctx->last_error))) return rc;
  std::ostringstream o;
  o << "{\"k\": " << k << ", \"start\": " << start << ", \"passes\": [";
  for (size_t i = 0; i < plan.size(); ++i)
    o << (i ? ", " : "") << "[" << plan[i].l << ", " << plan[i].F << ", " << (int)plan[i].kernel << "]";
  o << "], \"ms\": " << best[start] << ", \"default_ms\": " << def_ms << ", \"timings\": [" << timings.str() << "]}";
  json = o.str();
  return FLOCK_OK;
}

}  // namespace flock
