// GF(2^128) vector kernels for the sumcheck-style phases of the prover:
// multilinear zerocheck rounds, eq-table expansion, lincheck folds, ring
// switch, the Ligerito sumcheck, a generic additive-NTT layer (forward and
// transposed), the BLAKE3 proof-of-work scan and index gathers.
//
// Reductions write one partial per workgroup; xor_rows_device (zerocheck.hip)
// or the host XORs them.  Kernels whose products have both operands varying
// per lane use f128_mul with a 16 KiB per-wavefront LDS table; products by a
// workgroup-uniform scalar use the byte tables of gf128_dev.h.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "ctx.h"
#include "fulltab.h"
#include "gf128_dev.h"
#include "gf128_short.h"
#include "nibtab.h"
#include "prove_kernels.h"
#include "transcript_dev.h"

namespace flock {

#define MUL_WAVES 2
#define MUL_THREADS (64 * MUL_WAVES)

__device__ __forceinline__ void wg_xor_reduce(f128 v, f128 *red, f128 *out) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = f128_xor(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
  __syncthreads();
}

// ---------------------------------------------------------------- eq tables
// out[i] = lo[i & (2^k - 1)] * hi[i >> k].  Each workgroup covers a range
// inside one hi index, so hi is a workgroup-uniform multiplier: its full byte
// tables (fulltab.h) are built once per workgroup.
constexpr int FT_THREADS = 1024;

__global__ __launch_bounds__(FT_THREADS) void eq_expand_kernel(const f128 *__restrict__ lo, int k,
                                                               const f128 *__restrict__ hi, uint64_t per_wg,
                                                               f128 *__restrict__ out) {
  __shared__ f128 tab[4096];
  __shared__ f128 basis[128];
  const uint32_t tid = threadIdx.x, rot = tid & 15;
  const uint64_t lo_n = 1ull << k;
  const uint64_t chunks = lo_n / per_wg;
  const uint64_t h = blockIdx.x / chunks, c0 = (blockIdx.x % chunks) * per_wg;
  fulltab_basis(hi[h], basis, tid);
  __syncthreads();
  build_fulltab<FT_THREADS>(basis, tab, tid);
  __syncthreads();
  const FulltabLane L(rot);
  uint64_t i = c0 + tid;
  for (; i + FT_THREADS < c0 + per_wg; i += 2 * FT_THREADS) {
    f128 p0, p1;
    f128_mul2_fulltab(lo[i], lo[i + FT_THREADS], (const char *)tab, L, p0, p1);
    out[(h << k) + i] = p0;
    out[(h << k) + i + FT_THREADS] = p1;
  }
  if (i < c0 + per_wg) out[(h << k) + i] = f128_mul_fulltab(lo[i], (const char *)tab, L);
}

int eq_expand_device(Ctx *ctx, hipStream_t st, const void *lo, int k, const void *hi, int n_hi_log, void *out) {
  uint64_t lo_n = 1ull << k;
  uint64_t per = lo_n < 8192 ? lo_n : 8192;
  uint64_t wgs = (lo_n / per) << n_hi_log;
  hipLaunchKernelGGL(eq_expand_kernel, dim3((unsigned)wgs), dim3(FT_THREADS), 0, st, (const f128 *)lo, k,
                     (const f128 *)hi, per, (f128 *)out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

static unsigned grid_for(uint64_t work, unsigned threads, unsigned cap) {
  uint64_t g = (work + threads - 1) / threads;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

// ---------------------------------------------------------------- zerocheck rounds, split eq
// A = L(row) for round 0 read from the packed rows (see zerocheck round 0
// below) through nibble tables T[v * 16 + p] = sum_{t < 4, bit t of v}
// L_{4p + t}(z) (4 KiB), read in a lane-dependent order p = step ^ (lane mod
// 16) so lanes hit distinct bank groups (entry (v, p) sits in bank group p).
__device__ __forceinline__ f128 lagrange_row(uint64_t row, const f128 *__restrict__ T, uint32_t rot) {
  const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
  f128 acc = f128_zero();
#pragma unroll
  for (uint32_t step = 0; step < 16; step += 2) {
    const uint32_t p0 = step ^ rot, p1 = (step + 1) ^ rot;
    const uint32_t v0 = __builtin_amdgcn_ubfe(p0 < 8 ? lo : hi, 4 * (p0 & 7), 4);
    const uint32_t v1 = __builtin_amdgcn_ubfe(p1 < 8 ? lo : hi, 4 * (p1 & 7), 4);
    acc = f128_xor3(acc, T[v0 * 16 + p0], T[v1 * 16 + p1]);
  }
  return acc;
}

// The round sums with E[x] = Elo[x mod 2^LOG] Ehi[x >> LOG] (eq over the low
// LOG variables times eq over the rest):
//   G = sum_{x_lo} Elo[x_lo] sum_{x_hi} Ehi[x_hi] p[x_lo, x_hi].
// A workgroup takes one x_hi at a time, so Ehi[x_hi] is a workgroup constant:
// its nibble tables (8 KiB, nibtab.h) serve all 2^LOG products by it, with
// no per-product table build, and each thread keeps its 2^LOG / THREADS
// x_lo's sums, weighted by Elo once at the end.  Also folds Ehi for the next
// round, Ehi_next[y] = Ehi[2y] + Ehi[2y+1], unless Ehi_next is null.
template <int THREADS, int LOG>
__global__ __launch_bounds__(THREADS) void zc_round_split_kernel(const f128 *__restrict__ A,
                                                                 const f128 *__restrict__ B,
                                                                 const f128 *__restrict__ Elo,
                                                                 const f128 *__restrict__ Ehi, uint64_t n_hi,
                                                                 f128 *__restrict__ partial,
                                                                 f128 *__restrict__ Ehi_next) {
  constexpr int WAVES = THREADS / 64, PER = (1 << LOG) / THREADS;
  static_assert(THREADS == 256 || THREADS == 512, "the table builds assume 256 or 512 threads");
  __shared__ f128 mt[WAVES][1024];  // per-lane product tables; the reduction reuses them
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  const uint32_t tid = threadIdx.x;
  f128 *tab = &mt[0][0];
  const CombLane CL(tid);
  const NibtabLane NL(tid & 15u);
  f128 s1[PER], si[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) s1[s] = si[s] = f128_zero();
  // the next pair's loads are issued before this pair's products (the index
  // after the last one clamped to it, so no branch)
  f128 na0, na1, nb0, nb1;
  auto load = [&](uint64_t x) {
    na0 = A[2 * x];
    na1 = A[2 * x + 1];
    nb0 = B[2 * x];
    nb1 = B[2 * x + 1];
  };
  if (blockIdx.x < n_hi) load(((uint64_t)blockIdx.x << LOG) + tid);
  for (uint64_t xh = blockIdx.x; xh < n_hi; xh += gridDim.x) {
    const f128 eh = Ehi[xh];
    if (Ehi_next && !(xh & 1) && tid == 0) Ehi_next[xh >> 1] = f128_xor(eh, Ehi[xh + 1]);
    __syncthreads();  // the previous x_hi's table reads are done
    if (tid < 256) build_nibtab256<0>(eh, nt, tid);
    __syncthreads();
    const uint64_t xh_next = xh + gridDim.x < n_hi ? xh + gridDim.x : xh;
#pragma unroll 1
    for (int s = 0; s < PER; ++s) {
      const uint64_t x = (xh << LOG) + (uint64_t)s * THREADS + tid;
      const f128 a0 = na0, a1 = na1, b0 = nb0, b1 = nb1;
      load(s + 1 < PER ? x + THREADS : (xh_next == xh ? x : (xh_next << LOG) + tid));
      const f128 p1 = f128_mul(a1, b1, tab, CL);
      const f128 pi = f128_mul(f128_xor(a0, a1), f128_xor(b0, b1), tab, CL);
      const f128 w1 = f128_mul_nibtab<0>(p1, nt, NL), wi = f128_mul_nibtab<0>(pi, nt, NL);
      // register-indexed accumulators without dynamic indexing
#pragma unroll
      for (int q = 0; q < PER; ++q)
        if (q == s) {
          f128_xor_into(s1[q], w1);
          f128_xor_into(si[q], wi);
        }
    }
  }
  f128 g1 = f128_zero(), gi = f128_zero();
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    f128_mul_build(Elo[s * THREADS + tid], tab, CL);
    f128_xor_into(g1, f128_mul_comb(s1[s], tab, CL));
    f128_xor_into(gi, f128_mul_comb(si[s], tab, CL));
  }
  __syncthreads();  // every lane's table reads are done before the reduction reuses mt
  wg_xor_reduce(g1, tab, &partial[2 * blockIdx.x]);
  wg_xor_reduce(gi, tab, &partial[2 * blockIdx.x + 1]);
}

// Round 0 over the packed rows a, b (row pair x: bits of (x, 0) and (x, 1)).
// The tables are A = L(row) for the GF(2)-linear L of the nibble tables T
// (lagrange_row), so with r1 = row (x, 1) and ri = row (x, 0) ^ row (x, 1)
//   Ehi p1 = (Ehi L)(a_1) L(b_1),   Ehi p_inf = (Ehi L)(a_i) L(b_i):
// the weight Ehi[x_hi] moves into the tables of one factor, T' = Ehi T,
// rebuilt per x_hi from T by 256 products through Ehi's nibble tables, and a
// row pair costs four table evaluations and two general products, with no
// product by Ehi.  Same sums as zc_round_split_kernel over A = L(a), B = L(b).
template <int THREADS, int LOG>
__global__ __launch_bounds__(THREADS) void zc_round0_split_kernel(const ulonglong2 *__restrict__ a,
                                                                  const ulonglong2 *__restrict__ b,
                                                                  const f128 *__restrict__ lw,
                                                                  const f128 *__restrict__ Elo,
                                                                  const f128 *__restrict__ Ehi, uint64_t n_hi,
                                                                  f128 *__restrict__ partial) {
  constexpr int WAVES = THREADS / 64, PER = (1 << LOG) / THREADS;
  static_assert(THREADS == 256 || THREADS == 512, "the table builds assume 256 or 512 threads");
  __shared__ f128 mt[WAVES][1024];  // per-lane product tables; the reduction reuses them
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  __shared__ f128 T[256], Tw[256];  // L's nibble tables, and Ehi[x_hi] times them
  const uint32_t tid = threadIdx.x, rot = tid & 15u;
  if (tid < 256) T[tid] = lw[tid];
  f128 *tab = &mt[0][0];
  const CombLane CL(tid);
  const NibtabLane NL(tid & 15u);
  // the sums stay unreduced (256-bit) until the end: no reduction per product
  F256w s1[PER], si[PER];
#pragma unroll
  for (int s = 0; s < PER; ++s) s1[s] = si[s] = F256w{};
  // the next row pair's loads are issued before this pair's products (the
  // index after the last one clamped to it, so no branch)
  ulonglong2 nar, nbr;
  if (blockIdx.x < n_hi) {
    nar = a[((uint64_t)blockIdx.x << LOG) + tid];
    nbr = b[((uint64_t)blockIdx.x << LOG) + tid];
  }
  for (uint64_t xh = blockIdx.x; xh < n_hi; xh += gridDim.x) {
    const f128 eh = Ehi[xh];
    __syncthreads();  // the previous x_hi's reads of nt and Tw are done (and T is visible)
    if (tid < 256) build_nibtab256<0>(eh, nt, tid);
    __syncthreads();
    if (tid < 256) Tw[tid] = f128_mul_nibtab<0>(T[tid], nt, NL);
    __syncthreads();
    const uint64_t xh_next = xh + gridDim.x < n_hi ? xh + gridDim.x : xh;
#pragma unroll 1
    for (int s = 0; s < PER; ++s) {
      const uint64_t x = (xh << LOG) + (uint64_t)s * THREADS + tid;
      const ulonglong2 ar = nar, br = nbr;
      const uint64_t xn = s + 1 < PER ? x + THREADS : (xh_next == xh ? x : (xh_next << LOG) + tid);
      nar = a[xn];
      nbr = b[xn];
      const f128 a1 = lagrange_row(ar.y, Tw, rot), ai = lagrange_row(ar.x ^ ar.y, Tw, rot);
      const f128 b1 = lagrange_row(br.y, T, rot), bi = lagrange_row(br.x ^ br.y, T, rot);
      uint32_t w1[8], wi[8];
      f128_mul_build(a1, tab, CL);
      f128_comb256(b1, tab, CL, w1);
      f128_mul_build(ai, tab, CL);
      f128_comb256(bi, tab, CL, wi);
#pragma unroll
      for (int q = 0; q < PER; ++q)
        if (q == s) {
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            s1[q].d[i] ^= w1[i];
            si[q].d[i] ^= wi[i];
          }
        }
    }
  }
  f128 g1 = f128_zero(), gi = f128_zero();
#pragma unroll
  for (int s = 0; s < PER; ++s) {
    f128_mul_build(Elo[s * THREADS + tid], tab, CL);
    f128_xor_into(g1, f128_mul_comb(f128_reduce256w(s1[s]), tab, CL));
    f128_xor_into(gi, f128_mul_comb(f128_reduce256w(si[s]), tab, CL));
  }
  __syncthreads();  // every lane's table reads are done before the reduction reuses mt
  wg_xor_reduce(g1, tab, &partial[2 * blockIdx.x]);
  wg_xor_reduce(gi, tab, &partial[2 * blockIdx.x + 1]);
}

// Rounds after round 0: 256-thread workgroups (72 KiB of LDS, two per CU)
// over F128 tables, split at ZS_LOG.  Round 0: 512-thread workgroups (144 KiB,
// one per CU) over the rows, split at ZS_LOG + 1.
int zc_round_split_device(Ctx *ctx, hipStream_t st, const void *A, const void *B, const void *Elo, const void *Ehi,
                          uint64_t half, void *out2, void *Ehi_next, uint32_t seq, const FsStep &fs) {
  ScopedPhase ph(ctx, "zc_round", st);
  const uint64_t n_hi = half >> ZS_LOG;
  if ((half & ((1ull << ZS_LOG) - 1)) || n_hi < 2) {
    ctx->last_error = "zerocheck: split round needs at least two x_hi blocks";
    return FLOCK_E_SHAPE;
  }
  const unsigned g = (unsigned)(n_hi < 512 ? n_hi : 512);
  f128 *part = (f128 *)ctx->get_scratch("zc.round.partial", (size_t)g * 2 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  hipLaunchKernelGGL((zc_round_split_kernel<256, ZS_LOG>), dim3(g), dim3(256), 0, st, (const f128 *)A,
                     (const f128 *)B, (const f128 *)Elo, (const f128 *)Ehi, n_hi, part, (f128 *)Ehi_next);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return xor_rows_device(ctx, st, part, g, 2, out2, seq, fs);
}

int zc_round0_split_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, const void *d_lw, const void *Elo,
                           const void *Ehi, uint64_t half, void *out2, uint32_t seq, const FsStep &fs) {
  ScopedPhase ph(ctx, "zc_round", st);
  constexpr int LOG = ZS_LOG + 1;
  const uint64_t n_hi = half >> LOG;
  if ((half & ((1ull << LOG) - 1)) || n_hi < 2) {
    ctx->last_error = "zerocheck: split round 0 needs at least two x_hi blocks";
    return FLOCK_E_SHAPE;
  }
  const unsigned g = (unsigned)(n_hi < 256 ? n_hi : 256);
  f128 *part = (f128 *)ctx->get_scratch("zc.round.partial", (size_t)g * 2 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  hipLaunchKernelGGL((zc_round0_split_kernel<512, LOG>), dim3(g), dim3(512), 0, st, (const ulonglong2 *)a,
                     (const ulonglong2 *)b, (const f128 *)d_lw, (const f128 *)Elo, (const f128 *)Ehi, n_hi, part);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return xor_rows_device(ctx, st, part, g, 2, out2, seq, fs);
}

// ---------------------------------------------------------------- zerocheck fold after round 0
// After round 1 the multilinear tables are A[x] = sum_s bit_s(a[x]) L_s(z)
// (fold_at_z_naive, zerocheck/multilinear.rs:598-625), a GF(2)-linear image
// of the 64-bit row.  Round 0 (zc_round0_split_kernel) and its fold
// read the rows and evaluate that map on the fly instead of materializing A
// and B.  The fold uses byte tables T[v * 8 + p] = sum_{t < 8, bit t of v}
// L_{8p + t}(z) (32 KiB), read in a lane-dependent order p = step ^ (lane mod 8).
__device__ __forceinline__ f128 lagrange_row8(uint64_t row, const f128 *__restrict__ T, uint32_t rot8) {
  const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
  f128 acc = f128_zero();
#pragma unroll
  for (uint32_t step = 0; step < 8; step += 2) {
    const uint32_t p0 = step ^ rot8, p1 = (step + 1) ^ rot8;
    const uint32_t v0 = __builtin_amdgcn_ubfe(p0 < 4 ? lo : hi, 8 * (p0 & 3), 8);
    const uint32_t v1 = __builtin_amdgcn_ubfe(p1 < 4 ? lo : hi, 8 * (p1 & 3), 8);
    acc = f128_xor3(acc, T[v0 * 8 + p0], T[v1 * 8 + p1]);
  }
  return acc;
}

// A'[x] = A[2x] + rho (A[2x] + A[2x+1]) with A = L(rows): by linearity
// A'[x] = L(a[2x]) + (rho L)(a[2x] ^ a[2x+1]), each map evaluated through its
// byte tables (entry (v, p) at v * 8 + p, as in round 0): 8 lookups per row.
// Two 32 KiB tables per workgroup leave room for two workgroups per CU.
constexpr int ZF0_THREADS = 512;

__global__ __launch_bounds__(ZF0_THREADS) void zc_fold0_kernel(const uint64_t *__restrict__ a,
                                                               const uint64_t *__restrict__ b,
                                                               const f128 *__restrict__ lw8,
                                                               const f128 *__restrict__ lwr8, uint64_t half,
                                                               f128 *__restrict__ A2, f128 *__restrict__ B2) {
  __shared__ f128 T[2048], R[2048];
  for (uint32_t i = threadIdx.x; i < 2048; i += ZF0_THREADS) {
    T[i] = lw8[i];
    R[i] = lwr8[i];
  }
  __syncthreads();
  const uint32_t rot8 = threadIdx.x & 7;
  const ulonglong2 *a2 = (const ulonglong2 *)a, *b2 = (const ulonglong2 *)b;
  for (uint64_t x = (uint64_t)blockIdx.x * ZF0_THREADS + threadIdx.x; x < half;
       x += (uint64_t)gridDim.x * ZF0_THREADS) {
    const ulonglong2 av = a2[x], bv = b2[x];
    A2[x] = f128_xor(lagrange_row8(av.x, T, rot8), lagrange_row8(av.x ^ av.y, R, rot8));
    B2[x] = f128_xor(lagrange_row8(bv.x, T, rot8), lagrange_row8(bv.x ^ bv.y, R, rot8));
  }
}

int zc_fold0_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, const void *d_lw8, const void *d_lwr8,
                    uint64_t half, void *A2, void *B2) {
  ScopedPhase ph(ctx, "fold_pair", st);
  hipLaunchKernelGGL(zc_fold0_kernel, dim3(grid_for(half, ZF0_THREADS, 512)), dim3(ZF0_THREADS), 0, st,
                     (const uint64_t *)a, (const uint64_t *)b, (const f128 *)d_lw8, (const f128 *)d_lwr8, half,
                     (f128 *)A2, (f128 *)B2);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// v'[x] = v[2x] + r (v[2x] + v[2x+1]) for two vectors (fold_in_place_pair,
// multilinear.rs:3209-3230; partial_eval_lsb, pcs/ligerito.rs:1673).  The
// product by r uses the workgroup's full byte tables of r (fulltab.h).
constexpr int FOLD_UNROLL = 2;

__global__ __launch_bounds__(FT_THREADS) void fold_pair_kernel(const f128 *__restrict__ A, const f128 *__restrict__ B,
                                                               uint64_t half, f128 r, const f128 *__restrict__ r_dev,
                                                               f128 *__restrict__ A2, f128 *__restrict__ B2) {
  __shared__ f128 tab[4096];
  __shared__ f128 basis[128];
  const uint32_t tid = threadIdx.x, rot = tid & 15;
  fulltab_basis(r_dev ? *r_dev : r, basis, tid);
  __syncthreads();
  build_fulltab<FT_THREADS>(basis, tab, tid);
  __syncthreads();
  const FulltabLane L(rot);
  // FOLD_UNROLL elements per thread per pass, all loaded before any product,
  // so a pass waits for memory once
  const uint64_t stride = (uint64_t)gridDim.x * FT_THREADS;
  for (uint64_t x0 = (uint64_t)blockIdx.x * FT_THREADS + tid; x0 < half; x0 += FOLD_UNROLL * stride) {
    f128 a0[FOLD_UNROLL], a1[FOLD_UNROLL], b0[FOLD_UNROLL], b1[FOLD_UNROLL];
#pragma unroll
    for (int u = 0; u < FOLD_UNROLL; ++u) {
      const uint64_t x = x0 + u * stride;
      if (x < half) {
        a0[u] = f128_load_nt(A + 2 * x);
        a1[u] = f128_load_nt(A + 2 * x + 1);
        if (B) {
          b0[u] = f128_load_nt(B + 2 * x);
          b1[u] = f128_load_nt(B + 2 * x + 1);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < FOLD_UNROLL; ++u) {
      const uint64_t x = x0 + u * stride;
      if (x >= half) break;
      if (B) {
        f128 pa = a0[u], pb = b0[u];
        f128_mul2_fulltab_add(f128_xor(a0[u], a1[u]), f128_xor(b0[u], b1[u]), (const char *)tab, L, pa, pb);
        A2[x] = pa;
        B2[x] = pb;
      } else {
        A2[x] = f128_mul_fulltab_add(f128_xor(a0[u], a1[u]), (const char *)tab, L, a0[u]);
      }
    }
  }
}

int fold_pair_device(Ctx *ctx, hipStream_t st, const void *A, const void *B, uint64_t half, F128 r, void *A2, void *B2,
                     const void *d_r) {
  ScopedPhase ph(ctx, "fold_pair", st);
  f128 rr = make_uint4((uint32_t)r.lo, (uint32_t)(r.lo >> 32), (uint32_t)r.hi, (uint32_t)(r.hi >> 32));
  // FOLD_UNROLL elements per thread (one memory wait) until 512 workgroups
  unsigned g = grid_for(half, FT_THREADS * FOLD_UNROLL, 512);
  hipLaunchKernelGGL(fold_pair_kernel, dim3(g), dim3(FT_THREADS), 0, st, (const f128 *)A, (const f128 *)B, half, rr,
                     (const f128 *)d_r, (f128 *)A2, (f128 *)B2);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// ---------------------------------------------------------------- rounds fused with their fold
// A round of the zerocheck or Ligerito sumcheck can fold the previous
// round's tables by its challenge itself (writing them, as fold_pair_kernel
// does) and compute its message over the folded values.  Up to SMALL_PAIRS
// pairs one workgroup also reduces the message and publishes it to the host
// (as xor_rows_kernel does): one launch per round instead of three.  Pair x
// is folded positions 2x, 2x + 1, from positions 4x .. 4x + 3.
constexpr int SMALL_THREADS = 256;
static_assert(SMALL_PAIRS <= (uint64_t)SMALL_THREADS, "one pair per thread in the single-workgroup rounds");

// v[4x] + r (v[4x] + v[4x+1]) and v[4x+2] + r (v[4x+2] + v[4x+3]), the
// products through the workgroup's nibble tables of r (r is the round's one
// challenge: a table built once per workgroup, not a comb per product).
__device__ __forceinline__ void fold4(const f128 *__restrict__ v, uint64_t x, const char *nt, const NibtabLane &NL,
                                      f128 &o0, f128 &o1) {
  const f128 a0 = v[4 * x], a1 = v[4 * x + 1], a2 = v[4 * x + 2], a3 = v[4 * x + 3];
  o0 = f128_mul_nibtab_add<0>(f128_xor(a0, a1), nt, NL, a0);
  o1 = f128_mul_nibtab_add<0>(f128_xor(a2, a3), nt, NL, a2);
}

// XOR of s0, s1 over the workgroup into out[0], out[1] (out[2] = 0 when
// three are written), then seq published at *flag.  red: 512 F128 of LDS no
// longer in use.
__device__ __forceinline__ void small_publish(f128 s0, f128 s1, f128 *red, f128 *out, int n_out, uint32_t *flag,
                                              uint32_t seq) {
  const uint32_t tid = threadIdx.x;
  __syncthreads();  // every lane's comb table reads are done
  red[tid] = s0;
  red[SMALL_THREADS + tid] = s1;
  __syncthreads();
  for (uint32_t st = SMALL_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] = f128_xor(red[tid], red[tid + st]);
      red[SMALL_THREADS + tid] = f128_xor(red[SMALL_THREADS + tid], red[SMALL_THREADS + tid + st]);
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = red[0];
    out[1] = red[SMALL_THREADS];
    if (n_out > 2) out[2] = f128_zero();
    if (seq) {
      __threadfence_system();
      __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// Zerocheck round over `half` pairs after folding A, B (4 half entries each)
// by rho into A2, B2 (2 half): (G(1), G(inf)) of round_pair_naive
// (zerocheck/multilinear.rs:639-680), sum_x E[x] A1 B1 and
// sum_x E[x] (A0 + A1)(B0 + B1) over the folded pairs, and the next round's
// eq table E_next[y] = E[2y] + E[2y+1] (the two eq factors of the dropped
// variable sum to one) unless E_next is null.
__global__ __launch_bounds__(SMALL_THREADS) void zc_fold_round_small_kernel(
    const f128 *__restrict__ A, const f128 *__restrict__ B, f128 rho_arg, const f128 *__restrict__ rho_dev,
    const f128 *__restrict__ E, uint32_t half, f128 *__restrict__ A2, f128 *__restrict__ B2,
    f128 *__restrict__ E_next, f128 *out, uint32_t *flag, uint32_t seq, FsState *fs, f128 *rho_out) {
  const f128 rho = rho_dev ? *rho_dev : rho_arg;
  __shared__ f128 mt[SMALL_THREADS / 64][1024];
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  __shared__ FsState fss;
  f128 *tab = &mt[0][0];
  const CombLane CL(threadIdx.x);
  const NibtabLane NL(threadIdx.x & 15u);
  const uint32_t x = threadIdx.x;
  if (x < 64) build_nibtab<0>(rho, nt, x);
  __syncthreads();
  // the next round's eq table apart from the products (as in zc_fold_round_kernel)
  if (E_next && x < half / 2) E_next[x] = f128_xor(E[2 * x], E[2 * x + 1]);
  f128 g1 = f128_zero(), gi = f128_zero();
  if (x < half) {
    f128 a0, a1, b0, b1;
    fold4(A, x, nt, NL, a0, a1);
    fold4(B, x, nt, NL, b0, b1);
    A2[2 * x] = a0;
    A2[2 * x + 1] = a1;
    B2[2 * x] = b0;
    B2[2 * x + 1] = b1;
    const f128 p1 = f128_mul(a1, b1, tab, CL);
    const f128 pi = f128_mul(f128_xor(a0, a1), f128_xor(b0, b1), tab, CL);
    f128_mul_build(E[x], tab, CL);
    g1 = f128_mul_comb(p1, tab, CL);
    gi = f128_mul_comb(pi, tab, CL);
  }
  small_publish(g1, gi, tab, out, 2, flag, seq);
  if (fs) fs_round_block(fs, fss, out, rho_out);  // thread 0 reads back what it wrote
}

// Ligerito round over `half` pairs after folding f, b (4 half entries each)
// by r into f2, b2: (u0, u2) as lig_msg_kernel without the evaluation.
__global__ __launch_bounds__(SMALL_THREADS) void lig_fold_msg_small_kernel(const f128 *__restrict__ f,
                                                                           const f128 *__restrict__ b, f128 r_arg,
                                                                           const f128 *__restrict__ r_dev,
                                                                           uint32_t half, f128 *__restrict__ f2,
                                                                           f128 *__restrict__ b2,
                                                                           f128 *__restrict__ out, uint32_t *flag,
                                                                           uint32_t seq, FsState *fs, uint32_t *dg) {
  const f128 r = r_dev ? *r_dev : r_arg;
  __shared__ f128 mt[SMALL_THREADS / 64][1024];
  __shared__ FsState fss;
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  f128 *tab = &mt[0][0];
  const CombLane CL(threadIdx.x);
  const NibtabLane NL(threadIdx.x & 15u);
  const uint32_t x = threadIdx.x;
  if (x < 64) build_nibtab<0>(r, nt, x);
  __syncthreads();
  f128 u0 = f128_zero(), u2 = f128_zero();
  if (x < half) {
    f128 f0, f1, b0, b1;
    fold4(f, x, nt, NL, f0, f1);
    fold4(b, x, nt, NL, b0, b1);
    f2[2 * x] = f0;
    f2[2 * x + 1] = f1;
    b2[2 * x] = b0;
    b2[2 * x + 1] = b1;
    u0 = f128_mul(f0, b0, tab, CL);
    u2 = f128_mul(f128_xor(f0, f1), f128_xor(b0, b1), tab, CL);
  }
  small_publish(u0, u2, tab, out, 3, flag, seq);
  if (fs) fs_observe2_block(fs, fss, out, dg);  // thread 0 reads back what it wrote
}

// The same fused rounds over more pairs: a grid of MUL_THREADS-thread
// workgroups (grid-stride over the pairs), one partial per workgroup, reduced
// and published by xor_rows_kernel.  The fold's products use comb tables of
// the challenge (rebuilt per pair: the message's products overwrite them).
__global__ __launch_bounds__(MUL_THREADS) void zc_fold_round_kernel(const f128 *__restrict__ A,
                                                                    const f128 *__restrict__ B, f128 rho_arg,
                                                                    const f128 *__restrict__ rho_dev,
                                                                    const f128 *__restrict__ E, uint64_t half,
                                                                    f128 *__restrict__ A2, f128 *__restrict__ B2,
                                                                    f128 *__restrict__ E_next,
                                                                    f128 *__restrict__ partial) {
  const f128 rho = rho_dev ? *rho_dev : rho_arg;
  __shared__ f128 mt[MUL_WAVES][1024];
  __shared__ f128 red[MUL_THREADS];
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  f128 *tab = &mt[0][0];
  const CombLane CL(threadIdx.x);
  const NibtabLane NL(threadIdx.x & 15u);
  if (threadIdx.x < 64) build_nibtab<0>(rho, nt, threadIdx.x);
  __syncthreads();
  // the next round's eq table first, in a loop of its own: a lane-dependent
  // store in the products' loop leaves the compiler holding more than 256
  // registers (AGPR spills, one wavefront per SIMD)
  if (E_next)
    for (uint64_t y = (uint64_t)blockIdx.x * MUL_THREADS + threadIdx.x; y < half / 2;
         y += (uint64_t)gridDim.x * MUL_THREADS)
      E_next[y] = f128_xor(E[2 * y], E[2 * y + 1]);
  f128 g1 = f128_zero(), gi = f128_zero();
  for (uint64_t x = (uint64_t)blockIdx.x * MUL_THREADS + threadIdx.x; x < half; x += (uint64_t)gridDim.x * MUL_THREADS) {
    f128 a0, a1, b0, b1;
    fold4(A, x, nt, NL, a0, a1);
    fold4(B, x, nt, NL, b0, b1);
    A2[2 * x] = a0;
    A2[2 * x + 1] = a1;
    B2[2 * x] = b0;
    B2[2 * x + 1] = b1;
    const f128 p1 = f128_mul(a1, b1, tab, CL);
    const f128 pi = f128_mul(f128_xor(a0, a1), f128_xor(b0, b1), tab, CL);
    f128_mul_build(E[x], tab, CL);
    f128_xor_into(g1, f128_mul_comb(p1, tab, CL));
    f128_xor_into(gi, f128_mul_comb(pi, tab, CL));
  }
  wg_xor_reduce(g1, red, &partial[2 * blockIdx.x]);
  wg_xor_reduce(gi, red, &partial[2 * blockIdx.x + 1]);
}

__global__ __launch_bounds__(MUL_THREADS) void lig_fold_msg_kernel(const f128 *__restrict__ f,
                                                                   const f128 *__restrict__ b, f128 r_arg,
                                                                   const f128 *__restrict__ r_dev, uint64_t half,
                                                                   f128 *__restrict__ f2, f128 *__restrict__ b2,
                                                                   f128 *__restrict__ partial) {
  const f128 r = r_dev ? *r_dev : r_arg;
  __shared__ f128 mt[MUL_WAVES][1024];
  __shared__ f128 red[MUL_THREADS];
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  f128 *tab = &mt[0][0];
  const CombLane CL(threadIdx.x);
  const NibtabLane NL(threadIdx.x & 15u);
  if (threadIdx.x < 64) build_nibtab<0>(r, nt, threadIdx.x);
  __syncthreads();
  f128 u0 = f128_zero(), u2 = f128_zero();
  for (uint64_t x = (uint64_t)blockIdx.x * MUL_THREADS + threadIdx.x; x < half; x += (uint64_t)gridDim.x * MUL_THREADS) {
    f128 f0, f1, b0, b1;
    fold4(f, x, nt, NL, f0, f1);
    fold4(b, x, nt, NL, b0, b1);
    f2[2 * x] = f0;
    f2[2 * x + 1] = f1;
    b2[2 * x] = b0;
    b2[2 * x + 1] = b1;
    f128_xor_into(u0, f128_mul(f0, b0, tab, CL));
    f128_xor_into(u2, f128_mul(f128_xor(f0, f1), f128_xor(b0, b1), tab, CL));
  }
  wg_xor_reduce(u0, red, &partial[3 * blockIdx.x]);
  wg_xor_reduce(u2, red, &partial[3 * blockIdx.x + 1]);
  if (threadIdx.x == 0) partial[3 * blockIdx.x + 2] = f128_zero();
}

static f128 to_f128(F128 v) {
  return make_uint4((uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi, (uint32_t)(v.hi >> 32));
}

int zc_fold_round_device(Ctx *ctx, hipStream_t st, const void *A, const void *B, F128 rho, const void *E,
                               uint64_t half, void *A2, void *B2, void *E_next, void *out2, uint32_t seq,
                               const void *d_rho, const FsStep &fs) {
  ScopedPhase ph(ctx, "zc_round", st);
  if (half == 0) {
    ctx->last_error = "zerocheck: fused round needs at least one pair";
    return FLOCK_E_SHAPE;
  }
  if (half < 2) E_next = nullptr;
  if (half <= SMALL_PAIRS) {
    hipLaunchKernelGGL(zc_fold_round_small_kernel, dim3(1), dim3(SMALL_THREADS), 0, st, (const f128 *)A,
                       (const f128 *)B, to_f128(rho), (const f128 *)d_rho, (const f128 *)E, (uint32_t)half,
                       (f128 *)A2, (f128 *)B2, (f128 *)E_next, (f128 *)out2, ctx->readback_flag_dev(), seq,
                       (FsState *)fs.state, (f128 *)fs.rho);
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
    return FLOCK_OK;
  }
  const unsigned g = grid_for(half, MUL_THREADS, 4096);
  f128 *part = (f128 *)ctx->get_scratch("zc.round.partial", (size_t)g * 2 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  hipLaunchKernelGGL(zc_fold_round_kernel, dim3(g), dim3(MUL_THREADS), 0, st, (const f128 *)A, (const f128 *)B,
                     to_f128(rho), (const f128 *)d_rho, (const f128 *)E, half, (f128 *)A2, (f128 *)B2,
                     (f128 *)E_next, part);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return xor_rows_device(ctx, st, part, g, 2, out2, seq, fs);
}

int lig_fold_msg_device(Ctx *ctx, hipStream_t st, const void *f, const void *b, F128 r, uint64_t half,
                              void *f2, void *b2, void *out3, uint32_t seq, const void *d_r, const FsStep &fs) {
  if (fs.rho) {
    ctx->last_error = "ligerito: a round's transcript step observes its message only";
    return FLOCK_E_ARG;
  }
  ScopedPhase ph(ctx, "lig_msg", st);
  if (half == 0) {
    ctx->last_error = "ligerito: fused round needs at least one pair";
    return FLOCK_E_SHAPE;
  }
  if (half <= SMALL_PAIRS) {
    hipLaunchKernelGGL(lig_fold_msg_small_kernel, dim3(1), dim3(SMALL_THREADS), 0, st, (const f128 *)f,
                       (const f128 *)b, to_f128(r), (const f128 *)d_r, (uint32_t)half, (f128 *)f2, (f128 *)b2,
                       (f128 *)out3, ctx->readback_flag_dev(), seq, (FsState *)fs.state, (uint32_t *)fs.digest);
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
    return FLOCK_OK;
  }
  const unsigned g = grid_for(half, MUL_THREADS, 4096);
  f128 *part = (f128 *)ctx->get_scratch("lig.msg.partial", (size_t)g * 3 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  hipLaunchKernelGGL(lig_fold_msg_kernel, dim3(g), dim3(MUL_THREADS), 0, st, (const f128 *)f, (const f128 *)b,
                     to_f128(r), (const f128 *)d_r, half, (f128 *)f2, (f128 *)b2, part);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return xor_rows_device(ctx, st, part, g, 3, out3, seq, fs);
}

// ---------------------------------------------------------------- Ligerito sumcheck
// round_msg_lsb (pcs/ligerito.rs:3725-3772): u0 = sum f0 b0,
// u2 = sum (f0 + f1)(b0 + b1); with_eval adds y = sum f0 b0 + f1 b1
// (round_msg_and_eval_lsb, :3775-3816).
// (The launch bounds ask for the two wavefronts per SIMD the LDS allows:
// left to itself, the scheduler goes past 256 registers into AGPRs, which
// leaves one.)
template <bool EVAL>
__global__ __launch_bounds__(MUL_THREADS, 2) void lig_msg_kernel(const f128 *__restrict__ f, const f128 *__restrict__ b,
                                                              uint64_t half, f128 *__restrict__ partial) {
  __shared__ f128 mt[MUL_WAVES][1024];
  __shared__ f128 red[MUL_THREADS];
  const uint32_t lane = threadIdx.x & 63;
  f128 *tab = &mt[0][0];
  const CombLane CL(threadIdx.x);
  // the sums stay unreduced (256-bit) until the end; y = u0 + sum f1 b1
  F256w u0 = {}, u2 = {}, y1 = {};
  // the next pair's loads are issued before this pair's products (its index
  // clamped to this one at the end, so no branch), so their latency hides
  // behind the products
  const uint64_t stride = (uint64_t)gridDim.x * MUL_THREADS;
  uint64_t x = (uint64_t)blockIdx.x * MUL_THREADS + threadIdx.x;
  f128 nf0, nf1, nb0, nb1;
  if (x < half) {
    nf0 = f[2 * x];
    nf1 = f[2 * x + 1];
    nb0 = b[2 * x];
    nb1 = b[2 * x + 1];
  }
  for (; x < half; x += stride) {
    const f128 f0 = nf0, f1 = nf1, b0 = nb0, b1 = nb1;
    const uint64_t xn = x + stride < half ? x + stride : x;
    nf0 = f[2 * xn];
    nf1 = f[2 * xn + 1];
    nb0 = b[2 * xn];
    nb1 = b[2 * xn + 1];
    f128_mul_build(f0, tab, CL);
    f128_mul_comb_acc(b0, tab, CL, u0);
    f128_mul_build(f128_xor(f0, f1), tab, CL);
    f128_mul_comb_acc(f128_xor(b0, b1), tab, CL, u2);
    if constexpr (EVAL) {
      f128_mul_build(f1, tab, CL);
      f128_mul_comb_acc(b1, tab, CL, y1);
    }
  }
  const f128 r0 = f128_reduce256w(u0);
  wg_xor_reduce(r0, red, &partial[3 * blockIdx.x]);
  wg_xor_reduce(f128_reduce256w(u2), red, &partial[3 * blockIdx.x + 1]);
  wg_xor_reduce(EVAL ? f128_xor(r0, f128_reduce256w(y1)) : f128_zero(), red, &partial[3 * blockIdx.x + 2]);
}

int lig_msg_device(Ctx *ctx, hipStream_t st, const void *f, const void *b, uint64_t half, int with_eval, void *out3,
                   uint32_t seq, const FsStep &fs) {
  if (fs.rho) {
    ctx->last_error = "ligerito: a round's transcript step observes its message only";
    return FLOCK_E_ARG;
  }
  ScopedPhase ph(ctx, "lig_msg", st);
  // at most the 1024 workgroups the CUs hold at once (four each, set by LDS):
  // every workgroup starts at once and the partials to reduce stay few
  unsigned g = grid_for(half, MUL_THREADS, 1024);
  f128 *part = (f128 *)ctx->get_scratch("lig.msg.partial", (size_t)g * 3 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  if (with_eval)
    hipLaunchKernelGGL(lig_msg_kernel<true>, dim3(g), dim3(MUL_THREADS), 0, st, (const f128 *)f, (const f128 *)b, half,
                       part);
  else
    hipLaunchKernelGGL(lig_msg_kernel<false>, dim3(g), dim3(MUL_THREADS), 0, st, (const f128 *)f, (const f128 *)b,
                       half, part);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return xor_rows_device(ctx, st, part, g, 3, out3, seq, fs);
}

// acc[i] += beta * x[i]  (SumcheckProver::glue, pcs/ligerito.rs:5766-5790),
// the products through the workgroup's nibble tables of beta.
__global__ __launch_bounds__(256) void axpy_kernel(f128 *__restrict__ acc, const f128 *__restrict__ x, uint64_t n,
                                                   f128 beta) {
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  build_nibtab256<0>(beta, nt, threadIdx.x);
  __syncthreads();
  const NibtabLane NL(threadIdx.x & 15u);
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    f128_xor_into(acc[i], f128_mul_nibtab<0>(x[i], nt, NL));
}

int axpy_device(Ctx *ctx, hipStream_t st, void *acc, const void *x, uint64_t n, F128 beta) {
  f128 bb = make_uint4((uint32_t)beta.lo, (uint32_t)(beta.lo >> 32), (uint32_t)beta.hi, (uint32_t)(beta.hi >> 32));
  hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, (f128 *)acc, (const f128 *)x, n, bb);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// ---------------------------------------------------------------- lincheck
// comb[c] = alpha (A_0^T eq)[c] + (B_0^T eq)[c] (+ beta at the pinned column)
// (fold_alpha_batched, lincheck.rs:1158-1240; const pin, :1530-1545).
//
// The two transposes hold about 21M nonzeros (rows as 16-bit indices) against
// a 2^14-entry eq table (256 KiB).  A workgroup per CU stages half of eq in
// LDS at a time (128 KiB) and gathers from it: each column's rows are
// ascending, so its entries in the lower half come first (up to split[c]).
// A wavefront takes four columns; its lanes read the rows 8 at a time (one
// 16-byte load) and XOR the eq entries into per-lane sums, reduced across
// the wavefront at the end.  The product by alpha (a workgroup constant) uses
// its nibble tables, one per column.
constexpr int LC_THREADS = 1024, LC_WGS = 256, LC_HALF = 1 << 13;
constexpr int LC_COLS_PER_WAVE = (1 << 14) / (LC_WGS * (LC_THREADS / 64));

__device__ __forceinline__ f128 wave_xor(f128 v) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    v.x ^= __shfl_xor(v.x, d);
    v.y ^= __shfl_xor(v.y, d);
    v.z ^= __shfl_xor(v.z, d);
    v.w ^= __shfl_xor(v.w, d);
  }
  return v;
}

// acc ^= eqh[row - base] for the 8 rows packed in w, entries k .. k + 7 of
// the column's row array, those of them inside [b, e) (all in the staged half).
__device__ __forceinline__ void lc_take(uint4 w, uint32_t k, uint32_t b, uint32_t e, const f128 *eqh, uint32_t base,
                                        f128 &acc) {
  const uint32_t ww[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const uint32_t kj = k + (uint32_t)j;
    if (kj >= b && kj < e) f128_xor_into(acc, eqh[((ww[j >> 1] >> (16 * (j & 1))) & 0xFFFFu) - base]);
  }
}

// A's entries [ab, ae) into sa and B's [bb, be) into sb: lanes read the rows 8
// at a time (16-byte loads from ab, bb rounded down to 8 entries; the row
// arrays are padded so the last load stays inside), four loads in flight.
__device__ __forceinline__ void lc_gather(const uint16_t *__restrict__ ra, uint32_t ab, uint32_t ae,
                                          const uint16_t *__restrict__ rb, uint32_t bb, uint32_t be,
                                          const f128 *eqh, uint32_t base, uint32_t lane, f128 &sa, f128 &sb) {
  uint32_t ka = (ab & ~7u) + 8u * lane, kb = (bb & ~7u) + 8u * lane;
  while (ka < ae || kb < be) {
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    const uint4 a0 = ka < ae ? *(const uint4 *)(ra + ka) : z;
    const uint4 a1 = ka + 512 < ae ? *(const uint4 *)(ra + ka + 512) : z;
    const uint4 b0 = kb < be ? *(const uint4 *)(rb + kb) : z;
    const uint4 b1 = kb + 512 < be ? *(const uint4 *)(rb + kb + 512) : z;
    lc_take(a0, ka, ab, ae, eqh, base, sa);
    lc_take(a1, ka + 512, ab, ae, eqh, base, sa);
    lc_take(b0, kb, bb, be, eqh, base, sb);
    lc_take(b1, kb + 512, bb, be, eqh, base, sb);
    ka += 1024;
    kb += 1024;
  }
}

__global__ __launch_bounds__(LC_THREADS) void lc_comb_kernel(LcCsc A, LcCsc B, const f128 *__restrict__ eq,
                                                             f128 alpha, f128 beta, uint32_t pin,
                                                             f128 *__restrict__ comb) {
  __shared__ f128 eqh[LC_HALF];
  __shared__ __align__(16) char nt[NIBTAB_BYTES];
  __shared__ f128 cols[LC_COLS_PER_WAVE * (LC_THREADS / 64)][2];
  const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid < 256) build_nibtab256<0>(alpha, nt, tid);
  f128 sa[LC_COLS_PER_WAVE], sb[LC_COLS_PER_WAVE];
#pragma unroll
  for (int i = 0; i < LC_COLS_PER_WAVE; ++i) sa[i] = sb[i] = f128_zero();
#pragma unroll 1
  for (uint32_t h = 0; h < 2; ++h) {
    __syncthreads();  // the previous half's gathers are done
    for (uint32_t i = tid; i < LC_HALF; i += LC_THREADS) eqh[i] = eq[h * LC_HALF + i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < LC_COLS_PER_WAVE; ++i) {
      const uint32_t c = blockIdx.x + LC_WGS * (w + (LC_THREADS / 64) * (uint32_t)i);
      const uint32_t ab = h ? A.split[c] : A.ptr[c], ae = h ? A.ptr[c + 1] : A.split[c];
      const uint32_t bb = h ? B.split[c] : B.ptr[c], be = h ? B.ptr[c + 1] : B.split[c];
      lc_gather(A.rows, ab, ae, B.rows, bb, be, eqh, h * LC_HALF, lane, sa[i], sb[i]);
    }
  }
#pragma unroll
  for (int i = 0; i < LC_COLS_PER_WAVE; ++i) {
    const f128 ta = wave_xor(sa[i]), tb = wave_xor(sb[i]);
    if (lane == 0) {
      cols[w + (LC_THREADS / 64) * i][0] = ta;
      cols[w + (LC_THREADS / 64) * i][1] = tb;
    }
  }
  __syncthreads();
  constexpr uint32_t N_COLS = LC_COLS_PER_WAVE * (LC_THREADS / 64);
  if (tid < N_COLS) {
    const uint32_t c = blockIdx.x + LC_WGS * tid;
    f128 v = f128_mul_nibtab_add<0>(cols[tid][0], nt, NibtabLane(tid & 15u), cols[tid][1]);
    if (c == pin) f128_xor_into(v, beta);
    comb[c] = v;
  }
}

int lc_comb_device(Ctx *ctx, hipStream_t st, LcCsc a, LcCsc b, const void *eq, F128 alpha, F128 beta, uint32_t pin,
                   void *comb) {
  ScopedPhase ph(ctx, "lc_comb", st);
  static_assert(LC_COLS_PER_WAVE * (LC_THREADS / 64) * LC_WGS == 1 << 14, "columns per wavefront");
  auto cv = [](F128 v) { return make_uint4((uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi, (uint32_t)(v.hi >> 32)); };
  hipLaunchKernelGGL(lc_comb_kernel, dim3(LC_WGS), dim3(LC_THREADS), 0, st, a, b, (const f128 *)eq, cv(alpha),
                     cv(beta), pin, (f128 *)comb);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// ---------------------------------------------------------------- ring switch
// b[i] = sum_k M_k(eq_k[i]), M_k(v) = sum_bits bit_b(v) w_k[b] with
// w_k = gamma_k eq(r''_k) (fold_b128_elems, ring_switch.rs:1720-1760,
// combined as in pcs.rs:1025-1100); M_k comes as nibble tables Mt[k][p][v]
// (2 x 32 x 16 F128, host-built: entry = M_k(v x^(4p))).
//
// eq_k[i] = lo_k[i mod 2^k] hi_k[i >> k], and for one x_hi = i >> k the map
// N_k(v) = M_k(v hi_k[x_hi]) is GF(2)-linear in v, with images
// N_k(x^b) = M_k(x^b hi_k[x_hi]) of the basis x^b.  The workgroup of an x_hi
// computes those 2 x 128 images, builds full byte tables of N_0 and N_1 from
// them (fulltab.h, 64 KiB each) and applies them to lo_k[0 .. 2^k): two
// conflict-free 16-lookup products per element, and the 2^L-entry eq tables
// are never written or read.  The images of the workgroup's next x_hi are
// computed while the current tables are in use (two image buffers), so only
// the table builds sit between barriers.
constexpr int RS_THREADS = 1024;

__global__ __launch_bounds__(RS_THREADS) void rs_basis_kernel(const f128 *__restrict__ Mt, const f128 *__restrict__ lo0,
                                                              const f128 *__restrict__ lo1, int k,
                                                              const f128 *__restrict__ hi0,
                                                              const f128 *__restrict__ hi1, uint64_t n_hi,
                                                              f128 *__restrict__ out) {
  __shared__ f128 tab[2 * 4096];       // byte tables of N_0, then those of N_1 64 KiB above
  __shared__ f128 mtab[2 * 32 * 16];   // nibble tables of M_0, M_1
  __shared__ f128 basis[2][2][128];    // [buffer][claim][i * 16 + kb] = N_claim(x^(8 kb + i))
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < 2 * 32 * 16; i += RS_THREADS) mtab[i] = Mt[i];
  __syncthreads();
  // Four adjacent threads compute one image: claim c, power b; quarter q
  // evaluates M_c on the eight nibbles of word q of x^b hi_c[x_hi].
  const uint32_t im = tid >> 2, q = tid & 3, c = im >> 7, b = im & 127;
  auto images = [&](uint64_t xh, int buf) {
    const f128 y = f128_mul_xpow(c ? hi1[xh] : hi0[xh], b);
    const uint32_t w = f128_word(y, (int)q);
    const f128 *t = mtab + c * 512 + q * 128;
    f128 s = f128_zero();
#pragma unroll
    for (int j = 0; j < 8; ++j) f128_xor_into(s, t[j * 16 + ((w >> (4 * j)) & 15)]);
#pragma unroll
    for (int o = 1; o < 4; o <<= 1) {
      s.x ^= __shfl_xor(s.x, o);
      s.y ^= __shfl_xor(s.y, o);
      s.z ^= __shfl_xor(s.z, o);
      s.w ^= __shfl_xor(s.w, o);
    }
    if (q == 0) basis[buf][c][(b & 7) * 16 + (b >> 3)] = s;
  };
  const FulltabLane L(tid & 15);
  const uint64_t lo_n = 1ull << k;
  int buf = 0;
  if (blockIdx.x < n_hi) images(blockIdx.x, 0);
  for (uint64_t xh = blockIdx.x; xh < n_hi; xh += gridDim.x, buf ^= 1) {
    __syncthreads();  // basis[buf] is complete; the previous x_hi's table reads are done
    build_fulltab<RS_THREADS>(&basis[buf][0][0], tab, tid);
    build_fulltab<RS_THREADS>(&basis[buf][1][0], tab + 4096, tid);
    __syncthreads();
    if (xh + gridDim.x < n_hi) images(xh + gridDim.x, buf ^ 1);
    for (uint64_t il = tid; il < lo_n; il += RS_THREADS) {
      f128 p0, p1;
      f128_mul2_fulltab<false, true>(lo0[il], lo1[il], (const char *)tab, L, p0, p1);
      out[(xh << k) + il] = f128_xor(p0, p1);
    }
  }
}

int rs_basis_device(Ctx *ctx, hipStream_t st, const void *tables, const void *lo0, const void *lo1, int k,
                    const void *hi0, const void *hi1, int n_hi_log, void *out) {
  ScopedPhase ph(ctx, "rs_basis", st);
  const uint64_t n_hi = 1ull << n_hi_log;
  // one workgroup per CU (148 KiB of LDS), each keeping M_k for all its x_hi's
  const unsigned g = (unsigned)(n_hi < 256 ? n_hi : 256);
  hipLaunchKernelGGL(rs_basis_kernel, dim3(g), dim3(RS_THREADS), 0, st, (const f128 *)tables, (const f128 *)lo0,
                     (const f128 *)lo1, k, (const f128 *)hi0, (const f128 *)hi1, n_hi, (f128 *)out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// Several layers per launch for narrow codewords (2^log_nl interleaved lanes,
// log_nl small): thread = (group, lane) holds the 2^F positions of its group
// (the positions that exchange data in layers l .. l+F-1) in registers.
// Forward applies layers l .. l+F-1; transposed applies their transposes in
// reverse order (l+F-1 .. l), which is the next F steps of the transposed
// transform.
// Twiddle products: when the 2^(6 - log_nl) groups of a wavefront share
// their block (UT: 2^s >= groups per wavefront), each twiddle is
// wavefront-uniform, so the wavefront builds its nibble tables together (8
// entries per lane) and a product is 32 lookups of pre-reduced entries;
// otherwise (the deepest layers) each lane builds a comb table of its own
// twiddle (the general product).
template <int F, bool TRANSPOSED, bool UT>
__global__ __launch_bounds__(MUL_THREADS) void ntt_pass_narrow_kernel(f128 *__restrict__ d,
                                                                       const f128 *__restrict__ tw, int k, int l,
                                                                       int log_nl, uint64_t n_threads) {
  // comb: 256 B per lane; nibble tables: 8 KiB per wavefront
  __shared__ f128 mt[MUL_WAVES][UT ? NIBTAB_BYTES / 16 : 1024];
  const uint32_t lane = threadIdx.x & 63;
  f128 *tab = &mt[0][0];
  const CombLane CL(threadIdx.x);
  const NibtabLane NL(lane & 15u);
  char *wave_tab = (char *)&mt[threadIdx.x >> 6][0];  // UT: the wavefront's nibble tables
  const uint64_t t = (uint64_t)blockIdx.x * MUL_THREADS + threadIdx.x;
  if (t >= n_threads) return;  // with UT, n_threads is a multiple of 64: whole wavefronts return
  const uint64_t ln = t & ((1ull << log_nl) - 1), g = t >> log_nl;
  const int s = k - l - F;
  const uint64_t r = g & ((1ull << s) - 1), B = g >> s;
  const uint64_t base = (B << (k - l)) | r;
  f128 v[1 << F];
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) v[e] = d[((base | ((uint64_t)e << s)) << log_nl) + ln];
  if (TRANSPOSED) {
    // The induced-basis transform starts from a vector with a few hundred
    // nonzeros: an all-zero group maps to zeros (the map is linear).  With
    // UT the wavefront works together, so it returns only if all its
    // groups are zero.
    uint32_t any = 0;
#pragma unroll
    for (int e = 0; e < (1 << F); ++e) any |= v[e].x | v[e].y | v[e].z | v[e].w;
    if (UT ? __builtin_amdgcn_ballot_w64(any != 0) == 0 : !any) return;
  }
#pragma unroll
  for (int jj = 0; jj < F; ++jj) {
    const int j = TRANSPOSED ? F - 1 - jj : jj;
    const uint64_t tw_base = (1ull << (l + j)) - 1ull + (B << j);
#pragma unroll
    for (int c = 0; c < (1 << j); ++c) {
      if constexpr (UT) {
        // every lane's previous lookups precede the rebuild (a wavefront's
        // LDS accesses execute in order; the fences keep the compiler's)
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        build_nibtab<0>(tw[tw_base + c], wave_tab, lane);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        __builtin_amdgcn_sched_barrier(0);
      } else {
        f128_mul_build(tw[tw_base + c], tab, CL);  // one table per twiddle
      }
#pragma unroll
      for (int el = 0; el < (1 << (F - 1 - j)); ++el) {
        const int e0 = (c << (F - j)) | el, e1 = e0 | (1 << (F - 1 - j));
        if (!TRANSPOSED) {
          const f128 u = UT ? f128_mul_nibtab_add<0>(v[e1], wave_tab, NL, v[e0])
                            : f128_xor(v[e0], f128_mul_comb(v[e1], tab, CL));
          v[e0] = u;
          v[e1] = f128_xor(v[e1], u);
        } else {
          const f128 sum = f128_xor(v[e0], v[e1]);
          v[e0] = sum;
          v[e1] = UT ? f128_mul_nibtab_add<0>(sum, wave_tab, NL, v[e1]) : f128_xor(f128_mul_comb(sum, tab, CL), v[e1]);
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) d[((base | ((uint64_t)e << s)) << log_nl) + ln] = v[e];
}

// Forward (from layer l0 up to k-1) or transposed (from k-1 down to l0)
// transform over 2^k positions x 2^log_nl lanes, up to 4 layers per launch.
int ntt_narrow_device(Ctx *ctx, hipStream_t st, void *d, const void *tw, int k, int l0, int log_nl, int transposed) {
  ScopedPhase ph(ctx, "ntt_narrow", st);
  int done = 0;
  const int total = k - l0;
  // Layers per launch: each thread runs F layers of serial products, so keep
  // at least 2^18 threads in flight (fewer layers per launch when small),
  // but at least two layers: below that size a launch costs more than a
  // thread's two layers of products (independent within a layer).
  const int f_par = std::max(2, std::min(4, k + log_nl - 18));
  while (done < total) {
    const int F = std::min(f_par, total - done);
    // forward passes run l0, l0+4, ...; transposed ones run the same groups from the top
    const int l = transposed ? k - done - F : l0 + done;
    const uint64_t n_threads = 1ull << (k - F + log_nl);
    const dim3 grid((unsigned)((n_threads + MUL_THREADS - 1) / MUL_THREADS)), block(MUL_THREADS);
    f128 *dd = (f128 *)d;
    const f128 *t = (const f128 *)tw;
    // wavefront-uniform twiddles: the wavefront's 2^(6 - log_nl) groups share
    // their block, and every wavefront is whole
    const bool ut = log_nl <= 6 && k - l - F >= 6 - log_nl && k - F + log_nl >= 6;
#define NARROW_LAUNCH(FF, TR, U) \
  hipLaunchKernelGGL((ntt_pass_narrow_kernel<FF, TR, U>), grid, block, 0, st, dd, t, k, l, log_nl, n_threads)
#define NARROW_CASE(FF)                                                            \
  case FF:                                                                        \
    if (transposed) {                                                             \
      if (ut) NARROW_LAUNCH(FF, true, true); else NARROW_LAUNCH(FF, true, false);  \
    } else {                                                                      \
      if (ut) NARROW_LAUNCH(FF, false, true); else NARROW_LAUNCH(FF, false, false); \
    }                                                                             \
    break;
    switch (F) {
      NARROW_CASE(1)
      NARROW_CASE(2)
      NARROW_CASE(3)
      NARROW_CASE(4)
    }
#undef NARROW_CASE
#undef NARROW_LAUNCH
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
    done += F;
  }
  return FLOCK_OK;
}

// ---------------------------------------------------------------- proof of work
// Hand a proof-of-work search's result to the host and re-arm it: out =
// best, best = none (so the next search needs no upload for it).
__global__ void pow_take_kernel(unsigned long long *__restrict__ best, unsigned long long *__restrict__ out,
                                uint32_t *flag, uint32_t seq) {
  if (threadIdx.x == 0) {
    out[0] = best[0];
    best[0] = ~0ull;
    if (seq) {
      __threadfence_system();
      __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

int pow_take_device(Ctx *ctx, hipStream_t st, void *d_best, void *out, uint32_t seq) {
  hipLaunchKernelGGL(pow_take_kernel, dim3(1), dim3(64), 0, st, (unsigned long long *)d_best,
                     (unsigned long long *)out, ctx->readback_flag_dev(), seq);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// ---------------------------------------------------------------- readback signal
// Publish seq to the host: the stream runs this after every earlier kernel
// and copy has finished, so their results in host memory are complete.
__global__ void signal_kernel(uint32_t *flag, uint32_t seq) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

int signal_device(Ctx *ctx, hipStream_t st, void *d_flag, uint32_t seq) {
  hipLaunchKernelGGL(signal_kernel, dim3(1), dim3(64), 0, st, (uint32_t *)d_flag, seq);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// ---------------------------------------------------------------- gathers
__global__ void gather_kernel(const uint4 *__restrict__ src, const uint64_t *__restrict__ idx, uint32_t n,
                              uint32_t unit_u4, uint4 *__restrict__ out) {
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (uint64_t)n * unit_u4) return;
  const uint64_t i = t / unit_u4, j = t % unit_u4;
  out[t] = src[idx[i] * unit_u4 + j];
}

int gather_device(Ctx *ctx, hipStream_t st, const void *src, const uint64_t *d_idx, uint32_t n, uint32_t unit_bytes,
                  void *out) {
  const uint32_t u4 = unit_bytes / 16;
  const uint64_t total = (uint64_t)n * u4;
  if (!total) return FLOCK_OK;
  hipLaunchKernelGGL(gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, (const uint4 *)src, d_idx,
                     n, u4, (uint4 *)out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// replicate + scatter helpers
__global__ void scatter_kernel(f128 *__restrict__ dst, const uint64_t *__restrict__ idx, const f128 *__restrict__ val,
                               uint32_t n) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[idx[t]] = val[t];
}

int scatter_device(Ctx *ctx, hipStream_t st, void *dst, const uint64_t *d_idx, const void *d_val, uint32_t n) {
  if (!n) return FLOCK_OK;
  hipLaunchKernelGGL(scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, (f128 *)dst, d_idx, (const f128 *)d_val, n);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// out[j] = xor over replicas r of in[j + r * n]  (transpose of replicate)
__global__ void fold_replicas_kernel(const f128 *__restrict__ in, uint64_t n, uint32_t reps, f128 *__restrict__ out) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
    f128 s = f128_zero();
    for (uint32_t r = 0; r < reps; ++r) f128_xor_into(s, in[j + r * n]);
    out[j] = s;
  }
}

int fold_replicas_device(Ctx *ctx, hipStream_t st, const void *in, uint64_t n, uint32_t reps, void *out) {
  hipLaunchKernelGGL(fold_replicas_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, st, (const f128 *)in, n, reps,
                     (f128 *)out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock
