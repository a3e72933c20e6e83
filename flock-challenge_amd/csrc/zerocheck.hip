// Zerocheck kernels (crates/flock-core/src/zerocheck.rs:482-1364).
//
// Round 1 (univariate skip), product part.  Row x of the packed witness is
// one 64-bit word: its bits are the values on S = {0..63} of a degree-<64
// polynomial over GF(2^8); the round-1 message needs its values on
// Lambda = {64..127}.  With r_rest = [phi8(F7), phi8(53), phi8(B5),
// beta_1..beta_4, r_outer...] the first three eq coordinates are images of
// GF(2^8) constants, so the eq weight of the low three row bits, eps(x_lo),
// lives in GF(2^8) and
//   Y_j = sum_{x_lo} eps(x_lo) * ahat_j * bhat_j        (all in GF(2^8))
// is formed without touching GF(2^128).  The next four coordinates are fixed
// constants E4(x_mid); per outer index x_hi the workgroup scales the 16 x 8
// table E4(x_mid) * phi8(2^t) by the outer weight Ehi(x_hi) and folds Y_j in
// with two nibble lookups.  P^AB(lambda_j) = sum over everything.  This
// matches the reference round1_naive (zerocheck/univariate_skip.rs:84-138)
// exactly: every step is a regrouping of the same exact field sum.  P^C is
// linear in the bits of c and is computed from a bit-column fold (bitfold.hip).
//
// Thread layout: q4 = t & 3 picks lambda 16 q4 .. 16 q4 + 15 (one uint4 of
// extension bytes per nibble lookup), xs = (t >> 2) & 3 picks rows
// x_lo = 2 xs, 2 xs + 1, x_mid = t >> 4.  The four xs lanes of a (x_mid, q4)
// XOR their partial Y, then each keeps four lambdas for the F128 accumulation.
#include <hip/hip_runtime.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"
#include "gf128_dev.h"
#include "prove_kernels.h"
#include "zc_tables.h"

namespace flock {

__device__ __forceinline__ uint32_t f8_prod4(uint32_t ea, uint32_t eb, uint32_t le, const uint16_t *__restrict__ lg,
                                             const uint8_t *__restrict__ ex) {
  uint32_t y = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t va = (ea >> (8 * k)) & 255u, vb = (eb >> (8 * k)) & 255u;
    y |= (uint32_t)ex[lg[va] + lg[vb] + le] << (8 * k);
  }
  return y;
}

__device__ __forceinline__ uint4 ext16(uint64_t row, const uint4 *__restrict__ s_ext, uint32_t q4) {
  uint4 e = make_uint4(0, 0, 0, 0);
#pragma unroll
  for (int p = 0; p < 16; ++p) {
    const uint4 v = s_ext[(p * 16 + (uint32_t)((row >> (4 * p)) & 15)) * 4 + q4];
    e.x ^= v.x;
    e.y ^= v.y;
    e.z ^= v.z;
    e.w ^= v.w;
  }
  return e;
}

__global__ __launch_bounds__(256) void zc_round1_kernel(const uint64_t *__restrict__ a, const uint64_t *__restrict__ b,
                                                        const ZcTables *__restrict__ T, const f128 *__restrict__ ehi,
                                                        uint64_t n_hi, uint64_t hi_per_wg, f128 *__restrict__ partial) {
  __shared__ uint4 s_ext[16 * 16 * 4];
  __shared__ uint16_t s_log[256];
  __shared__ uint8_t s_exp[3 * 1024 + 4];
  __shared__ f128 s_tab[512];     // byte table of Ehi(x_hi)
  __shared__ f128 s_pt[16][8];    // E4 * phi8(2^t), unscaled
  __shared__ f128 s_pts[16][8];   // scaled by Ehi(x_hi)
  __shared__ f128 s_ptn[16][2][16];
  __shared__ uint32_t s_leps[8];
  const uint32_t tid = threadIdx.x;
  {
    const uint4 *src = (const uint4 *)T->ext;
    for (uint32_t i = tid; i < 16 * 16 * 4; i += 256) s_ext[i] = src[i];
  }
  s_log[tid] = T->log8[tid];
  for (uint32_t i = tid; i < 3 * 1024 + 4; i += 256) s_exp[i] = T->exp8[i];
  if (tid < 128) s_pt[tid >> 3][tid & 7] = *(const f128 *)T->pt[tid >> 3][tid & 7];
  if (tid < 8) s_leps[tid] = T->log_eps[tid];
  const uint32_t q4 = tid & 3, xs = (tid >> 2) & 3, x_mid = tid >> 4;
  f128 acc[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) acc[k] = f128_zero();
  __syncthreads();

  const uint64_t hi0 = (uint64_t)blockIdx.x * hi_per_wg;
  uint64_t hi1 = hi0 + hi_per_wg;
  if (hi1 > n_hi) hi1 = n_hi;
  for (uint64_t xh = hi0; xh < hi1; ++xh) {
    // (1) two rows -> partial GF(2^8) Y for 16 lambdas
    const uint64_t row0 = (xh << 7) | ((uint64_t)x_mid << 3) | (2 * xs);
    uint4 y = make_uint4(0, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 2; ++r) {
      const uint4 ea = ext16(a[row0 + r], s_ext, q4);
      const uint4 eb = ext16(b[row0 + r], s_ext, q4);
      const uint32_t le = s_leps[2 * xs + r];
      y.x ^= f8_prod4(ea.x, eb.x, le, s_log, s_exp);
      y.y ^= f8_prod4(ea.y, eb.y, le, s_log, s_exp);
      y.z ^= f8_prod4(ea.z, eb.z, le, s_log, s_exp);
      y.w ^= f8_prod4(ea.w, eb.w, le, s_log, s_exp);
    }
    // (2) combine the four xs lanes (lanes differing in bits 2 and 3)
#pragma unroll
    for (int d = 4; d <= 8; d <<= 1) {
      y.x ^= __shfl_xor(y.x, d);
      y.y ^= __shfl_xor(y.y, d);
      y.z ^= __shfl_xor(y.z, d);
      y.w ^= __shfl_xor(y.w, d);
    }
    const uint32_t yw = xs == 0 ? y.x : xs == 1 ? y.y : xs == 2 ? y.z : y.w;
    // (3) scale E4 * phi8(2^t) by Ehi(x_hi) and expand to nibble tables
    __syncthreads();
    if (tid < 64) f128_build_bytetab(ehi[xh], s_tab, tid);
    __syncthreads();
    if (tid < 128) s_pts[tid >> 3][tid & 7] = f128_mul_bytetab(s_pt[tid >> 3][tid & 7], s_tab);
    __syncthreads();
    for (uint32_t e = tid; e < 512; e += 256) {
      const uint32_t xm = e >> 5, h = (e >> 4) & 1, v = e & 15;
      f128 s = f128_zero();
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if ((v >> t) & 1) f128_xor_into(s, s_pts[xm][4 * h + t]);
      s_ptn[xm][h][v] = s;
    }
    __syncthreads();
    // (4) accumulate lambda = 16 q4 + 4 xs + k
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t yb = (yw >> (8 * k)) & 255u;
      f128_xor_into(acc[k], s_ptn[x_mid][0][yb & 15]);
      f128_xor_into(acc[k], s_ptn[x_mid][1][yb >> 4]);
    }
  }
  // reduce over the 16 x_mid values; reuse s_ext (16 KiB = 1024 f128) as scratch
  __syncthreads();
  f128 *red = (f128 *)s_ext;
#pragma unroll
  for (int k = 0; k < 4; ++k) red[k * 256 + tid] = acc[k];
  __syncthreads();
  if (tid < 64) {
    // output lambda j = tid: q4 = j >> 4, xs = (j >> 2) & 3, k = j & 3
    const uint32_t j = tid, jq4 = j >> 4, jxs = (j >> 2) & 3, k = j & 3;
    f128 s = f128_zero();
    for (uint32_t xm = 0; xm < 16; ++xm) f128_xor_into(s, red[k * 256 + (xm << 4) + (jxs << 2) + jq4]);
    partial[(uint64_t)blockIdx.x * 64 + j] = s;
  }
}

// XOR-sum `count` rows of `width` F128 each: out[w] = xor_r in[r * width + w].
__global__ void xor_rows_kernel(const f128 *__restrict__ in, uint64_t count, uint32_t width, f128 *__restrict__ out) {
  const uint32_t w = blockIdx.x;
  f128 s = f128_zero();
  for (uint64_t r = threadIdx.x; r < count; r += blockDim.x) f128_xor_into(s, in[r * width + w]);
  __shared__ f128 red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t st = blockDim.x / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = f128_xor(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[w] = red[0];
}

int xor_rows_device(Ctx *ctx, hipStream_t st, const void *in, uint64_t count, uint32_t width, void *out) {
  hipLaunchKernelGGL(xor_rows_kernel, dim3(width), dim3(256), 0, st, (const f128 *)in, count, width, (f128 *)out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int zc_round1_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, int m, const void *d_tables,
                     const void *d_ehi, void *d_out64) {
  ScopedPhase ph(ctx, "zc_round1", st);
  const uint64_t n_hi = 1ull << (m - 6 - 7);
  uint64_t wgs = 2048;
  if (wgs > n_hi) wgs = n_hi;
  const uint64_t per = (n_hi + wgs - 1) / wgs;
  wgs = (n_hi + per - 1) / per;
  f128 *part = (f128 *)ctx->get_scratch("zc.round1.partial", wgs * 64 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  hipLaunchKernelGGL(zc_round1_kernel, dim3((unsigned)wgs), dim3(256), 0, st, (const uint64_t *)a,
                     (const uint64_t *)b, (const ZcTables *)d_tables, (const f128 *)d_ehi, n_hi, per, part);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return xor_reduce_device(ctx, st, part, 1, (uint32_t)wgs, 64, d_out64);
}

size_t zc_tables_bytes() { return sizeof(ZcTables); }

// ---------------------------------------------------------------- fold at z
// A[x] = sum_s bit_s(a[x]) L_s(z) (zerocheck/multilinear.rs:598-625) through
// eight byte tables of the 64 Lagrange weights, for a and b together.
__global__ __launch_bounds__(256) void fold_at_z_kernel(const uint64_t *__restrict__ a,
                                                        const uint64_t *__restrict__ b,
                                                        const f128 *__restrict__ w, uint64_t rows,
                                                        f128 *__restrict__ A, f128 *__restrict__ B) {
  __shared__ f128 tab[8][256];
  for (uint32_t e = threadIdx.x; e < 2048; e += 256) {
    const uint32_t p = e >> 8, v = e & 255;
    f128 s = f128_zero();
    for (int t = 0; t < 8; ++t)
      if ((v >> t) & 1) f128_xor_into(s, w[8 * p + t]);
    tab[p][v] = s;
  }
  __syncthreads();
  for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < rows; x += (uint64_t)gridDim.x * 256) {
    const uint64_t ra = a[x], rb = b[x];
    f128 sa = f128_zero(), sb = f128_zero();
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      f128_xor_into(sa, tab[p][(ra >> (8 * p)) & 255]);
      f128_xor_into(sb, tab[p][(rb >> (8 * p)) & 255]);
    }
    A[x] = sa;
    B[x] = sb;
  }
}

int fold_at_z_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, const void *d_w64, uint64_t rows, void *A,
                     void *B) {
  ScopedPhase ph(ctx, "zc_fold_at_z", st);
  hipLaunchKernelGGL(fold_at_z_kernel, dim3(4096), dim3(256), 0, st, (const uint64_t *)a, (const uint64_t *)b,
                     (const f128 *)d_w64, rows, (f128 *)A, (f128 *)B);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock
