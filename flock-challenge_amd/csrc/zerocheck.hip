// Zerocheck kernels (crates/flock-core/src/zerocheck.rs:482-1364).
//
// Round 1 (univariate skip), product part.  Row x of the packed witness is
// one 64-bit word: its bits are the values on S = {0..63} of a degree-<64
// polynomial over GF(2^8); the round-1 message needs its values on
// Lambda = {64..127}.  With r_rest = [phi8(F7), phi8(53), phi8(B5),
// beta_1..beta_4, r_outer...] the first three eq coordinates are images of
// GF(2^8) constants, so the eq weight of the low three row bits, eps(x_lo),
// lives in GF(2^8) and
//   Y_j(x_mid, x_hi) = sum_{x_lo} eps(x_lo) * ahat_j * bhat_j    (in GF(2^8))
// is formed without touching GF(2^128).  zc_y_kernel writes these bytes;
// then P^AB(lambda_j) = sum_{x_mid} E4(x_mid) sum_t phi8(2^t)
// sum_{x_hi} Ehi(x_hi) bit_t(Y_j(x_mid, x_hi)), whose inner sums are a
// bit-column fold of Y with the weights Ehi (bitfold.hip).  This matches the
// reference round1_naive (zerocheck/univariate_skip.rs:84-138) exactly: every
// step is a regrouping of the same exact field sum.  P^C is linear in the
// bits of c and is computed from a bit-column fold of c.
//
// Extension by byte tables: ahat over all 64 lambdas is the XOR of 8 table
// rows, one per byte of the row.  Thread (x_mid, q4) handles lambdas
// 16 q4 .. 16 q4 + 15 (one 16-byte LDS read per byte).  Entry (v, p, q4) is
// at chunk (v * 8 + p) * 4 + q4, so its bank group (mod 16) is
// 4 (p mod 4) + q4; a lane visits byte positions p = step ^ (x_mid mod 4),
// which makes the 16 lanes served together hit 16 distinct bank groups.
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

__device__ __forceinline__ uint4 ext8(uint64_t row, const uint4 *__restrict__ s_ext, uint32_t q4, uint32_t sw) {
  const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
  uint4 e = make_uint4(0, 0, 0, 0);
#pragma unroll
  for (uint32_t step = 0; step < 8; ++step) {
    const uint32_t p = step ^ sw;
    const uint32_t w = step < 4 ? lo : hi;
    const uint32_t v = __builtin_amdgcn_ubfe(w, 8 * (p & 3), 8);
    const uint4 t = s_ext[(v * 8 + p) * 4 + q4];
    e.x ^= t.x;
    e.y ^= t.y;
    e.z ^= t.z;
    e.w ^= t.w;
  }
  return e;
}

constexpr int ZCY_THREADS = 1024;

__global__ __launch_bounds__(ZCY_THREADS) void zc_y_kernel(const uint64_t *__restrict__ a,
                                                           const uint64_t *__restrict__ b,
                                                           const ZcTables *__restrict__ T, uint64_t n_blocks,
                                                           uint4 *__restrict__ Y) {
  __shared__ uint4 s_ext[256 * 8 * 4];
  __shared__ uint16_t s_log[256];
  __shared__ uint8_t s_exp[3 * 1024 + 4];
  __shared__ uint32_t s_leps[8];
  const uint32_t tid = threadIdx.x;
  {
    const uint4 *src = (const uint4 *)T->ext8;
    for (uint32_t i = tid; i < 256 * 8 * 4; i += ZCY_THREADS) s_ext[i] = src[i];
  }
  if (tid < 256) s_log[tid] = T->log8[tid];
  for (uint32_t i = tid; i < 3 * 1024 + 4; i += ZCY_THREADS) s_exp[i] = T->exp8[i];
  if (tid < 8) s_leps[tid] = T->log_eps[tid];
  __syncthreads();
  const uint32_t q4 = tid & 3, xm = (tid >> 2) & 15, xo = tid >> 6, sw = xm & 3;
  for (uint64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const uint64_t xh = blk * 16 + xo;
    const uint64_t row0 = (xh << 7) | ((uint64_t)xm << 3);
    uint4 y = make_uint4(0, 0, 0, 0);
#pragma unroll 1
    for (uint32_t xl = 0; xl < 8; ++xl) {
      const uint4 ea = ext8(a[row0 + xl], s_ext, q4, sw);
      const uint4 eb = ext8(b[row0 + xl], s_ext, q4, sw);
      const uint32_t le = s_leps[xl];
      y.x ^= f8_prod4(ea.x, eb.x, le, s_log, s_exp);
      y.y ^= f8_prod4(ea.y, eb.y, le, s_log, s_exp);
      y.z ^= f8_prod4(ea.z, eb.z, le, s_log, s_exp);
      y.w ^= f8_prod4(ea.w, eb.w, le, s_log, s_exp);
    }
    Y[((xh << 4) + xm) * 4 + q4] = y;
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
                     const void *d_ehi, void *d_q) {
  ScopedPhase ph(ctx, "zc_round1", st);
  const uint64_t n_hi = 1ull << (m - 13);
  if (n_hi < 16) {
    ctx->last_error = "zerocheck: m too small";
    return FLOCK_E_SHAPE;
  }
  const uint64_t n_blocks = n_hi / 16;
  uint4 *Y = (uint4 *)ctx->get_scratch("zc.round1.y", (size_t)n_hi * 16 * 64);
  if (!Y) return FLOCK_E_OOM;
  const unsigned grid = (unsigned)(n_blocks < 1024 ? n_blocks : 1024);
  hipLaunchKernelGGL(zc_y_kernel, dim3(grid), dim3(ZCY_THREADS), 0, st, (const uint64_t *)a, (const uint64_t *)b,
                     (const ZcTables *)d_tables, n_blocks, Y);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  // Q[4 x_mid + q4][8 j + t] = sum_{x_hi} Ehi(x_hi) bit_t(Y_{16 q4 + j}(x_mid, x_hi))
  return bitfold_device(ctx, st, Y, 6, n_hi, d_ehi, nullptr, d_q);
}

size_t zc_tables_bytes() { return sizeof(ZcTables); }

}  // namespace flock
