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
// is formed without touching GF(2^128).  zc_y_kernel writes these values;
// then P^AB(lambda_j) = sum_{x_mid} E4(x_mid) sum_t phi8(2^t)
// sum_{x_hi} Ehi(x_hi) bit_t(Y_j(x_mid, x_hi)), whose inner sums are a
// bit-column fold of Y with the weights Ehi (bitfold.hip).  This matches the
// reference round1_naive (zerocheck/univariate_skip.rs:84-138) exactly: every
// step is a regrouping of the same exact field sum.  P^C is linear in the
// bits of c and is computed from a bit-column fold of c.
//
// Bitsliced: a lane handles 32 lambdas of one (x_mid, x_hi) and keeps each
// GF(2^8) value as 8 bit-planes of 32 lanes.  The extension of a row is the
// XOR of 8 table entries (one per row byte, 8 planes each); the product
// ahat * bhat is a schoolbook AND/XOR network (64 partial products, then
// x^8 .. x^14 folded back), and the constant eps(x_lo) is a fixed XOR
// network, with no table lookup at data-dependent addresses.  The bit-column
// fold reads Y's bits in that plane order (element 4 x_mid + 2 h + t / 4,
// bit 32 (t mod 4) + lambda mod 32); the host combination follows it.
#include <hip/hip_runtime.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"
#include "transcript_dev.h"
#include "gf128_dev.h"
#include "prove_kernels.h"
#include "zc_tables.h"

namespace flock {

// GF(2^8) (AES polynomial 0x11B) at compile time: the eq weights eps(x_lo)
// of the three low row bits are products of the small challenges
// (univariate_skip_optimized.rs:103-106), fixed constants.
__host__ __device__ constexpr uint8_t f8_mul_const(uint8_t a, uint8_t b) {
  uint8_t r = 0;
  for (int i = 0; i < 8; ++i) {
    if (b & 1) r ^= a;
    b >>= 1;
    a = (a & 0x80) ? (uint8_t)((a << 1) ^ 0x1B) : (uint8_t)(a << 1);
  }
  return r;
}
__host__ __device__ constexpr uint8_t zc_eps(int xl) {
  const uint8_t c[3] = {0xF7, 0x53, 0xB5};
  uint8_t e = 1;
  for (int i = 0; i < 3; ++i) e = f8_mul_const(e, ((xl >> i) & 1) ? c[i] : (uint8_t)(c[i] ^ 1));
  return e;
}
// x^k mod 0x11B for k = 8..14 (the reduction of a bitsliced product)
__host__ __device__ constexpr uint8_t f8_xpow(int k) {
  uint8_t v = 1;
  for (int i = 0; i < k; ++i) v = f8_mul_const(v, 2);
  return v;
}

// acc ^= terms[i] for every i with bit i of `mask` (a compile-time constant
// after unrolling), two terms per three-input XOR.
template <int N>
__device__ __forceinline__ uint32_t xor_masked(uint32_t acc, const uint32_t (&terms)[N], uint32_t mask) {
  uint32_t pend = 0;
  bool has = false;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if ((mask >> i) & 1) {
      if (has) {
        acc = xor3_u32(acc, pend, terms[i]);
        has = false;
      } else {
        pend = terms[i];
        has = true;
      }
    }
  return has ? acc ^ pend : acc;
}

// Bitsliced GF(2^8) product of 32 lanes: plane t of a dword holds bit t of
// the 32 values.  Schoolbook partial products, then x^k (k >= 8) folded back.
__device__ __forceinline__ void f8_bs_mul(const uint32_t (&A)[8], const uint32_t (&B)[8], uint32_t (&C)[8]) {
  uint32_t c[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) c[k] = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) c[i + j] = __builtin_amdgcn_bitop3_b32(A[i], B[j], c[i + j], 0x6a);  // (A & B) ^ c
  uint32_t hi[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) hi[k] = c[8 + k];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    uint32_t mask = 0;
#pragma unroll
    for (int k = 8; k < 15; ++k) mask |= (uint32_t)((f8_xpow(k) >> t) & 1) << (k - 8);
    C[t] = xor_masked(c[t], hi, mask);
  }
}

// Bitsliced extension of one row for this lane's 32 lambdas: 8 byte
// positions, each one entry (2 x 16 B) of the table, XORed into 8 planes.
// Lane rotation: step s reads byte position p = s ^ r (r = x_mid mod 8), so
// the 16 lanes the LDS serves together (8 x_mid x 2 halves) hit 16 distinct
// bank groups 2 p + h.  Addresses are one v_perm each: byte 0 = (2 p + h) * 16
// from the lane's offset registers, byte 1 = the row byte, byte 2 = k.
__device__ __forceinline__ void ext_bs(uint64_t row, const char *__restrict__ tab, const uint32_t (&O)[4],
                                       uint32_t sel_dw, uint32_t sel_b, uint32_t (&P)[8]) {
  // byte s of (w0, w1) = byte s ^ r of the row
  const uint32_t lo = (uint32_t)row, hi = (uint32_t)(row >> 32);
  const uint32_t w0 = __builtin_amdgcn_perm(0u, __builtin_amdgcn_perm(hi, lo, sel_dw), sel_b);
  const uint32_t w1 = __builtin_amdgcn_perm(0u, __builtin_amdgcn_perm(lo, hi, sel_dw), sel_b);
#pragma unroll
  for (int t = 0; t < 8; ++t) P[t] = 0;
#pragma unroll
  for (int s = 0; s < 8; s += 2) {
    uint4 e[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int st = s + u;
      const uint32_t word = st < 4 ? w0 : w1;
      const uint32_t base = (4u + (uint32_t)(st & 1)) | ((uint32_t)(st & 3) << 8) | (12u << 24);
      e[u][0] = *(const uint4 *)(tab + __builtin_amdgcn_perm(O[st >> 1], word, base | (12u << 16)));
      e[u][1] = *(const uint4 *)(tab + __builtin_amdgcn_perm(O[st >> 1], word, base | (6u << 16)));
    }
    P[0] = xor3_u32(P[0], e[0][0].x, e[1][0].x);
    P[1] = xor3_u32(P[1], e[0][0].y, e[1][0].y);
    P[2] = xor3_u32(P[2], e[0][0].z, e[1][0].z);
    P[3] = xor3_u32(P[3], e[0][0].w, e[1][0].w);
    P[4] = xor3_u32(P[4], e[0][1].x, e[1][1].x);
    P[5] = xor3_u32(P[5], e[0][1].y, e[1][1].y);
    P[6] = xor3_u32(P[6], e[0][1].z, e[1][1].z);
    P[7] = xor3_u32(P[7], e[0][1].w, e[1][1].w);
  }
}

constexpr int ZCY_THREADS = 1024;
constexpr int ZCY_XH_PER_WG = ZCY_THREADS / 32;  // 32 lanes (16 x_mid x 2 halves) per x_hi

// Y_lambda(x_mid, x_hi) = sum_{x_lo} eps(x_lo) ahat_lambda bhat_lambda, bitsliced:
// lane (x_mid, h) of x_hi computes lambdas 32 h .. 32 h + 31 and writes
// element 2 h + g (g = 0, 1: planes 4 g .. 4 g + 3) of (x_hi, x_mid).
__global__ __launch_bounds__(ZCY_THREADS) void zc_y_kernel(const uint64_t *__restrict__ a,
                                                           const uint64_t *__restrict__ b,
                                                           const ZcTables *__restrict__ T, uint64_t n_blocks,
                                                           uint4 *__restrict__ Y) {
  __shared__ uint4 s_ext[2 * 4096];
  const uint32_t tid = threadIdx.x;
  {
    const uint4 *src = (const uint4 *)T->extb;
    for (uint32_t i = tid; i < 2 * 4096; i += ZCY_THREADS) s_ext[i] = src[i];
  }
  __syncthreads();
  const uint32_t h = tid & 1, xm = (tid >> 1) & 15, xo = tid >> 5, r = xm & 7;
  uint32_t O[4];  // [off(2 i), off(2 i + 1), 1, 0], off(s) = ((s ^ r) * 2 + h) * 16
#pragma unroll
  for (uint32_t i = 0; i < 4; ++i)
    O[i] = ((((2 * i) ^ r) * 2 + h) << 4) | (((((2 * i + 1) ^ r) * 2 + h) << 4) << 8) | (1u << 16);
  const uint32_t sel_dw = (r & 4) ? 0x07060504u : 0x03020100u;
  uint32_t sel_b = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) sel_b |= (j ^ (r & 3)) << (8 * j);
  const char *tab = (const char *)s_ext;
  for (uint64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    const uint64_t xh = blk * ZCY_XH_PER_WG + xo;
    const uint64_t row0 = (xh << 7) | ((uint64_t)xm << 3);
    uint32_t Yp[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) Yp[t] = 0;
#pragma unroll
    for (int xl = 0; xl < 8; ++xl) {
      uint32_t A[8], B[8], C[8];
      ext_bs(a[row0 + xl], tab, O, sel_dw, sel_b, A);
      ext_bs(b[row0 + xl], tab, O, sel_dw, sel_b, B);
      f8_bs_mul(A, B, C);
      // Yp += eps(x_lo) C: plane t gathers the planes s with bit t of eps x^s
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        uint32_t mask = 0;
#pragma unroll
        for (int s2 = 0; s2 < 8; ++s2) mask |= (uint32_t)((f8_mul_const(zc_eps(xl), (uint8_t)(1u << s2)) >> t) & 1) << s2;
        Yp[t] = xor_masked(Yp[t], C, mask);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    uint4 *dst = Y + (((xh << 4) + xm) << 2) + 2 * h;
    dst[0] = make_uint4(Yp[0], Yp[1], Yp[2], Yp[3]);
    dst[1] = make_uint4(Yp[4], Yp[5], Yp[6], Yp[7]);
  }
}

// XOR-sum `count` rows of W F128 each, out[w] = xor_r in[r * W + w], in one
// workgroup (a 256-thread tree per column; four rows per thread in flight).  With seq != 0 it then publishes
// seq at *flag after out is written (as signal_kernel does), so a host
// waiting for out needs no further launch.  With fs (W = 2) the workgroup
// then absorbs out into the transcript state and samples *rho.
template <int W>
__global__ __launch_bounds__(256) void xor_rows_kernel(const f128 *__restrict__ in, uint64_t count, f128 *out,
                                                       uint32_t *flag, uint32_t seq, FsState *fs, f128 *rho,
                                                       uint32_t *dg) {
  __shared__ f128 red[W][256];
  const uint32_t tid = threadIdx.x;
  f128 s[W];
#pragma unroll
  for (int w = 0; w < W; ++w) s[w] = f128_zero();
  for (uint64_t r0 = 0; r0 < count; r0 += 1024) {
    f128 v[4][W];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t r = r0 + (uint64_t)u * 256 + tid;
#pragma unroll
      for (int w = 0; w < W; ++w) v[u][w] = r < count ? in[r * W + w] : f128_zero();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int w = 0; w < W; ++w) f128_xor_into(s[w], v[u][w]);
  }
#pragma unroll
  for (int w = 0; w < W; ++w) red[w][tid] = s[w];
  __syncthreads();
  for (uint32_t st = 128; st > 0; st >>= 1) {
    if (tid < st)
#pragma unroll
      for (int w = 0; w < W; ++w) red[w][tid] = f128_xor(red[w][tid], red[w][tid + st]);
    __syncthreads();
  }
  if (tid == 0) {
#pragma unroll
    for (int w = 0; w < W; ++w) out[w] = red[w][0];
    if (seq) {
      __threadfence_system();
      __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
  if constexpr (W == 2 || W == 3) {
    __shared__ FsState fss;
    if (fs) {  // thread 0 reads back what it wrote
      if (rho)
        fs_round_block(fs, fss, out, rho);
      else
        fs_observe2_block(fs, fss, out, dg);
    }
  }
}

int xor_rows_device(Ctx *ctx, hipStream_t st, const void *in, uint64_t count, uint32_t width, void *out,
                    uint32_t seq, const FsStep &fs) {
  const f128 *i = (const f128 *)in;
  f128 *o = (f128 *)out;
  uint32_t *flag = ctx->readback_flag_dev();
  if (fs.state && width != 2 && !(width == 3 && !fs.rho)) {
    ctx->last_error = "xor_rows: the transcript step takes two values";
    return FLOCK_E_ARG;
  }
  FsState *S = (FsState *)fs.state;
  f128 *R = (f128 *)fs.rho;
  uint32_t *D = (uint32_t *)fs.digest;
  switch (width) {
    case 1: hipLaunchKernelGGL(xor_rows_kernel<1>, dim3(1), dim3(256), 0, st, i, count, o, flag, seq, S, R, D); break;
    case 2: hipLaunchKernelGGL(xor_rows_kernel<2>, dim3(1), dim3(256), 0, st, i, count, o, flag, seq, S, R, D); break;
    case 3: hipLaunchKernelGGL(xor_rows_kernel<3>, dim3(1), dim3(256), 0, st, i, count, o, flag, seq, S, R, D); break;
    case 4: hipLaunchKernelGGL(xor_rows_kernel<4>, dim3(1), dim3(256), 0, st, i, count, o, flag, seq, S, R, D); break;
    default:
      ctx->last_error = "xor_rows: width out of range";
      return FLOCK_E_SHAPE;
  }
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// Y of every (x_mid, x_hi): depends on the rows only, not on the challenges,
// so it can run while the host derives them.
int zc_y_device(Ctx *ctx, hipStream_t st, const void *a, const void *b, int m, const void *d_tables) {
  ScopedPhase ph(ctx, "zc_round1", st);
  const uint64_t n_hi = 1ull << (m - 13);
  if (n_hi < 16) {
    ctx->last_error = "zerocheck: m too small";
    return FLOCK_E_SHAPE;
  }
  if (n_hi % ZCY_XH_PER_WG) {
    ctx->last_error = "zerocheck: x_hi count not a multiple of the workgroup's";
    return FLOCK_E_SHAPE;
  }
  const uint64_t n_blocks = n_hi / ZCY_XH_PER_WG;
  uint4 *Y = (uint4 *)ctx->get_scratch("zc.round1.y", (size_t)n_hi * 16 * 64);
  if (!Y) return FLOCK_E_OOM;
  const unsigned grid = (unsigned)(n_blocks < 1024 ? n_blocks : 1024);
  hipLaunchKernelGGL(zc_y_kernel, dim3(grid), dim3(ZCY_THREADS), 0, st, (const uint64_t *)a, (const uint64_t *)b,
                     (const ZcTables *)d_tables, n_blocks, Y);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

// Q[4 x_mid + q4][8 j + t] = sum_{x_hi} Ehi(x_hi) bit_t(Y_{16 q4 + j}(x_mid, x_hi)),
// from the Y of zc_y_device.
int zc_round1_device(Ctx *ctx, hipStream_t st, int m, const void *d_ehi, void *d_q) {
  ScopedPhase ph(ctx, "zc_round1", st);
  const uint64_t n_hi = 1ull << (m - 13);
  uint4 *Y = (uint4 *)ctx->get_scratch("zc.round1.y", (size_t)n_hi * 16 * 64);
  if (!Y) return FLOCK_E_OOM;
  return bitfold_device(ctx, st, Y, 6, n_hi, d_ehi, nullptr, d_q);
}

size_t zc_tables_bytes() { return sizeof(ZcTables); }

}  // namespace flock
