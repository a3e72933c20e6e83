// Bit-column folds: Q_k[il][r] = sum_ih bit_r(elem[il + IL * ih]) * hi_k[ih]
// for 128-bit elements, one or two weight vectors hi_k.
//
// Three prover steps are this computation with eq tables as weights:
//  - the ring-switch fold s_hat_v[r] = sum_i bit_r(z[i]) eq(x, i)
//    (fold_1b_rows, crates/flock-core/src/pcs/ring_switch.rs:1568-1618), with
//    eq(x, i) = eq_lo(i mod 16) eq_hi(i / 16);
//  - the lincheck batch fold z_vec[i] = sum_blk eq(x_outer, blk) bit_i(z_blk)
//    (partial_fold_packed_z, crates/flock-core/src/lincheck.rs:660-845), with
//    elements = 16-byte chunks of a 2 KiB block (IL = 128);
//  - the linear part of zerocheck round 1, sum_x eq(x) bit_s(c[x])
//    (round1_naive's P^C, zerocheck/univariate_skip.rs:84-138 regrouped).
// The caller combines Q with the small eq_lo factor.
//
// Method of four Russians over the weights: for 8 consecutive ih the 256
// XOR-combinations of hi_k[ih..ih+7] form a table shared by every (il, r).
// Thread (il, b) owns bits r = 8b..8b+7 of elements il: it takes byte b of
// the 8 elements, transposes the 8x8 bit block so byte c holds bit 8b+c of
// the 8 elements, and adds table[byte c] into accumulator c.  Per element
// that is 16 table reads of 16 B instead of 128 masked 128-bit XORs.  A group
// of 8 ih covers 8 x IL x 16 B of contiguous input, so reads coalesce.
#include <hip/hip_runtime.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"
#include "gf128_dev.h"
#include "prove_kernels.h"

namespace flock {

// Transpose of an 8x8 bit matrix held as bytes (bit 8r + c <-> bit 8c + r).
__device__ __forceinline__ uint64_t transpose8x8(uint64_t x) {
  uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x = x ^ t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x = x ^ t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x = x ^ t ^ (t << 28);
  return x;
}

// The grid runs about four wavefronts per SIMD, so the kernel may use up to
// 128 VGPRs to keep loads and table reads in flight.
template <int NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void bitfold_kernel(const uint32_t *el, int il_log, const f128 *hi0, const f128 *hi1,
                                                      uint64_t n_ih, uint64_t ih_per_wg, f128 *__restrict__ partial) {
  __shared__ f128 tab[2][NC][256];
  const uint32_t t = threadIdx.x, il = t >> 4, b = t & 15;
  const uint64_t ilg = (uint64_t)blockIdx.y * 16 + il;
  const uint64_t ih_begin = (uint64_t)blockIdx.x * ih_per_wg;
  uint64_t ih_end = ih_begin + ih_per_wg;
  if (ih_end > n_ih) ih_end = n_ih;
  // Table-build role: NC == 2 -> claim t >> 7 builds entries v, v | 128 (v = t & 127);
  // NC == 1 -> entry t.  t >> 7 is constant over a wavefront.
  const uint32_t bk = NC == 2 ? (t >> 7) : 0u;
  const uint32_t bv = NC == 2 ? (t & 127u) : t;
  // The 8 weights of a group are 32 dwords; lane j (mod 32) of each wavefront
  // loads dword j and the table build reads them back with v_readlane.  Vector
  // loads, unlike scalar ones, stay in flight across the LDS barrier.
  const uint32_t *hw = (const uint32_t *)(bk ? hi1 : hi0) + (t & 31u);
  const uint32_t sh = 8 * (b & 3);
  // byte offset of this lane's dword within a row of 2^il_log elements
  const uint32_t lane_off = ((uint32_t)ilg << 4) + 4u * (b >> 2);
  const uint32_t row_shift = (uint32_t)il_log + 4u;

  f128 acc[NC][8];
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[k][c] = f128_zero();

  auto load = [&](uint64_t ih, uint32_t (&w)[8], uint32_t &hv) {
#pragma unroll
    for (int j = 0; j < 8; ++j) w[j] = *(const uint32_t *)((const char *)el + ((ih + j) << row_shift) + lane_off);
    hv = hw[4 * ih];
  };
  auto group = [&](const uint32_t (&w)[8], uint32_t hv, int buf) {
    // (1) this group's table: entry bv = sum_j bit_j(bv) h_j
    {
      f128 s = f128_zero();
#pragma unroll
      for (int j = 0; j < (NC == 2 ? 7 : 8); ++j) {
        const uint32_t m = 0u - ((bv >> j) & 1u);
        // (h & m) ^ s in one three-input op per dword
        s.x = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 0), m, s.x, 0x6a);
        s.y = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 1), m, s.y, 0x6a);
        s.z = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 2), m, s.z, 0x6a);
        s.w = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 3), m, s.w, 0x6a);
      }
      tab[buf][bk][bv] = s;
      if (NC == 2)
        tab[buf][bk][bv | 128u] =
            make_uint4(s.x ^ __builtin_amdgcn_readlane(hv, 28), s.y ^ __builtin_amdgcn_readlane(hv, 29),
                       s.z ^ __builtin_amdgcn_readlane(hv, 30), s.w ^ __builtin_amdgcn_readlane(hv, 31));
    }
    // (2) byte b of the 8 elements
    uint32_t lo = 0, hi = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const uint32_t byte = (w[j] >> sh) & 255u;
      if (j < 4)
        lo |= byte << (8 * j);
      else
        hi |= byte << (8 * (j - 4));
    }
    __syncthreads();
    const uint64_t x = transpose8x8((uint64_t)lo | ((uint64_t)hi << 32));
    // reads issued in batches of B before their XORs (B = 4 for two weight
    // vectors, whose 64 accumulator registers leave less room)
    constexpr int B = NC == 2 ? 4 : 8;
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int c0 = 0; c0 < 8; c0 += B) {
        f128 e[B];
#pragma unroll
        for (int c = 0; c < B; ++c) e[c] = tab[buf][k][(uint32_t)(x >> (8 * (c0 + c))) & 255u];
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int c = 0; c < B; ++c) f128_xor_into(acc[k][c0 + c], e[c]);
      }
  };

  // Two register sets alternate: the next group's loads are issued before
  // this group's work, so their latency overlaps it.  The last group
  // reloads itself (in bounds, unused).
  // The empty asm statements keep the compiler from moving the loads into
  // the other half of the loop (the input pointers are deliberately not
  // __restrict__, so the memory clobber applies to them).
  uint32_t wa[8], wb[8], ha, hb;
  load(ih_begin, wa, ha);
  for (uint64_t ih0 = ih_begin;;) {
    load(ih0 + 8 < ih_end ? ih0 + 8 : ih0, wb, hb);
    asm volatile("" ::: "memory");
    group(wa, ha, 0);
    ih0 += 8;
    if (ih0 >= ih_end) break;
    load(ih0 + 8 < ih_end ? ih0 + 8 : ih0, wa, ha);
    asm volatile("" ::: "memory");
    group(wb, hb, 1);
    ih0 += 8;
    if (ih0 >= ih_end) break;
  }
  f128 *out = partial + (uint64_t)(blockIdx.y * gridDim.x + blockIdx.x) * NC * 2048;
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int c = 0; c < 8; ++c) out[k * 2048 + il * 128 + 8 * b + c] = acc[k][c];
}

// out[y][w] = xor_x in[(y * nx + x) * width + w]: thread per column, rows split
// into chunks of at most `rows_per` per workgroup row.
__global__ __launch_bounds__(256) void xor_reduce_kernel(const f128 *__restrict__ in, uint32_t nx, uint32_t width,
                                                         uint32_t rows_per, f128 *__restrict__ out) {
  const uint32_t w = blockIdx.x * 256 + threadIdx.x;
  const uint32_t chunk = blockIdx.y, y = blockIdx.z;
  if (w >= width) return;
  const uint32_t x0 = chunk * rows_per;
  uint32_t x1 = x0 + rows_per;
  if (x1 > nx) x1 = nx;
  f128 s = f128_zero();
  for (uint32_t x = x0; x < x1; ++x) f128_xor_into(s, in[((uint64_t)y * nx + x) * width + w]);
  out[((uint64_t)y * gridDim.y + chunk) * width + w] = s;
}

// Reduce ny independent stacks of nx rows (width F128 each) to ny rows.
int xor_reduce_device(Ctx *ctx, hipStream_t st, const void *in, uint32_t ny, uint32_t nx, uint32_t width, void *out) {
  const f128 *src = (const f128 *)in;
  int pass = 0;
  while (nx > 1) {
    const uint32_t rows_per = nx > 16 ? 16 : nx;
    const uint32_t chunks = (nx + rows_per - 1) / rows_per;
    f128 *dst;
    if (chunks == 1) {
      dst = (f128 *)out;
    } else {
      dst = (f128 *)ctx->get_scratch(pass & 1 ? "xor_reduce.b" : "xor_reduce.a",
                                     (size_t)ny * chunks * width * sizeof(f128));
      if (!dst) return FLOCK_E_OOM;
    }
    hipLaunchKernelGGL(xor_reduce_kernel, dim3((width + 255) / 256, chunks, ny), dim3(256), 0, st, src, nx, width,
                       rows_per, dst);
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
    src = dst;
    nx = chunks;
    ++pass;
  }
  if (pass == 0) FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(out, in, (size_t)ny * width * sizeof(f128), hipMemcpyDeviceToDevice, st));
  return FLOCK_OK;
}

int bitfold_device(Ctx *ctx, hipStream_t st, const void *elems, int il_log, uint64_t n_ih, const void *hi0,
                   const void *hi1, void *out) {
  ScopedPhase ph(ctx, "bitfold", st);
  const int nc = hi1 ? 2 : 1;
  if (il_log < 4 || (n_ih & 7) || n_ih == 0) {
    ctx->last_error = "bitfold: needs IL >= 16 and a multiple of 8 weight rows";
    return FLOCK_E_SHAPE;
  }
  const uint32_t ny = 1u << (il_log - 4);
  // About 1024 workgroups in all, each a multiple of 8 weight rows.
  uint64_t nx = 1024 / ny;
  if (nx < 1) nx = 1;
  uint64_t per = (n_ih + nx - 1) / nx;
  per = (per + 7) & ~7ull;
  nx = (n_ih + per - 1) / per;
  f128 *part = (f128 *)ctx->get_scratch("bitfold.partial", (size_t)ny * nx * nc * 2048 * sizeof(f128));
  if (!part) return FLOCK_E_OOM;
  if (nc == 2)
    hipLaunchKernelGGL(bitfold_kernel<2>, dim3((unsigned)nx, ny), dim3(256), 0, st, (const uint32_t *)elems, il_log,
                       (const f128 *)hi0, (const f128 *)hi1, n_ih, per, part);
  else
    hipLaunchKernelGGL(bitfold_kernel<1>, dim3((unsigned)nx, ny), dim3(256), 0, st, (const uint32_t *)elems, il_log,
                       (const f128 *)hi0, (const f128 *)hi0, n_ih, per, part);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  // partial layout [y][x][k][2048] -> out [y][k][2048]
  return xor_reduce_device(ctx, st, part, ny, (uint32_t)nx, (uint32_t)(nc * 2048), out);
}

}  // namespace flock
