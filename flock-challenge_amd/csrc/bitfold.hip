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
// Method of four Russians over the weights: for 8 consecutive ih, the 16
// XOR-combinations of hi_k[ih..ih+3] and those of hi_k[ih+4..ih+7] form two
// small tables shared by every (il, r).  Thread (il, b) owns bits
// r = 8b..8b+7 of elements il: it takes byte b of the 8 elements, transposes
// the two 4x4 bit blocks of each half so that each nibble indexes one table
// for one bit column, and adds table_lo[n_lo] ^ table_hi[n_hi] into the
// accumulator of that column.  Each table is stored in 16 copies, copy L mod
// 16 read by lane L, so the 16 lanes the LDS serves together hit 16 distinct
// bank groups whatever the data.  Per element that is 32 table reads of 16 B
// instead of 128 masked 128-bit XORs.  The group's input words pass through
// LDS: each wavefront loads two full rows of the workgroup's 16 elements
// (256 B contiguous per load, no dword fetched twice) and each lane reads back
// its dword of the 8 rows.
#include <hip/hip_runtime.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"
#include "gf128_dev.h"
#include "prove_kernels.h"

namespace flock {

// In each 32-bit word (bytes r = 0..3, bit c of byte r = element r, bit c)
// transpose the two 4x4 bit blocks: afterwards bit c of byte r is element c
// (c < 4) bit r, and bit 4 + c of byte r is element c bit 4 + r.
__device__ __forceinline__ uint32_t transpose4x4_pair(uint32_t x) {
  uint32_t t = __builtin_amdgcn_bitop3_b32(x, x >> 7, 0x00AA00AAu, 0x28);  // (x ^ (x >> 7)) & m
  x = __builtin_amdgcn_bitop3_b32(x, t, t << 7, 0x96);
  t = __builtin_amdgcn_bitop3_b32(x, x >> 14, 0x0000CCCCu, 0x28);
  return __builtin_amdgcn_bitop3_b32(x, t, t << 14, 0x96);
}

// The grid is sized for six (one weight vector, at most 80 VGPRs) or four
// (two, at most 128 VGPRs) resident wavefronts per SIMD.
template <int NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NC == 1 ? 6 : 4, NC == 1 ? 6 : 4))) void bitfold_kernel(
    const uint32_t *el, int il_log, const f128 *hi0, const f128 *hi1, uint64_t n_ih, uint64_t ih_per_wg,
    f128 *__restrict__ partial) {
  // entry (k, half, n, copy): half 0 combines weights ih0..ih0+3, half 1
  // ih0+4..ih0+7, by the bits of n; bank group = copy
  __shared__ f128 tab[2][NC][2][16][16];
  // The group's input words, dword (il * 4 + d) * 8 + row: element il's
  // dword d in the group's 8 rows.
  __shared__ uint4 stage[2][128];
  const uint32_t t = threadIdx.x, il = t >> 4, b = t & 15, lane = t & 63, wave = t >> 6;
  const uint64_t ih_begin = (uint64_t)blockIdx.x * ih_per_wg;
  uint64_t ih_end = ih_begin + ih_per_wg;
  if (ih_end > n_ih) ih_end = n_ih;
  // Table-build role: wave -> (k, half); lane -> value n = lane mod 16 and
  // copies n + lane / 16 + {0, 8} (NC = 1, the two wave pairs of a half adding
  // 0 / 4) or n + lane / 16 + {0, 4, 8, 12} (NC = 2), mod 16.  The 16 lanes
  // of a write thus hit 16 distinct copies, i.e. bank groups.
  const uint32_t bk = NC == 2 ? (wave >> 1) : 0u;
  const uint32_t half = wave & 1u;
  const uint32_t bn = lane & 15u;
  const uint32_t bc = (lane >> 4) + (NC == 2 ? 0u : 4u * (wave >> 1));
  // The 4 weights of a half are 16 dwords: lane i loads dword i mod 16 and the
  // build reads them back with v_readlane.  Vector loads, unlike scalar ones,
  // stay in flight across the LDS barrier.
  const uint32_t *hw = (const uint32_t *)(bk ? hi1 : hi0) + 16u * half + (lane & 15u);
  // Input role: the workgroup's 16 elements of a row are 64 dwords; wavefront
  // m loads rows 2m and 2m+1, lane i dword i, so each load is 256 B
  // contiguous and no dword is fetched twice.
  const uint32_t ld_off = ((uint32_t)blockIdx.y << 8) + (lane << 2);
  const uint32_t row_shift = (uint32_t)il_log + 4u;
  uint2 *const st_w = (uint2 *)&stage[0][0] + lane * 4 + wave;  // uint2 (rows 2m, 2m+1) of dword lane
  const uint4 *const st_r = &stage[0][0] + (il * 4 + (b >> 2)) * 2;
  // v_perm selector taking byte b mod 4 of two words into bytes 0, 1
  const uint32_t gsel = (b & 3u) | ((4u + (b & 3u)) << 8) | 0x0C0C0000u;
  const uint32_t copy16 = b << 4;  // this lane's copy, as a byte offset

  f128 acc[NC][8];
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[k][c] = f128_zero();

  auto load = [&](uint64_t ih, uint2 &v, uint32_t &hv) {
    const uint64_t r = ih + 2 * wave;
    v.x = *(const uint32_t *)((const char *)el + (r << row_shift) + ld_off);
    v.y = *(const uint32_t *)((const char *)el + ((r + 1) << row_shift) + ld_off);
    hv = hw[4 * ih];
  };
  auto group = [&](const uint2 v, uint32_t hv, int buf) {
    // (1) the tables: value n = sum_j bit_j(n) h_{4 half + j}
    {
      f128 s = f128_zero();
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t m = 0u - ((bn >> j) & 1u);
        // (h & m) ^ s in one three-input op per dword
        s.x = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 0), m, s.x, 0x6a);
        s.y = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 1), m, s.y, 0x6a);
        s.z = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 2), m, s.z, 0x6a);
        s.w = __builtin_amdgcn_bitop3_b32(__builtin_amdgcn_readlane(hv, 4 * j + 3), m, s.w, 0x6a);
      }
#pragma unroll
      for (int q = 0; q < (NC == 2 ? 4 : 2); ++q) tab[buf][bk][half][bn][(bn + bc + (NC == 2 ? 4 : 8) * q) & 15u] = s;
    }
    st_w[buf * 256] = v;
    __syncthreads();
    // (2) byte b of the 8 elements: lo = elements 0..3, hi = elements 4..7
    const uint4 r03 = st_r[buf * 128], r47 = st_r[buf * 128 + 1];
    const uint32_t w[8] = {r03.x, r03.y, r03.z, r03.w, r47.x, r47.y, r47.z, r47.w};
    const uint32_t p01 = __builtin_amdgcn_perm(w[1], w[0], gsel), p23 = __builtin_amdgcn_perm(w[3], w[2], gsel);
    const uint32_t p45 = __builtin_amdgcn_perm(w[5], w[4], gsel), p67 = __builtin_amdgcn_perm(w[7], w[6], gsel);
    const uint32_t lo = transpose4x4_pair(__builtin_amdgcn_perm(p23, p01, 0x05040100u));
    const uint32_t hi = transpose4x4_pair(__builtin_amdgcn_perm(p67, p45, 0x05040100u));
    // byte r of N[0] / N[1]: index of table half 0 for column r / 4 + r;
    // N[2] / N[3]: the same for half 1
    const uint32_t N[4] = {lo & 0x0F0F0F0Fu, (lo >> 4) & 0x0F0F0F0Fu, hi & 0x0F0F0F0Fu, (hi >> 4) & 0x0F0F0F0Fu};
    const char *tb = (const char *)&tab[buf][0][0][0][0];
    // reads issued in batches of B columns before their XORs
    constexpr int B = 2;
#pragma unroll
    for (int c0 = 0; c0 < 8; c0 += B) {
      f128 e[NC][B][2];
#pragma unroll
      for (int c = c0; c < c0 + B; ++c) {
        // address n * 256 + copy * 16: byte 1 = n, byte 0 = copy16
        const uint32_t sel = 0x0C0C0000u | ((4u + (uint32_t)(c & 3)) << 8);
        const uint32_t a0 = __builtin_amdgcn_perm(N[c >> 2], copy16, sel);
        const uint32_t a1 = __builtin_amdgcn_perm(N[2 + (c >> 2)], copy16, sel);
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          e[k][c - c0][0] = *(const f128 *)(tb + k * 8192 + a0);
          e[k][c - c0][1] = *(const f128 *)(tb + k * 8192 + 4096 + a1);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int c = c0; c < c0 + B; ++c) acc[k][c] = f128_xor3(acc[k][c], e[k][c - c0][0], e[k][c - c0][1]);
    }
  };

  // Two register sets alternate: the next group's loads are issued before
  // this group's work, so their latency overlaps it.  The last group
  // reloads itself (in bounds, unused).  The empty asm statements keep the
  // compiler from moving the loads into the other half of the loop (the input
  // pointers are deliberately not __restrict__, so the memory clobber applies
  // to them).
  uint2 wa, wb;
  uint32_t ha, hb;
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
  // About 6 (one weight vector) or 4 (two) wavefronts per SIMD of the 256
  // CUs, each workgroup a multiple of 8 weight rows.
  uint64_t nx = (nc == 1 ? 1536 : 1024) / ny;
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
