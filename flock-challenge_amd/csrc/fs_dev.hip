// Device-transcript helpers of the zerocheck.  The transcript itself
// (transcript_dev.h) runs in each round's reduction kernel: it absorbs the
// round's message (G(1), G(inf)) and squeezes the round's challenge into
// device memory, and the next round's kernels read the challenge from
// there, so the rounds run back to back on the stream with no host round
// trip, and the whole sequence can be captured in a HIP graph.
#include <hip/hip_runtime.h>
#include <string.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"
#include "fs_dev.h"
#include "gf128_dev.h"

namespace flock {

namespace {

// a * b by shifts (one product per lane; for a few dozen products).
__device__ f128 f128_mul_shift(f128 a, f128 b) {
  f128 acc = f128_zero();
  const uint32_t bw[4] = {b.x, b.y, b.z, b.w};
  for (int i = 127; i >= 0; --i) {
    acc = f128_mulx(acc);
    if ((bw[i >> 5] >> (i & 31)) & 1) f128_xor_into(acc, a);
  }
  return acc;
}

// Byte tables of rho L_s(z): out[v * 8 + p] = sum over the set bits t of v
// of rho ws[8 p + t] (the fold after zerocheck round 0 evaluates
// (rho L)(row) through them).
__global__ __launch_bounds__(256) void rho_bytetab_kernel(const f128 *__restrict__ ws, const f128 *__restrict__ rho,
                                                          f128 *__restrict__ out) {
  __shared__ f128 wr[64];
  const uint32_t t = threadIdx.x;
  if (t < 64) wr[t] = f128_mul_shift(*rho, ws[t]);
  __syncthreads();
  for (uint32_t e = t; e < 2048; e += 256) {
    const uint32_t v = e >> 3, p = e & 7;
    f128 acc = f128_zero();
    for (int b = 0; b < 8; ++b)
      if ((v >> b) & 1) f128_xor_into(acc, wr[8 * p + b]);
    out[e] = acc;
  }
}

}  // namespace

int rho_bytetab_device(Ctx *ctx, hipStream_t st, const void *d_ws, const void *d_rho, void *d_out) {
  hipLaunchKernelGGL(rho_bytetab_kernel, dim3(1), dim3(256), 0, st, (const f128 *)d_ws, (const f128 *)d_rho,
                     (f128 *)d_out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock
