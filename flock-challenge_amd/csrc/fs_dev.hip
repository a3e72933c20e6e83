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
#include "gf128_short.h"

namespace flock {

namespace {

// Byte tables of rho L_s(z): out[v * 8 + p] = sum over the set bits t of v
// of rho ws[8 p + t] (the fold after zerocheck round 0 evaluates
// (rho L)(row) through them).
__global__ __launch_bounds__(256) void rho_bytetab_kernel(const f128 *__restrict__ ws, const f128 *__restrict__ rho,
                                                          f128 *__restrict__ out) {
  __shared__ f128 wr[64];
  const uint32_t t = threadIdx.x;
  if (t < 64) {
    // rho is one value for every lane: its product with ws[t] by shifted
    // XORs over rho's set bits (a scalar loop)
    const f128 r = *rho, w = ws[t];
    const uint32_t wv[4] = {w.x, w.y, w.z, w.w};
    uint32_t o[4];
    shortmul::mul(wv, __builtin_amdgcn_readfirstlane(r.x), __builtin_amdgcn_readfirstlane(r.y),
                  __builtin_amdgcn_readfirstlane(r.z), __builtin_amdgcn_readfirstlane(r.w), o);
    wr[t] = make_uint4(o[0], o[1], o[2], o[3]);
  }
  __syncthreads();
  for (uint32_t e = t; e < 2048; e += 256) {
    const uint32_t v = e >> 3, p = e & 7;
    f128 acc = f128_zero();
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint32_t mk = 0u - ((v >> b) & 1u);
      const f128 x = wr[8 * p + b];
      acc = f128_xor(acc, make_uint4(x.x & mk, x.y & mk, x.z & mk, x.w & mk));
    }
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
