// Device-transcript helpers of the zerocheck and the Ligerito fold rounds.
// The transcript itself (transcript_dev.h) runs in each round's reduction
// kernel: it absorbs the round's message and, in the zerocheck, squeezes the
// round's challenge into device memory; a Ligerito round's challenge comes
// after a proof of work, which the kernels below run on the device.  The
// next round's kernels read the challenge from device memory, so the rounds
// run back to back on the stream with no host round trip.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "ctx.h"
#include "fs_dev.h"
#include "gf128_dev.h"
#include "gf128_short.h"
#include "transcript_dev.h"

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

// ---------------------------------------------------------------- Ligerito challenges
// The proof-of-work base: the transcript's 32-byte state digest into dg.
__global__ __launch_bounds__(64) void fs_digest_kernel(FsState *g, uint32_t *dg) {
  __shared__ FsState s;
  fs_block(g, s, [&](FsState &st, FsRegs &r, uint32_t *pm, uint32_t lane) {
    uint32_t d[8];
    fs_digest32(st, r, pm, lane, d);
    if (lane < 8) {
      uint32_t v = d[0];
#pragma unroll
      for (int i = 1; i < 8; ++i)
        if (lane == (uint32_t)i) v = d[i];
      dg[lane] = v;
    }
  });
}

// The smallest nonce n >= start with BLAKE3(dg | n_le | 0^24) starting with
// `bits` zero bits (challenger.rs:1246-1300), into *best (atomicMin; it holds
// ~0 on entry).  Windows of `window` nonces are scanned in order, each split
// evenly over the workgroups; a workgroup stops after a window once *best
// lies in it or an earlier one, so every nonce below the final *best has
// been tried by the workgroup that owns it, and every workgroup stops after
// max_windows windows at the latest.
constexpr uint32_t POW_THREADS = 256, POW_WINDOWS = 64;
// The digest comes from dg (device memory) or, with dg null, as the kernel
// arguments d0, d1.
__global__ __launch_bounds__(POW_THREADS) void pow_chain_kernel(const uint32_t *__restrict__ dg, uint4 d0, uint4 d1,
                                                                uint32_t bits, uint64_t start, uint64_t window,
                                                                uint32_t max_windows, unsigned long long *best) {
  __shared__ unsigned long long seen;
  uint32_t d[8] = {d0.x, d0.y, d0.z, d0.w, d1.x, d1.y, d1.z, d1.w};
  if (dg) {
#pragma unroll
    for (int i = 0; i < 8; ++i) d[i] = dg[i];
  }
  const uint64_t per_wg = window / gridDim.x;
  for (uint32_t w = 0; w < max_windows; ++w) {
    const uint64_t base = start + (uint64_t)w * window + (uint64_t)blockIdx.x * per_wg;
    for (uint64_t i = threadIdx.x; i < per_wg; i += POW_THREADS) {
      const uint64_t nonce = base + i;
      if (b3::pow_ok(d, nonce, bits)) atomicMin(best, (unsigned long long)nonce);
    }
    __syncthreads();
    if (threadIdx.x == 0) seen = __hip_atomic_load(best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (seen < start + (uint64_t)(w + 1) * window) break;  // workgroup-uniform
    __syncthreads();  // seen is read before thread 0 rewrites it
  }
}

// Up to FLOCK_POW_WAVE_BITS bits the wavefront that runs the transcript
// searches the nonces itself, 64 at a time (about 2^bits / 64 steps), with
// no other launch.
constexpr uint32_t POW_WAVE_BITS = FLOCK_POW_WAVE_BITS;

// Observe the nonce (when bits > 0; none found: status 1, nonce 0) and
// re-arm *best, then sample the round's challenge into *r.  With bits <=
// POW_WAVE_BITS the nonce is searched here (from the state's digest), else
// it is *best, left by pow_chain_kernel.
__global__ __launch_bounds__(64) void lig_nonce_sample_kernel(FsState *g, unsigned long long *best, uint32_t bits,
                                                              uint64_t *nonce_out, f128 *r, uint32_t *status) {
  __shared__ FsState s;
  fs_block(g, s, [&](FsState &st, FsRegs &rg, uint32_t *pm, uint32_t lane) {
    if (bits && bits <= POW_WAVE_BITS) {
      uint32_t d[8];
      fs_digest32(st, rg, pm, lane, d);
      uint64_t nonce = ~0ull;
      for (uint32_t it = 0; it < (1u << 12); ++it) {  // 2^18 nonces at most
        const uint64_t n = (uint64_t)it * 64 + lane;
        const uint64_t hit = __builtin_amdgcn_ballot_w64(b3::pow_ok(d, n, bits));
        if (hit) {  // wavefront-uniform
          nonce = (uint64_t)it * 64 + (uint64_t)__builtin_ctzll(hit);
          break;
        }
      }
      if (nonce == ~0ull) {
        if (lane == 0) *status = 1u;
        nonce = 0;
      }
      fs_observe_u64(st, rg, nonce, pm, lane);
      if (lane == 0) *nonce_out = nonce;
    } else if (bits) {
      uint64_t nonce = *best;
      if (nonce == ~0ull) {
        if (lane == 0) *status = 1u;
        nonce = 0;
      }
      fs_observe_u64(st, rg, nonce, pm, lane);
      if (lane == 0) {
        *nonce_out = nonce;
        *best = ~0ull;
      }
    }
    const f128 v = fs_sample(st, rg, pm, lane);
    if (lane == 0) *r = v;
  });
}

}  // namespace

int lig_challenge_device(Ctx *ctx, hipStream_t st, void *d_fs, uint32_t bits, void *d_work, bool digest_ready,
                         void *d_nonce, void *d_r, void *d_status) {
  if (bits > 40) {
    ctx->last_error = "proof of work: more than 40 bits";
    return FLOCK_E_ARG;
  }
  unsigned long long *best = (unsigned long long *)d_work;
  uint32_t *dg = (uint32_t *)lig_pow_digest(d_work);
  if (bits > POW_WAVE_BITS) {
    if (!digest_ready) hipLaunchKernelGGL(fs_digest_kernel, dim3(1), dim3(64), 0, st, (FsState *)d_fs, dg);
    // one search from nonce 0: with bits <= 30 it finds nothing with
    // probability below e^-64 (the status word then says so)
    if (int rc = pow_chain_device(ctx, st, d_work, nullptr, bits, 0, nullptr)) return rc;
  }
  hipLaunchKernelGGL(lig_nonce_sample_kernel, dim3(1), dim3(64), 0, st, (FsState *)d_fs, best, bits,
                     (uint64_t *)d_nonce, (f128 *)d_r, (uint32_t *)d_status);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int pow_chain_device(Ctx *ctx, hipStream_t st, void *d_work, const uint8_t *digest, uint32_t bits, uint64_t start,
                     uint64_t *span) {
  if (bits == 0 || bits > 40) {
    ctx->last_error = "proof of work: 1 to 40 bits";
    return FLOCK_E_ARG;
  }
  // windows of one expected solution (at least 2^10 nonces, at most 2^30):
  // about 1.6 windows are scanned on average
  const uint64_t window = 1ull << std::min<uint32_t>(std::max<uint32_t>(bits, 10), 30);
  const uint64_t wgs = std::min<uint64_t>(window / POW_THREADS, 1024);
  uint4 d0 = make_uint4(0u, 0u, 0u, 0u), d1 = d0;
  if (digest) {
    memcpy(&d0, digest, 16);
    memcpy(&d1, digest + 16, 16);
  }
  hipLaunchKernelGGL(pow_chain_kernel, dim3((unsigned)wgs), dim3(POW_THREADS), 0, st,
                     digest ? nullptr : (const uint32_t *)lig_pow_digest(d_work), d0, d1, bits, start, window,
                     POW_WINDOWS, (unsigned long long *)d_work);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  if (span) *span = window * POW_WINDOWS;
  return FLOCK_OK;
}

int rho_bytetab_device(Ctx *ctx, hipStream_t st, const void *d_ws, const void *d_rho, void *d_out) {
  hipLaunchKernelGGL(rho_bytetab_kernel, dim3(1), dim3(256), 0, st, (const f128 *)d_ws, (const f128 *)d_rho,
                     (f128 *)d_out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock
