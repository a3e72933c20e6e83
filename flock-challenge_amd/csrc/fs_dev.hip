// The Fiat-Shamir transcript on the device: the BLAKE3 transcript of
// transcript.h (FsChallenger::with_hash(.., Blake3), crates/flock-core/src/
// challenger.rs:246-437) as an incremental BLAKE3 hasher held in device
// memory, with the same state layout as the host's hb3::Hasher, so a proof
// moves its transcript between host and device with one copy.
//
// Used by the zerocheck's multilinear rounds: each round's kernels write
// its message (G(1), G(inf)) into device memory, fs_round_kernel absorbs it
// and squeezes the round's challenge into device memory, and the next
// round's kernels read the challenge from there, so the rounds run back to
// back on the stream with no host round trip, and the whole sequence can
// be captured in a HIP graph.
//
// The hasher runs on one lane (the transcript is a serial chain of a few
// compressions per round) with its state staged in LDS.
#include <hip/hip_runtime.h>
#include <string.h>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "ctx.h"
#include "fs_dev.h"
#include "gf128_dev.h"

namespace flock {

namespace {

// The host hb3::Hasher's fields, in its order.
struct FsState {
  uint32_t cv[8];
  uint64_t chunk;
  uint8_t block[64];
  uint32_t block_len, blocks_done;
  uint32_t stack[54][8];
  uint32_t depth;
};
static_assert(sizeof(FsState) == FS_STATE_BYTES, "transcript state layout");

__device__ void fs_words(const uint8_t *b, uint32_t w[16]) {
#pragma unroll
  for (int i = 0; i < 16; ++i)
    w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
           ((uint32_t)b[4 * i + 3] << 24);
}

// Absorb n bytes (Hasher::update): a full block is compressed only once
// more input arrives, so the last block is always there to finalize; the
// 16th block of a chunk closes it and merges completed subtrees.
__device__ void fs_update(FsState &s, const uint8_t *p, uint32_t n) {
  while (n) {
    if (s.block_len == 64) {
      uint32_t w[16], o[16];
      fs_words(s.block, w);
      if (s.blocks_done == 15) {
        b3::compress_full(s.cv, w, s.chunk, 64, (s.blocks_done == 0 ? b3::CHUNK_START : 0u) | b3::CHUNK_END, o);
        uint32_t c[8];
        for (int i = 0; i < 8; ++i) c[i] = o[i];
        uint64_t total = s.chunk + 1;
        while ((total & 1) == 0) {
          uint32_t pm[16], po[16];
          for (int i = 0; i < 8; ++i) {
            pm[i] = s.stack[s.depth - 1][i];
            pm[8 + i] = c[i];
          }
          b3::compress_full(b3::IV, pm, 0, 64, b3::PARENT, po);
          for (int i = 0; i < 8; ++i) c[i] = po[i];
          --s.depth;
          total >>= 1;
        }
        for (int i = 0; i < 8; ++i) s.stack[s.depth][i] = c[i];
        ++s.depth;
        ++s.chunk;
        for (int i = 0; i < 8; ++i) s.cv[i] = b3::IV[i];
        s.blocks_done = 0;
      } else {
        b3::compress_full(s.cv, w, s.chunk, 64, s.blocks_done == 0 ? b3::CHUNK_START : 0u, o);
        for (int i = 0; i < 8; ++i) s.cv[i] = o[i];
        ++s.blocks_done;
      }
      for (int i = 0; i < 64; ++i) s.block[i] = 0;
      s.block_len = 0;
    }
    uint32_t take = 64 - s.block_len;
    if (take > n) take = n;
    for (uint32_t i = 0; i < take; ++i) s.block[s.block_len + i] = p[i];
    s.block_len += take;
    p += take;
    n -= take;
  }
}

// The first 16 bytes of the root output of the current state
// (Hasher::finalize_xof): the open chunk's output merged with the stack,
// then the root compression at output block 0.
__device__ void fs_squeeze16(const FsState &s, uint8_t out[16]) {
  uint32_t cv[8], m[16], o[16];
  for (int i = 0; i < 8; ++i) cv[i] = s.cv[i];
  fs_words(s.block, m);
  uint64_t counter = s.chunk;
  uint32_t len = s.block_len, flags = (s.blocks_done == 0 ? b3::CHUNK_START : 0u) | b3::CHUNK_END;
  for (int i = (int)s.depth - 1; i >= 0; --i) {
    b3::compress_full(cv, m, counter, len, flags, o);
    for (int j = 0; j < 8; ++j) {
      m[j] = s.stack[i][j];
      m[8 + j] = o[j];
      cv[j] = b3::IV[j];
    }
    counter = 0;
    len = 64;
    flags = b3::PARENT;
  }
  b3::compress_full(cv, m, 0, len, flags | b3::ROOT, o);
  for (int i = 0; i < 4; ++i)
    for (int b = 0; b < 4; ++b) out[4 * i + b] = (uint8_t)(o[i] >> (8 * b));
}

__device__ void fs_observe_f128(FsState &s, const f128 v) {
  uint8_t b[18] = {3, 1};
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
  for (int i = 0; i < 4; ++i)
    for (int k = 0; k < 4; ++k) b[2 + 4 * i + k] = (uint8_t)(w[i] >> (8 * k));
  fs_update(s, b, 18);
}

__device__ f128 fs_sample(FsState &s) {
  const uint8_t tag[2] = {4, 1};
  fs_update(s, tag, 2);
  uint8_t v[16];
  fs_squeeze16(s, v);
  fs_update(s, v, 16);
  f128 r;
  r.x = (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
  r.y = (uint32_t)v[4] | ((uint32_t)v[5] << 8) | ((uint32_t)v[6] << 16) | ((uint32_t)v[7] << 24);
  r.z = (uint32_t)v[8] | ((uint32_t)v[9] << 8) | ((uint32_t)v[10] << 16) | ((uint32_t)v[11] << 24);
  r.w = (uint32_t)v[12] | ((uint32_t)v[13] << 8) | ((uint32_t)v[14] << 16) | ((uint32_t)v[15] << 24);
  return r;
}

// One round's transcript steps: observe the message's n_msg values, then
// sample the challenge into *rho.  The state is staged through LDS.
__global__ __launch_bounds__(64) void fs_round_kernel(FsState *__restrict__ g, const f128 *__restrict__ msg, int n_msg,
                                                     f128 *__restrict__ rho) {
  __shared__ FsState s;
  const uint32_t *src = (const uint32_t *)g;
  uint32_t *dst = (uint32_t *)&s;
  for (uint32_t i = threadIdx.x; i < sizeof(FsState) / 4; i += 64) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < n_msg; ++i) fs_observe_f128(s, msg[i]);
    *rho = fs_sample(s);
  }
  __syncthreads();
  uint32_t *back = (uint32_t *)g;
  for (uint32_t i = threadIdx.x; i < sizeof(FsState) / 4; i += 64) back[i] = dst[i];
}

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

int fs_round_device(Ctx *ctx, hipStream_t st, void *d_state, const void *d_msg, int n_msg, void *d_rho) {
  hipLaunchKernelGGL(fs_round_kernel, dim3(1), dim3(64), 0, st, (FsState *)d_state, (const f128 *)d_msg, n_msg,
                     (f128 *)d_rho);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int rho_bytetab_device(Ctx *ctx, hipStream_t st, const void *d_ws, const void *d_rho, void *d_out) {
  hipLaunchKernelGGL(rho_bytetab_kernel, dim3(1), dim3(256), 0, st, (const f128 *)d_ws, (const f128 *)d_rho,
                     (f128 *)d_out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock
