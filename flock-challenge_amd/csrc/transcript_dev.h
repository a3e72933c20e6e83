// The Fiat-Shamir transcript's BLAKE3 hasher on the device: the BLAKE3
// transcript of transcript.h (FsChallenger::with_hash(.., Blake3),
// crates/flock-core/src/challenger.rs:246-437) as an incremental hasher whose
// state has the layout of the host's hb3::Hasher, so a proof moves its
// transcript between host and device with one copy.
//
// The hasher is a serial chain of a few compressions per step and runs on
// one lane.  While a step runs, the scalar fields (chaining value, counters,
// stack depth) live in registers and the 64-byte block and the subtree stack
// in LDS; inputs arrive as little-endian words in registers, so no byte of
// them goes through per-lane scratch memory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "blake3_dev.h"
#include "fs_dev.h"
#include "gf128_dev.h"

namespace flock {

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

// The scalar fields of an FsState while a step runs.
struct FsRegs {
  uint32_t cv[8];
  uint64_t chunk;
  uint32_t len, done, depth;
};

__device__ __forceinline__ void fs_regs_load(const FsState &s, FsRegs &r) {
#pragma unroll
  for (int i = 0; i < 8; ++i) r.cv[i] = s.cv[i];
  r.chunk = s.chunk;
  r.len = s.block_len;
  r.done = s.blocks_done;
  r.depth = s.depth;
}

__device__ __forceinline__ void fs_regs_store(FsState &s, const FsRegs &r) {
#pragma unroll
  for (int i = 0; i < 8; ++i) s.cv[i] = r.cv[i];
  s.chunk = r.chunk;
  s.block_len = r.len;
  s.blocks_done = r.done;
  s.depth = r.depth;
}

// The block as 16 little-endian words (it sits 8-byte aligned in FsState).
__device__ __forceinline__ void fs_block_words(const FsState &s, uint32_t w[16]) {
  const uint2 *b = (const uint2 *)s.block;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const uint2 v = b[i];
    w[2 * i] = v.x;
    w[2 * i + 1] = v.y;
  }
}

// The block is full and more input follows (Hasher::update compresses a
// full block only then, so the last block is always there to finalize):
// compress it; the 16th block of a chunk closes the chunk and merges the
// completed subtrees on the stack.  (One compression call site for both
// cases keeps the inlined code small.)
__device__ __forceinline__ void fs_flush(FsState &s, FsRegs &r) {
  uint32_t w[16], o[16];
  fs_block_words(s, w);
  const bool last = r.done == 15;
  b3::compress_full(r.cv, w, r.chunk, 64, last ? b3::CHUNK_END : r.done == 0 ? b3::CHUNK_START : 0u, o);
  if (last) {
    uint32_t c[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) c[i] = o[i];
    uint64_t total = r.chunk + 1;
    while ((total & 1) == 0) {
      uint32_t pm[16], po[16];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        pm[i] = s.stack[r.depth - 1][i];
        pm[8 + i] = c[i];
      }
      b3::compress_full(b3::IV, pm, 0, 64, b3::PARENT, po);
#pragma unroll
      for (int i = 0; i < 8; ++i) c[i] = po[i];
      --r.depth;
      total >>= 1;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) s.stack[r.depth][i] = c[i];
    ++r.depth;
    ++r.chunk;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.cv[i] = b3::IV[i];
    r.done = 0;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) r.cv[i] = o[i];
    ++r.done;
  }
  uint2 *b = (uint2 *)s.block;
#pragma unroll
  for (int i = 0; i < 8; ++i) b[i] = make_uint2(0u, 0u);
  r.len = 0;
}

// Absorb the first n bytes of the little-endian words w (Hasher::update).
template <int NW>
__device__ __forceinline__ void fs_absorb(FsState &s, FsRegs &r, const uint32_t (&w)[NW], uint32_t n) {
#pragma unroll 1
  for (uint32_t i = 0; i < n; ++i) {
    if (r.len == 64) fs_flush(s, r);
    uint32_t word = w[0];
#pragma unroll
    for (int k = 1; k < NW; ++k)
      if ((i >> 2) == (uint32_t)k) word = w[k];
    s.block[r.len++] = (uint8_t)(word >> (8 * (i & 3)));
  }
}

// The first 16 bytes of the root output of the current state
// (Hasher::finalize_xof): the open chunk's output merged with the stack,
// then the root compression at output block 0.
__device__ __forceinline__ void fs_squeeze16(const FsState &s, const FsRegs &r, uint32_t out[4]) {
  uint32_t cv[8], m[16], o[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) cv[i] = r.cv[i];
  fs_block_words(s, m);
  uint64_t counter = r.chunk;
  uint32_t len = r.len, flags = (r.done == 0 ? b3::CHUNK_START : 0u) | b3::CHUNK_END;
  for (int i = (int)r.depth - 1; i >= 0; --i) {
    b3::compress_full(cv, m, counter, len, flags, o);
#pragma unroll
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
#pragma unroll
  for (int i = 0; i < 4; ++i) out[i] = o[i];
}

// One zerocheck round's transcript steps, Transcript::observe of the two
// F128 values m0, m1 (tag bytes 3, 1, then the value's 16 bytes, each) and
// Transcript::sample of one F128 (tag bytes 4, 1, squeeze 16 bytes, absorb
// them; they are the sample).  The 38 bytes before the squeeze and the 16
// after go through one absorb site, so the hasher is inlined once.
__device__ __forceinline__ f128 fs_round2(FsState &s, FsRegs &r, const f128 m0, const f128 m1) {
  uint32_t w[10] = {3u | (1u << 8) | (m0.x << 16), (m0.x >> 16) | (m0.y << 16), (m0.y >> 16) | (m0.z << 16),
                    (m0.z >> 16) | (m0.w << 16),   (m0.w >> 16) | (3u << 16) | (1u << 24), m1.x, m1.y, m1.z, m1.w,
                    4u | (1u << 8)};
  uint32_t n = 38;
  f128 out = f128_zero();
#pragma unroll 1
  for (int phase = 0; phase < 2; ++phase) {
    fs_absorb(s, r, w, n);
    if (phase == 0) {
      uint32_t v[4];
      fs_squeeze16(s, r, v);
      out = make_uint4(v[0], v[1], v[2], v[3]);
#pragma unroll
      for (int i = 0; i < 4; ++i) w[i] = v[i];
      n = 16;
    }
  }
  return out;
}

// One zerocheck round's transcript steps, run by every thread of a
// workgroup: the state at g is staged through the LDS copy s, thread 0
// observes msg[0], msg[1] and samples the challenge into *rho, and the state
// goes back to g.  msg may be memory thread 0 itself has just written.
__device__ __forceinline__ void fs_round_block(FsState *g, FsState &s, const f128 *msg, f128 *rho) {
  const uint32_t *src = (const uint32_t *)g;
  uint32_t *dst = (uint32_t *)&s;
  for (uint32_t i = threadIdx.x; i < sizeof(FsState) / 4; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    FsRegs r;
    fs_regs_load(s, r);
    *rho = fs_round2(s, r, msg[0], msg[1]);
    fs_regs_store(s, r);
  }
  __syncthreads();
  uint32_t *back = (uint32_t *)g;
  for (uint32_t i = threadIdx.x; i < sizeof(FsState) / 4; i += blockDim.x) back[i] = dst[i];
}

}  // namespace flock
