// The Fiat-Shamir transcript's BLAKE3 hasher on the device: the BLAKE3
// transcript of transcript.h (FsChallenger::with_hash(.., Blake3),
// crates/flock-core/src/challenger.rs:246-437) as an incremental hasher whose
// state has the layout of the host's hb3::Hasher, so a proof moves its
// transcript between host and device with one copy.
//
// The hasher is a serial chain of a few compressions per step and runs on
// one wavefront, each compression spread over a quad of lanes (one state
// column per lane).  While a step runs, the scalar fields (chaining value,
// counters, stack depth) live in scalar registers and the 64-byte block and
// the subtree stack in LDS; inputs arrive as little-endian words in
// registers.
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

// ---------------------------------------------------------------- quad compression
// One BLAKE3 compression (public BLAKE3 specification) on a quad of lanes:
// lane i of each quad holds column i of the 4 x 4 state (a = s[i],
// b = s[4 + i], c = s[8 + i], d = s[12 + i]), so the four G functions of a
// column step run side by side; for the diagonal step, lane i takes b, c, d
// from lanes i + 1, i + 2, i + 3 of its quad (DPP quad permutations) and
// gives them back afterwards.  A compression is then a chain of about a
// quarter of the one-lane instructions.  The message words (16, in LDS) are
// read by each lane at its own schedule positions.  Every quad of the
// wavefront computes the same values.

// quad_perm controls: lane i reads lane (i + k) mod 4 of its quad
constexpr int QUAD_NEXT1 = 0x39, QUAD_NEXT2 = 0x4E, QUAD_NEXT3 = 0x93;

__device__ __forceinline__ uint32_t quad_take(uint32_t x, const int ctrl) {
  switch (ctrl) {
    case QUAD_NEXT1: return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, QUAD_NEXT1, 0xF, 0xF, false);
    case QUAD_NEXT2: return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, QUAD_NEXT2, 0xF, 0xF, false);
    default: return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, QUAD_NEXT3, 0xF, 0xF, false);
  }
}

// The word of schedule position j in round r: m[PERM^r(j)], with the
// permutation PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
// the specification applies between rounds.
__host__ __device__ constexpr uint32_t b3_sched(int r, int j) {
  constexpr uint8_t P[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  int k = j;
  for (int i = 0; i < r; ++i) k = P[k];
  return (uint32_t)k;
}
// Byte i of quad_msg_idx(r, slot) is the word lane i reads in round r:
// slots 0, 1 the column step's two words (positions 2i, 2i + 1), slots 2, 3
// the diagonal step's (8 + 2i, 9 + 2i).
__host__ __device__ constexpr uint32_t quad_msg_idx(int r, int slot) {
  uint32_t v = 0;
  for (int i = 0; i < 4; ++i) {
    const int j = slot < 2 ? 2 * i + slot : 8 + 2 * i + (slot - 2);
    v |= b3_sched(r, j) << (8 * i);
  }
  return v;
}

// The compression of the 16 message words at m (LDS) with the uniform
// chaining value cv: lane i returns output words i, 4 + i, 8 + i, 12 + i
// (the last two only matter for XOF blocks).
__device__ __forceinline__ void compress_quad(const uint32_t (&cv)[8], const uint32_t *m, uint64_t counter,
                                              uint32_t len, uint32_t flags, uint32_t qi, uint32_t out[4]) {
  // lane qi's column, selected with lane masks (an index into cv would put
  // cv in scratch memory)
  const uint32_t k0 = 0u - (uint32_t)(qi == 0), k1 = 0u - (uint32_t)(qi == 1), k2 = 0u - (uint32_t)(qi == 2),
                 k3 = 0u - (uint32_t)(qi == 3);
  const uint32_t c0 = (cv[0] & k0) | (cv[1] & k1) | (cv[2] & k2) | (cv[3] & k3);
  const uint32_t c4 = (cv[4] & k0) | (cv[5] & k1) | (cv[6] & k2) | (cv[7] & k3);
  uint32_t a = c0, b = c4;
  uint32_t c = (0x6A09E667u & k0) | (0xBB67AE85u & k1) | (0x3C6EF372u & k2) | (0xA54FF53Au & k3);
  uint32_t d = ((uint32_t)counter & k0) | ((uint32_t)(counter >> 32) & k1) | (len & k2) | (flags & k3);
  const uint32_t sh = 8 * qi;
  using b3::rotr;
#pragma unroll
  for (int r = 0; r < 7; ++r) {
    const uint32_t m0 = m[(quad_msg_idx(r, 0) >> sh) & 15u], m1 = m[(quad_msg_idx(r, 1) >> sh) & 15u];
    const uint32_t m2 = m[(quad_msg_idx(r, 2) >> sh) & 15u], m3 = m[(quad_msg_idx(r, 3) >> sh) & 15u];
    B3_G(a, b, c, d, m0, m1);
    b = quad_take(b, QUAD_NEXT1);
    c = quad_take(c, QUAD_NEXT2);
    d = quad_take(d, QUAD_NEXT3);
    B3_G(a, b, c, d, m2, m3);
    b = quad_take(b, QUAD_NEXT3);
    c = quad_take(c, QUAD_NEXT2);
    d = quad_take(d, QUAD_NEXT1);
  }
  out[0] = a ^ c;
  out[1] = b ^ d;
  out[2] = c ^ c0;
  out[3] = d ^ c4;
}

// The uniform 8-word chaining value from the quad's output (words 0..7 sit
// in lanes 0..3 of out[0] and out[1]).
__device__ __forceinline__ void quad_cv(const uint32_t out[4], uint32_t (&cv)[8]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    cv[k] = __builtin_amdgcn_readlane(out[0], k);
    cv[4 + k] = __builtin_amdgcn_readlane(out[1], k);
  }
}

// ---------------------------------------------------------------- the hasher
// Lane 0's LDS writes before this point are seen by the wavefront's LDS reads
// after it: a wavefront's LDS accesses execute in order, and the fences keep
// the compiler from moving any access across.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The block is full and more input follows (Hasher::update compresses a
// full block only then, so the last block is always there to finalize):
// compress it; the 16th block of a chunk closes the chunk and merges the
// completed subtrees on the stack.  pm is 16 words of LDS for the parents'
// messages.  Lane 0 writes the LDS.
__device__ __forceinline__ void fs_flush(FsState &s, FsRegs &r, uint32_t *pm, uint32_t lane) {
  const uint32_t qi = lane & 3u;
  uint32_t o[4];
  const bool last = r.done == 15;
  compress_quad(r.cv, (const uint32_t *)s.block, r.chunk, 64,
                last ? b3::CHUNK_END : r.done == 0 ? b3::CHUNK_START : 0u, qi, o);
  uint32_t c[8];
  quad_cv(o, c);
  if (last) {
    uint64_t total = r.chunk + 1;
    while ((total & 1) == 0) {
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          pm[i] = s.stack[r.depth - 1][i];
          pm[8 + i] = c[i];
        }
      }
      wave_lds_sync();  // the quads read what lane 0 wrote
      uint32_t cv0[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) cv0[i] = b3::IV[i];
      compress_quad(cv0, pm, 0, 64, b3::PARENT, qi, o);
      quad_cv(o, c);
      --r.depth;
      total >>= 1;
    }
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < 8; ++i) s.stack[r.depth][i] = c[i];
    }
    ++r.depth;
    ++r.chunk;
#pragma unroll
    for (int i = 0; i < 8; ++i) r.cv[i] = b3::IV[i];
    r.done = 0;
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) r.cv[i] = c[i];
    ++r.done;
  }
  if (lane == 0) {
    uint2 *b = (uint2 *)s.block;
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = make_uint2(0u, 0u);
  }
  r.len = 0;
}

// Absorb the N bytes of the little-endian words w (Hasher::update, N <= 64):
// the bytes that fit go into the open block, then, if more follow, the block
// is flushed and the rest start a new one.
template <int N, int NW>
__device__ __forceinline__ void fs_absorb(FsState &s, FsRegs &r, const uint32_t (&w)[NW], uint32_t *pm,
                                          uint32_t lane) {
  static_assert(N <= 64 && 4 * NW >= N, "one flush at most");
  const uint32_t n1 = 64u - r.len < (uint32_t)N ? 64u - r.len : (uint32_t)N;
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < N; ++i)
      if ((uint32_t)i < n1) s.block[r.len + i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
  }
  r.len += n1;
  if (n1 < (uint32_t)N) {
    wave_lds_sync();  // the block's bytes are in LDS
    fs_flush(s, r, pm, lane);
    if (lane == 0) {
#pragma unroll
      for (int i = 0; i < N; ++i)
        if ((uint32_t)i >= n1) s.block[i - n1] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
    }
    r.len = (uint32_t)N - n1;
  }
  wave_lds_sync();
}

// The root compression of the current state (Hasher::finalize_xof, output
// block 0): the open chunk's output merged with the stack, then the root
// compression.  Lane i of each quad gets its output words i, 4 + i (o[0],
// o[1]).
__device__ __forceinline__ void fs_root_output(const FsState &s, const FsRegs &r, uint32_t *pm, uint32_t lane,
                                               uint32_t o[4]) {
  const uint32_t qi = lane & 3u;
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cv[i] = r.cv[i];
  const uint32_t *m = (const uint32_t *)s.block;
  uint64_t counter = r.chunk;
  uint32_t len = r.len, flags = (r.done == 0 ? b3::CHUNK_START : 0u) | b3::CHUNK_END;
  for (int i = (int)r.depth - 1; i >= 0; --i) {
    compress_quad(cv, m, counter, len, flags, qi, o);
    uint32_t c[8];
    quad_cv(o, c);
    if (lane == 0) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        pm[j] = s.stack[i][j];
        pm[8 + j] = c[j];
      }
    }
    wave_lds_sync();
    m = pm;
#pragma unroll
    for (int j = 0; j < 8; ++j) cv[j] = b3::IV[j];
    counter = 0;
    len = 64;
    flags = b3::PARENT;
  }
  compress_quad(cv, m, 0, len, flags | b3::ROOT, qi, o);
}

// The first 16 bytes of the root output (Hasher::finalize_xof of 16 bytes),
// in every lane.
__device__ __forceinline__ void fs_squeeze16(const FsState &s, const FsRegs &r, uint32_t *pm, uint32_t lane,
                                             uint32_t out[4]) {
  uint32_t o[4];
  fs_root_output(s, r, pm, lane, o);
#pragma unroll
  for (int k = 0; k < 4; ++k) out[k] = __builtin_amdgcn_readlane(o[0], k);
}

// The first 32 bytes (Transcript::state_digest, the proof-of-work base), in
// every lane.
__device__ __forceinline__ void fs_digest32(const FsState &s, const FsRegs &r, uint32_t *pm, uint32_t lane,
                                            uint32_t (&out)[8]) {
  uint32_t o[4];
  fs_root_output(s, r, pm, lane, o);
  quad_cv(o, out);
}

// One zerocheck round's transcript steps, Transcript::observe of the two
// F128 values m0, m1 (tag bytes 3, 1, then the value's 16 bytes, each) and
// Transcript::sample of one F128 (tag bytes 4, 1, squeeze 16 bytes, absorb
// them; they are the sample).  Run by one wavefront.
__device__ __forceinline__ f128 fs_round2(FsState &s, FsRegs &r, const f128 m0, const f128 m1, uint32_t *pm,
                                          uint32_t lane) {
  const uint32_t w[10] = {3u | (1u << 8) | (m0.x << 16), (m0.x >> 16) | (m0.y << 16), (m0.y >> 16) | (m0.z << 16),
                          (m0.z >> 16) | (m0.w << 16),   (m0.w >> 16) | (3u << 16) | (1u << 24), m1.x, m1.y, m1.z,
                          m1.w, 4u | (1u << 8)};
  fs_absorb<38>(s, r, w, pm, lane);
  uint32_t v[4];
  fs_squeeze16(s, r, pm, lane, v);
  fs_absorb<16>(s, r, v, pm, lane);
  return make_uint4(v[0], v[1], v[2], v[3]);
}

// Transcript::observe of m0 and m1 (tag bytes 3, 1, then the value's 16
// bytes, each), run by one wavefront.
__device__ __forceinline__ void fs_observe2(FsState &s, FsRegs &r, const f128 m0, const f128 m1, uint32_t *pm,
                                            uint32_t lane) {
  const uint32_t w[9] = {3u | (1u << 8) | (m0.x << 16), (m0.x >> 16) | (m0.y << 16), (m0.y >> 16) | (m0.z << 16),
                         (m0.z >> 16) | (m0.w << 16),   (m0.w >> 16) | (3u << 16) | (1u << 24), m1.x, m1.y, m1.z,
                         m1.w};
  fs_absorb<36>(s, r, w, pm, lane);
}

// Transcript::sample of one F128 (tag bytes 4, 1, squeeze 16 bytes, absorb
// them), run by one wavefront.
__device__ __forceinline__ f128 fs_sample(FsState &s, FsRegs &r, uint32_t *pm, uint32_t lane) {
  const uint32_t t[1] = {4u | (1u << 8)};
  fs_absorb<2>(s, r, t, pm, lane);
  uint32_t v[4];
  fs_squeeze16(s, r, pm, lane, v);
  fs_absorb<16>(s, r, v, pm, lane);
  return make_uint4(v[0], v[1], v[2], v[3]);
}

// Transcript::observe_bytes of a u64 (tag byte 5, the length 8 as a u64,
// the 8 bytes), run by one wavefront.
__device__ __forceinline__ void fs_observe_u64(FsState &s, FsRegs &r, uint64_t v, uint32_t *pm, uint32_t lane) {
  const uint32_t w[5] = {5u | (8u << 8), 0u, (uint32_t)(v << 8), (uint32_t)(v >> 24), (uint32_t)(v >> 56)};
  fs_absorb<17>(s, r, w, pm, lane);
}

// A transcript step run by a workgroup: the state at g is staged through
// the LDS copy s, wavefront 0 runs body(s, r, pm, lane) and stores the
// scalar fields back, and the state returns to g.
template <class Body>
__device__ __forceinline__ void fs_block(FsState *g, FsState &s, Body body) {
  __shared__ uint32_t pm[16];
  const uint32_t *src = (const uint32_t *)g;
  uint32_t *dst = (uint32_t *)&s;
  for (uint32_t i = threadIdx.x; i < sizeof(FsState) / 4; i += blockDim.x) dst[i] = src[i];
  __syncthreads();
  if (threadIdx.x < 64) {
    const uint32_t lane = threadIdx.x;
    FsRegs r;
    fs_regs_load(s, r);
    // uniform copies of the scalar fields (every lane read the same LDS words)
#pragma unroll
    for (int i = 0; i < 8; ++i) r.cv[i] = __builtin_amdgcn_readfirstlane(r.cv[i]);
    r.chunk = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(r.chunk >> 32)) << 32) |
              __builtin_amdgcn_readfirstlane((uint32_t)r.chunk);
    r.len = __builtin_amdgcn_readfirstlane(r.len);
    r.done = __builtin_amdgcn_readfirstlane(r.done);
    r.depth = __builtin_amdgcn_readfirstlane(r.depth);
    body(s, r, pm, lane);
    if (lane == 0) fs_regs_store(s, r);
  }
  __syncthreads();
  uint32_t *back = (uint32_t *)g;
  for (uint32_t i = threadIdx.x; i < sizeof(FsState) / 4; i += blockDim.x) back[i] = dst[i];
}

// Observe msg[0], msg[1] with the state at g (a Ligerito round message: its
// challenge comes after the next proof of work).  msg may be memory thread 0
// itself has just written.
// With dg, the state's digest (the next proof of work's base) then goes to
// dg[0..7].
__device__ __forceinline__ void fs_observe2_block(FsState *g, FsState &s, const f128 *msg, uint32_t *dg = nullptr) {
  fs_block(g, s, [&](FsState &st, FsRegs &r, uint32_t *pm, uint32_t lane) {
    fs_observe2(st, r, msg[0], msg[1], pm, lane);
    if (dg) {
      uint32_t d[8];
      fs_digest32(st, r, pm, lane, d);
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 8; ++i) dg[i] = d[i];
      }
    }
  });
}

// One zerocheck round's transcript steps, run by a workgroup: the state at g
// is staged through the LDS copy s, wavefront 0 observes msg[0], msg[1] and
// samples the challenge into *rho (lane 0 stores it), and the state goes
// back to g.  msg may be memory thread 0 itself has just written.
__device__ __forceinline__ void fs_round_block(FsState *g, FsState &s, const f128 *msg, f128 *rho) {
  fs_block(g, s, [&](FsState &st, FsRegs &r, uint32_t *pm, uint32_t lane) {
    const f128 rv = fs_round2(st, r, msg[0], msg[1], pm, lane);
    if (lane == 0) *rho = rv;
  });
}

}  // namespace flock
