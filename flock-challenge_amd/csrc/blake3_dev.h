// BLAKE3 compression on one lane (public BLAKE3 specification).  Used by the
// Merkle leaf/parent kernels, the proof-of-work scan and the witness kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b3 {

enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };

__device__ __constant__ static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                                      0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t n) { return __builtin_amdgcn_alignbit(x, x, n); }

#define B3_G(a, b, c, d, mx, my)        \
  do {                                  \
    a = a + b + (mx);                   \
    d = rotr(d ^ a, 16);                \
    c = c + d;                          \
    b = rotr(b ^ c, 12);                \
    a = a + b + (my);                   \
    d = rotr(d ^ a, 8);                 \
    c = c + d;                          \
    b = rotr(b ^ c, 7);                 \
  } while (0)

// One round over the message schedule m[] already permuted for this round.
#define B3_ROUND(m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15) \
  do {                                                                               \
    B3_G(s0, s4, s8, s12, m0, m1);                                                   \
    B3_G(s1, s5, s9, s13, m2, m3);                                                   \
    B3_G(s2, s6, s10, s14, m4, m5);                                                  \
    B3_G(s3, s7, s11, s15, m6, m7);                                                  \
    B3_G(s0, s5, s10, s15, m8, m9);                                                  \
    B3_G(s1, s6, s11, s12, m10, m11);                                                \
    B3_G(s2, s7, s8, s13, m12, m13);                                                 \
    B3_G(s3, s4, s9, s14, m14, m15);                                                 \
  } while (0)

// The seven rounds with the message permutation applied symbolically
// (round r reads m[PERM^r[i]]).
#define B3_ROUNDS(m)                                                                                             \
  B3_ROUND(m[0], m[1], m[2], m[3], m[4], m[5], m[6], m[7], m[8], m[9], m[10], m[11], m[12], m[13], m[14], m[15]); \
  B3_ROUND(m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13], m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]); \
  B3_ROUND(m[3], m[4], m[10], m[12], m[13], m[2], m[7], m[14], m[6], m[5], m[9], m[0], m[11], m[15], m[8], m[1]); \
  B3_ROUND(m[10], m[7], m[12], m[9], m[14], m[3], m[13], m[15], m[4], m[0], m[11], m[2], m[5], m[8], m[1], m[6]); \
  B3_ROUND(m[12], m[13], m[9], m[11], m[15], m[10], m[14], m[8], m[7], m[2], m[5], m[3], m[0], m[1], m[6], m[4]); \
  B3_ROUND(m[9], m[14], m[11], m[5], m[8], m[12], m[15], m[1], m[13], m[3], m[0], m[10], m[2], m[6], m[4], m[7]); \
  B3_ROUND(m[11], m[15], m[5], m[0], m[1], m[9], m[8], m[6], m[14], m[10], m[2], m[12], m[3], m[4], m[7], m[13]);

// Chaining-value compression: cv <- first 8 words of compress(cv, m, ...).
__device__ __forceinline__ void compress_cv(uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                                            uint32_t block_len, uint32_t flags) {
  uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
  uint32_t s8 = IV[0], s9 = IV[1], s10 = IV[2], s11 = IV[3];
  uint32_t s12 = (uint32_t)counter, s13 = (uint32_t)(counter >> 32), s14 = block_len, s15 = flags;
  B3_ROUNDS(m)
  cv[0] = s0 ^ s8; cv[1] = s1 ^ s9; cv[2] = s2 ^ s10; cv[3] = s3 ^ s11;
  cv[4] = s4 ^ s12; cv[5] = s5 ^ s13; cv[6] = s6 ^ s14; cv[7] = s7 ^ s15;
}

// Full 16-word output (for XOF blocks).
__device__ __forceinline__ void compress_full(const uint32_t cv[8], const uint32_t m[16], uint64_t counter,
                                              uint32_t block_len, uint32_t flags, uint32_t out[16]) {
  uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
  uint32_t s8 = IV[0], s9 = IV[1], s10 = IV[2], s11 = IV[3];
  uint32_t s12 = (uint32_t)counter, s13 = (uint32_t)(counter >> 32), s14 = block_len, s15 = flags;
  B3_ROUNDS(m)
  out[0] = s0 ^ s8; out[1] = s1 ^ s9; out[2] = s2 ^ s10; out[3] = s3 ^ s11;
  out[4] = s4 ^ s12; out[5] = s5 ^ s13; out[6] = s6 ^ s14; out[7] = s7 ^ s15;
  out[8] = s8 ^ cv[0]; out[9] = s9 ^ cv[1]; out[10] = s10 ^ cv[2]; out[11] = s11 ^ cv[3];
  out[12] = s12 ^ cv[4]; out[13] = s13 ^ cv[5]; out[14] = s14 ^ cv[6]; out[15] = s15 ^ cv[7];
}

// Proof of work (challenger.rs:1246-1300): BLAKE3(digest | nonce_le | 0^24)
// as one CHUNK_START|CHUNK_END|ROOT block; the leading `bits` bits of its
// output bytes must be zero.
__device__ __forceinline__ bool pow_ok(const uint32_t (&dg)[8], uint64_t nonce, uint32_t bits) {
  uint32_t m[16] = {dg[0], dg[1], dg[2], dg[3], dg[4], dg[5], dg[6], dg[7], (uint32_t)nonce, (uint32_t)(nonce >> 32),
                    0, 0, 0, 0, 0, 0};
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cv[i] = IV[i];
  compress_cv(cv, m, 0, 64, CHUNK_START | CHUNK_END | ROOT);
  const uint64_t be = ((uint64_t)__builtin_bswap32(cv[0]) << 32) | __builtin_bswap32(cv[1]);
  return bits == 0 || (bits <= 64 && (be >> (64 - bits)) == 0);
}

}  // namespace b3
