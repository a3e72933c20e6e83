// Whole-proof entry points (prover.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/flock_gpu.h"
#include "ctx.h"

namespace flock {

// Proof bundle bytes (FLOCK v4 header + bincode payload) for n_blocks device-
// resident compressions, Fiat-Shamir domain `domain`, BLAKE3 transcript and
// Merkle hash, Fast profile, rate 1/2 (the benchmark setup).
int prove_blake3_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_blocks,
                        const uint8_t *domain, size_t domain_len, std::vector<uint8_t> &out);

// The buffers the commitment step leaves for the rest of the proof: packed
// z, a, b (2^(m-6) u64 words each), the rate-1/2 codeword (2^(m-12) positions
// x 64 lanes F128) and its Merkle tree (2^(m-11) - 1 nodes).
struct ProofFront {
  void *z = nullptr, *a = nullptr, *b = nullptr, *codeword = nullptr, *tree = nullptr;
};
// The proof after the commitment, whose root is already known: the same
// bytes as prove_blake3_device for the compressions the front came from.
int prove_after_commit_device(Ctx *ctx, hipStream_t st, int m, const ProofFront &fb, const uint8_t root[32],
                              const uint8_t *domain, size_t domain_len, std::vector<uint8_t> &out);

// generate_compressions(log2, seed) on the device (benchmark-tools/common/src/lib.rs:20-33).
int generate_compressions_device(Ctx *ctx, hipStream_t st, uint32_t log2, uint64_t seed, flock_compression *d_out);
// The commitment root of the witness of n_blocks compressions in device
// memory (the prover's witness + commit steps); waits for the stream.
int commitment_root_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_blocks,
                           uint8_t root[32]);

}  // namespace flock
