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

// generate_compressions(log2, seed) on the device (benchmark-tools/common/src/lib.rs:20-33).
int generate_compressions_device(Ctx *ctx, hipStream_t st, uint32_t log2, uint64_t seed, flock_compression *d_out);
// The commitment root of the witness of n_blocks compressions in device
// memory (the prover's witness + commit steps); waits for the stream.
int commitment_root_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_blocks,
                           uint8_t root[32]);

}  // namespace flock
