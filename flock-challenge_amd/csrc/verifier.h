// Host verifier of a proof bundle (Blake3Setup::verify,
// crates/flock-prover/src/r1cs_hashes/blake3.rs:3698).
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>

#include "../../include/flock_gpu.h"

namespace flock {

// FLOCK_OK when the bundle of a proof of n_blocks compressions verifies under
// the Fiat-Shamir domain; FLOCK_E_VERIFY (reason in err) when it does not;
// FLOCK_E_SHAPE for an unsupported size.  recomputed_root, when given, is the
// commitment root of the compressions' witness, which the bundle's must equal.
int verify_blake3(const uint8_t *bytes, size_t len, uint64_t n_blocks, const uint8_t *domain, size_t domain_len,
                  const uint8_t *recomputed_root, std::string &err);

}  // namespace flock
