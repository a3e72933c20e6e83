"""Record the oracle's proof of the reference's ranked canary input.

Input: 2^18 compressions from SplitMix64 state 0xC0FFEE_BEEF ^ 2^18 (the
reference test rng, crates/flock-prover/src/r1cs_hashes/blake3.rs:4920-4931;
note: no rotl(log2, 29) mixing, unlike generate_compressions), Fiat-Shamir
domain b"flock-bench-v0", BLAKE3 transcript and Merkle hash.  The reference
asserts only the proof LENGTH (437,551 bytes, blake3.rs:4914); the digest
recorded here is the oracle's own and serves as the GPU parity fixture.

    python tests/golden/make_canary_golden.py   (about 3 minutes)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402

LOG2 = 18
RAW_SEED = 0xC0FFEEBEEF ^ (1 << LOG2)


def canary_blocks():
    # generate_compressions mixes seed ^ rotl(log2, 29); undo it to start the
    # stream at RAW_SEED.
    return O.generate_compressions(LOG2, RAW_SEED ^ ((LOG2 << 29) & 0xFFFFFFFFFFFFFFFF))


def main():
    blocks = canary_blocks()
    pf = O.prove(blocks)
    rc = O.verify(pf, blocks.size, None)
    out = {"log2": LOG2, "raw_seed": RAW_SEED, "proof_len": len(pf), "proof_blake3": O.blake3(pf).hex(),
           "reference_expected_len": 437551, "oracle_verify_rc": rc}
    json.dump(out, open(os.path.join(HERE, "canary_m32.json"), "w"), indent=1)
    print(out)


if __name__ == "__main__":
    main()
