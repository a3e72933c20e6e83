"""Parity of the GPU prover (libflockgpu.so, through the C ABI) with the CPU
oracle: identical proof bundle bytes for identical compressions, accepted by
the oracle's verifier.  The m=32 case (2^18 compressions, the benchmark size)
is checked against the committed canary fixture: its length equals the
reference's own canary assertion (crates/flock-prover/src/r1cs_hashes/
blake3.rs:4910-4939) and its BLAKE3 digest equals the oracle's proof.
"""
import json
import os

import numpy as np
import pytest

import gpu_util
import flockgpu
import oracle_lib as O
import bundle_layout as BL
import proof_format as PF

pytestmark = pytest.mark.gpu

WORKER_SEED = 0x00C0FFEEBEEFD15C  # worker/src/main.rs:28
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ctx():
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    c = flockgpu.Context(0)
    yield c
    c.close()


def _say(msg):
    print(msg, flush=True)


def _gpu_prove(ctx, blocks):
    s = flockgpu.Blake3Setup(ctx, blocks.size)
    try:
        return s.prove_fast(blocks)
    finally:
        s.close()


def _check_same(gpu, ref):
    diff = PF.first_difference(gpu, ref)
    assert diff is None, diff


@pytest.mark.parametrize("log2,seed", [(8, WORKER_SEED), (9, 7), (10, 1)])
def test_proof_bytes_match_oracle(ctx, log2, seed):
    _say(f"prove log2={log2} seed={seed:#x}")
    blocks = O.generate_compressions(log2, seed)
    gpu = _gpu_prove(ctx, blocks)
    ref = O.prove(blocks)
    _check_same(gpu, ref)
    assert O.verify(gpu, blocks.size, blocks) == 0


@pytest.mark.parametrize("n", [200, 3000])
def test_non_power_of_two_batches_pad_like_the_oracle(ctx, n):
    _say(f"prove n={n}")
    log2 = int(np.ceil(np.log2(n)))
    blocks = O.generate_compressions(log2, 0x9A110F11)[:n].copy()
    gpu = _gpu_prove(ctx, blocks)
    ref = O.prove(blocks)
    _check_same(gpu, ref)
    assert O.verify(gpu, n, blocks) == 0


def test_config0_worker_seed_matches_oracle(ctx):
    """2^12 compressions (m = 26), the worker warm-up seed."""
    _say("prove log2=12 worker seed")
    blocks = O.generate_compressions(12, WORKER_SEED)
    gpu = _gpu_prove(ctx, blocks)
    ref = O.prove(blocks)
    _check_same(gpu, ref)


def test_seeded_and_device_paths_match_host_blocks(ctx):
    _say("seeded/device paths log2=11")
    log2, seed = 11, 0x1234
    blocks = O.generate_compressions(log2, seed)
    s = flockgpu.Blake3Setup(ctx, 1 << log2)
    try:
        host = s.prove_fast(blocks)
        seeded = s.prove_seeded(log2, seed)
        d = flockgpu.generate_compressions_device(ctx, log2, seed)
        assert np.array_equal(d.download(np.uint8), blocks.view(np.uint8).reshape(-1))
        dev = s.prove_fast_device(d.ptr, 1 << log2)
        again = s.prove_fast(blocks)
    finally:
        s.close()
    assert seeded == host and dev == host and again == host


def test_canary_m32_matches_fixture(ctx):
    """2^18 compressions from the reference canary seed: 437,551 bytes, and
    the same bytes as the oracle (BLAKE3 digest of its proof)."""
    g = json.load(open(os.path.join(GOLDEN, "canary_m32.json")))
    _say("prove canary m=32")
    s = flockgpu.Blake3Setup(ctx, 1 << g["log2"])
    seed = g["raw_seed"] ^ (g["log2"] << 29)
    try:
        proof = s.prove_seeded(g["log2"], seed)
        assert len(proof) == g["reference_expected_len"]
        assert O.blake3(proof).hex() == g["proof_blake3"]
        # field order and counts fixed by the declarations and m32_fast.toml
        BL.check_layout(proof, 32)
        # the library's own verifier, with the root recomputed on the GPU
        _say("verify canary m=32")
        assert s.commitment_root_seeded(g["log2"], seed) == proof[7:39]
        s.verify_seeded(g["log2"], seed, proof)
    finally:
        s.close()


def test_commitment_root_and_verifier_on_host_blocks(ctx):
    _say("commitment root + verify log2=10")
    blocks = O.generate_compressions(10, 0x5EED)
    s = flockgpu.Blake3Setup(ctx, blocks.size)
    try:
        proof = s.prove_fast(blocks)
        assert s.commitment_root(blocks) == proof[7:39]
        s.verify(proof, blocks)
        other = O.generate_compressions(10, 0x5EEE)
        with pytest.raises(flockgpu.VerifyError, match="root"):
            s.verify(proof, other)
    finally:
        s.close()


def test_shape_errors_are_reported(ctx):
    with pytest.raises(flockgpu.FlockError) as e:
        flockgpu.Blake3Setup(ctx, 4)  # m = 17: no Ligerito configuration
    assert e.value.code == -1
    s = flockgpu.Blake3Setup(ctx, 256)
    try:
        with pytest.raises(flockgpu.FlockError) as e:
            s.prove_fast(O.generate_compressions(7, 1))
        assert e.value.code == -1
    finally:
        s.close()


@pytest.mark.parametrize("log2", [8, 12, 16, 18])
def test_device_and_host_transcripts_give_the_same_bytes(ctx, log2):
    """The Fiat-Shamir steps of the zerocheck rounds and of the Ligerito
    fold rounds (their proofs of work included) on the device (the first
    proof of a size captures the zerocheck rounds in a HIP graph, later ones
    replay it) and on the host (the default) give identical proofs, for three
    inputs in a row; at 2^8 and 2^12 they also equal the oracle's."""
    _say(f"device vs host transcript log2={log2}")
    s = flockgpu.Blake3Setup(ctx, 1 << log2)
    try:
        for k in range(3):
            seed = 0xD7F5_0000 | (k << 8) | log2
            ctx.set_option("device_transcript", 1)
            dev = s.prove_seeded(log2, seed)
            ctx.set_option("device_transcript", 0)
            host = s.prove_seeded(log2, seed)
            ctx.set_option("device_transcript", 1)
            ctx.set_option("zc_graph", 0)
            direct = s.prove_seeded(log2, seed)
            ctx.set_option("zc_graph", 1)
            _check_same(dev, host)
            _check_same(direct, host)
            if log2 <= 12 and k == 0:
                _check_same(dev, O.prove(O.generate_compressions(log2, seed)))
    finally:
        ctx.set_option("device_transcript", 0)
        ctx.set_option("zc_graph", 1)
        s.close()
