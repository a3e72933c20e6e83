"""The library's host verifier (flock_verify_blake3, the checks of
Blake3Setup::verify, crates/flock-prover/src/r1cs_hashes/blake3.rs:3698) on
proofs made by the oracle: it accepts honest proofs, rejects a proof with
any single field corrupted, a wrong domain, size or commitment root, and
truncated or padded bytes, and agrees with the oracle's verifier on every
case.  Host only: runs without a GPU."""
import pytest

import gpu_util  # noqa: F401  (puts the package on sys.path)
import flockgpu
import oracle_lib as O
import proof_format as PF

DOMAIN = flockgpu.BENCH_DOMAIN


@pytest.fixture(scope="module")
def proof22():
    blocks = O.generate_compressions(8, 0x00C0FFEEBEEFD15C)
    return blocks, O.prove(blocks)


def _accepts(pf, n, domain=DOMAIN, root=None):
    try:
        flockgpu.verify(pf, n, domain, root)
        return True
    except flockgpu.VerifyError:
        return False


def test_accepts_honest_proof(proof22):
    blocks, pf = proof22
    flockgpu.verify(pf, blocks.size)
    # the commitment root the proof carries (bytes 7..39), as a recomputed root would be
    flockgpu.verify(pf, blocks.size, DOMAIN, pf[7:39])


def test_rejects_other_root_domain_and_size(proof22):
    blocks, pf = proof22
    bad_root = bytes([pf[7] ^ 1]) + pf[8:39]
    assert not _accepts(pf, blocks.size, root=bad_root)
    assert not _accepts(pf, blocks.size, domain=b"flock-bench-v1")
    assert O.verify(pf, blocks.size, None, b"flock-bench-v1") != 0
    assert not _accepts(pf, 2 * blocks.size)  # another m: parameters differ


def test_rejects_truncated_and_padded_bytes(proof22):
    blocks, pf = proof22
    for bad in (pf[:-1], pf + b"\x00", pf[:100], b""):
        assert not _accepts(bad, blocks.size)
        assert O.verify(bad, blocks.size, None) != 0


def test_error_message_names_the_failed_check(proof22):
    blocks, pf = proof22
    fields = {name: (off, ln) for name, off, ln in PF.fields(pf)}
    off, _ = fields["zerocheck.final_a"]
    bad = bytearray(pf)
    bad[off] ^= 1
    with pytest.raises(flockgpu.VerifyError, match="zerocheck"):
        flockgpu.verify(bytes(bad), blocks.size)


# One corrupted bit in each kind of field; each must be rejected by both
# verifiers (length fields are left alone: they only make the bytes unparsable).
TAMPER_FIELDS = ["commitment.root", "zerocheck.round1_ab", "zerocheck.round1_c", "zerocheck.rounds[3]",
                 "zerocheck.final_b", "zerocheck.final_c", "lincheck.rounds[5]", "lincheck.z_partial",
                 "ring_switch.s_hat_v[0]", "ring_switch.s_hat_v[1]", "ligerito.initial_root",
                 "ligerito.initial_proof.rows[0]", "ligerito.initial_proof.hashes", "ligerito.roots[0]",
                 "ligerito.recursive_proofs[0].rows[2]", "ligerito.final.yr", "ligerito.final.opening.rows[1]",
                 "ligerito.final.opening.hashes", "ligerito.sumcheck_msgs[0]", "ligerito.sumcheck_msgs[9]",
                 "ligerito.nonces", "ligerito.ood", "ligerito.fold_nonces"]


@pytest.mark.parametrize("field", TAMPER_FIELDS)
def test_rejects_corrupted_field(proof22, field):
    blocks, pf = proof22
    fields = {name: (off, ln) for name, off, ln in PF.fields(pf)}
    if field not in fields:
        pytest.skip(f"{field} absent at this size")
    off, ln = fields[field]
    if ln == 0:
        pytest.skip(f"{field} is empty at this size")
    bad = bytearray(pf)
    bad[off + ln // 2] ^= 0x10
    bad = bytes(bad)
    assert not _accepts(bad, blocks.size), field
    assert O.verify(bad, blocks.size, None) != 0, field


def test_accepts_padded_instance_and_other_sizes():
    # a non-power-of-two count (padding blocks) and the config-0 size
    for log2, seed, n in ((9, 0x57, 300), (12, 0x00C0FFEEBEEFD15C, None)):
        blocks = O.generate_compressions(log2, seed)
        if n is not None:
            blocks = blocks[:n].copy()
        pf = O.prove(blocks)
        flockgpu.verify(pf, blocks.size)
        assert O.verify(pf, blocks.size, blocks) == 0


def test_blake3_known_answer():
    # BLAKE3("") from the public BLAKE3 specification
    assert flockgpu.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    data = bytes(range(251)) * 7
    assert flockgpu.blake3(data) == O.blake3(data)
