"""The bundle's field order and counts (tests/bundle_layout.py) on an oracle
proof of 2^12 compressions (m = 26, the m26_fast ladder), and that the check
fails when two sections of different lengths trade places or a count is
wrong.  The m = 32 proof is checked the same way in the GPU canary test."""
import struct

import pytest

import bundle_layout as BL
import oracle_lib as O
import proof_format as PF


@pytest.fixture(scope="module")
def proof26():
    return O.prove(O.generate_compressions(12, 0x00C0FFEEBEEFD15C))


def test_expected_counts_from_the_ladders():
    e26, e32 = BL.expected_counts(26), BL.expected_counts(32)
    assert (e26["roots"], e26["recursive_proofs"], e26["nonces"], e26["ood"]) == (3, 2, 4, 3)
    assert (e32["roots"], e32["recursive_proofs"], e32["nonces"], e32["ood"]) == (5, 4, 6, 5)
    # tapered fold grinds 13..8 | 8,7,6 | 5,4,3 | 2,1 (the 0-bit one emits none)
    assert e26["fold_nonces"] == 14
    assert e32["fold_nonces"] == 21
    assert e26["sumcheck_msgs"] == 1 + 15 + 3 + 3
    assert e32["sumcheck_msgs"] == 1 + 21 + 5 + 5


def test_oracle_m26_proof_has_the_fixed_layout(proof26):
    BL.check_layout(proof26, 26)


def _span(proof, first, last):
    """Byte range covering the proof_format fields whose names start with
    `first` through the last field starting with `last`."""
    fs = PF.fields(proof)
    lo = min(o for n, o, _ in fs if n.startswith(first))
    hi = max(o + ln for n, o, ln in fs if n.startswith(last))
    return lo, hi


def _swap(proof, a, b):
    (a0, a1), (b0, b1) = a, b
    assert a1 <= b0
    return proof[:a0] + proof[b0:b1] + proof[a1:b0] + proof[a0:a1] + proof[b1:]


@pytest.mark.parametrize("x,y", [("ligerito.nonces", "ligerito.ood"),
                                 ("ligerito.ood", "ligerito.fold_nonces"),
                                 ("ligerito.roots", "ligerito.recursive_proofs"),
                                 ("zerocheck.rounds", "zerocheck.final_c")])
def test_swapped_sections_fail(proof26, x, y):
    a, b = _span(proof26, x, x), _span(proof26, y, y)
    bad = _swap(proof26, a, b)
    assert len(bad) == len(proof26)
    with pytest.raises(AssertionError):
        BL.check_layout(bad, 26)


def test_wrong_count_fails(proof26):
    off = dict((n, o) for n, o, _ in PF.fields(proof26))["ligerito.sumcheck_msgs.len"]
    n = struct.unpack_from("<Q", proof26, off)[0]
    bad = bytearray(proof26)
    struct.pack_into("<Q", bad, off, n + 1)
    with pytest.raises(AssertionError, match="sumcheck_transcript"):
        BL.check_layout(bytes(bad), 26)
