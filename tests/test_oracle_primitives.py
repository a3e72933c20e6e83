"""Pin the C oracle's primitives against known answers, the reference's own
fixtures and independent restatements (tests/pyref.py)."""
import json
import os
import random

import numpy as np
import pytest

import oracle_lib as O
import pyref

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "field_golden.json")))


# ----------------------------------------------------------------- GF(2^128)
def test_ghash_known_answers():
    # field/gf2_128.rs test ghash_reduction_smoking_gun
    x = 2
    assert O.mul(x, 1 << 127) == 0x87
    assert O.mul(x, 1 << 63) == 1 << 64
    assert O.mul(1 << 64, 1 << 64) == 0x87
    assert O.mul(x, x) == 4


def test_f128_mul_matches_independent_and_portable():
    rng = random.Random(1)
    for _ in range(300):
        a, b = rng.getrandbits(128), rng.getrandbits(128)
        want = pyref.gf_mul(a, b)
        assert O.mul(a, b) == want
        assert O.i(O.lib().of128_mul_portable(O.f(a), O.f(b))) == want
        u = O.lib().of128_mul_unreduced(O.f(a), O.f(b))
        assert O.i(O.lib().of256_reduce(u)) == want


def test_f128_inverse_and_mul_by_x():
    rng = random.Random(2)
    for a in [1, 2, 1 << 127, (1 << 128) - 1] + [rng.getrandbits(128) for _ in range(8)]:
        inv = O.i(O.lib().of128_inv(O.f(a)))
        assert O.mul(a, inv) == 1
        assert O.i(O.lib().of128_mul_by_x(O.f(a))) == pyref.gf_mul(a, 2)


# ----------------------------------------------------------------- GF(2^8), phi8
def test_f8_mul_inv():
    for a in range(256):
        for b in range(0, 256, 7):
            assert O.lib().of8_mul(a, b) == pyref.gf8_mul(a, b)
        if a:
            assert O.lib().of8_mul(a, O.lib().of8_inv(a)) == 1


def test_phi8_matches_reference_table():
    table = GOLD["phi8_table"]
    for a in range(256):
        lo, hi = (int(x, 16) for x in table[a])
        assert O.i(O.lib().of_phi8(a)) == lo | (hi << 64), a


def test_phi8_is_a_field_homomorphism():
    phi = [O.i(O.lib().of_phi8(a)) for a in range(256)]
    for a in range(0, 256, 5):
        for b in range(256):
            assert phi[O.lib().of8_mul(a, b)] == pyref.gf_mul(phi[a], phi[b])
            assert phi[a ^ b] == phi[a] ^ phi[b]


# ----------------------------------------------------------------- BLAKE3
def test_blake3_empty_and_one_byte_known_digests():
    # Published BLAKE3 digests (official test vectors): "" and the single byte 0x00.
    assert O.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert O.blake3(b"\x00").hex() == "2d3adedff11b61f14c886e35afa036736dcd87a74d27b5c1510225d0f592e213"


def test_blake3_compress_of_empty_is_hash_of_empty():
    # r1cs_hashes/blake3.rs:4353-4369: compress(IV, 0^16, 0, 0, CS|CE|ROOT) = BLAKE3("").
    cv = np.array(pyref.IV, np.uint32)
    blk = np.zeros(16, np.uint32)
    out = np.zeros(16, np.uint32)
    O.lib().ob3_compress(O.ptr(cv), O.ptr(blk), 0, 0, 1 | 2 | 8, O.ptr(out))
    assert out[:8].tobytes() == O.blake3(b"")


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 127, 128, 1023, 1024, 1025, 2048, 2049, 3072, 4097, 5 * 1024 + 3])
def test_blake3_matches_independent(n):
    data = bytes((i % 251) for i in range(n))
    assert O.blake3(data) == pyref.blake3(data)
    assert O.leaf_cv(data) == pyref.blake3_non_root_cv(data)


def test_blake3_xof_and_parent():
    a, b = bytes(range(32)), bytes(range(32, 64))
    out = np.zeros(32, np.uint8)
    an, bn = np.frombuffer(a, np.uint8).copy(), np.frombuffer(b, np.uint8).copy()
    O.lib().ob3_parent_cv(O.ptr(an), O.ptr(bn), O.ptr(out))
    assert out.tobytes() == pyref.blake3_parent_cv(a, b)


# ----------------------------------------------------------------- generator
def test_splitmix64_known_outputs():
    # SplitMix64 from state 0: the published first outputs.
    st = O.ctypes.c_uint64(0)
    outs = [O.lib().o_splitmix_next(O.ctypes.byref(st)) for _ in range(3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generate_compressions_closed_form():
    for log2, seed in [(3, 0x00C0FFEEBEEFD15C), (5, 7)]:
        got = O.generate_compressions(log2, seed)
        st = seed ^ (((log2 << 29) | (log2 >> 35)) & 0xFFFFFFFFFFFFFFFF)
        for k in range(1 << log2):
            vals = []
            for _ in range(25):
                st, z = pyref.splitmix64(st)
                vals.append(z & 0xFFFFFFFF)
            assert list(got[k]["cv"]) == vals[:8]
            assert list(got[k]["m"]) == vals[8:24]
            assert int(got[k]["counter"]) == vals[24]  # upper 32 bits zero
            assert got[k]["block_len"] == 64 and got[k]["flags"] == 11
