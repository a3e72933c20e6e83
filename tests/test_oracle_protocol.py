"""Pin the oracle's BLAKE3 R1CS, witness, F8 extension, and whole-proof
transcript (prove -> verify round trips, tamper rejection)."""
import ctypes
import os
import random

import numpy as np
import pytest

import oracle_lib as O
import pyref

K = 1 << 14


def _sparse():
    a = ctypes.c_void_p()
    b = ctypes.c_void_p()
    O.lib().o_r1cs_matrices(ctypes.byref(a), ctypes.byref(b))
    return a, b


def _apply(s, z):
    out = np.zeros(256, np.uint64)
    O.lib().o_apply_block(s, O.ptr(z), O.ptr(out))
    return out


def _bits(words, base, n=32):
    v = 0
    for i in range(n):
        idx = base + i
        v |= int((int(words[idx >> 6]) >> (idx & 63)) & 1) << i
    return v


def _rand_comp(rng, flags=11):
    c = np.zeros(1, O.COMPRESSION_DTYPE)[0]
    c["cv"] = [rng.getrandbits(32) for _ in range(8)]
    c["m"] = [rng.getrandbits(32) for _ in range(16)]
    c["counter"] = rng.getrandbits(64)
    c["block_len"] = 64
    c["flags"] = flags
    return c


def test_witness_out_slots_equal_blake3_compression():
    # blake3.rs:4393-4419 with the reference's own seed.
    rng = random.Random(0x12345678)
    for _ in range(4):
        c = _rand_comp(rng, flags=1 | 2 | 8)
        z, a, b = O.block_witness(c)
        want = pyref.compress(list(map(int, c["cv"])), list(map(int, c["m"])), int(c["counter"]), 64, 11)
        for w in range(8):
            assert _bits(z, 256 + 32 * w) == want[w]
            assert _bits(z, 15153 + 32 * w) == want[w + 8]
        # inputs at their slots (blake3.rs:178-189)
        assert _bits(z, 512, 1) == 1
        assert _bits(z, 513) == int(c["m"][0])
        assert _bits(z, 1025) == int(c["counter"]) & 0xFFFFFFFF
        assert _bits(z, 1057) == int(c["counter"]) >> 32
        assert _bits(z, 1089) == 64 and _bits(z, 1121) == 11
        # padding bits are zero
        assert _bits(z, 15409, 32) == 0 and _bits(z, K - 32, 32) == 0


def test_row_witnesses_equal_matrix_products_and_satisfy_r1cs():
    """a = A_0 z, b = B_0 z (the fast row semantics equal the matrices) and
    (A z) & (B z) == z (C = I); blake3.rs:4421-4440, 4464+."""
    A, B = _sparse()
    rng = random.Random(0xCAFEF00D)
    for i in range(3):
        c = _rand_comp(rng)
        z, a, b = O.block_witness(c)
        assert np.array_equal(_apply(A, z), a)
        assert np.array_equal(_apply(B, z), b)
        assert np.array_equal(a & b, z)


def test_mutated_witness_fails_r1cs():
    # blake3.rs:4443-4462: flip carry bit 5 of ADD_A2 in G #10.
    A, B = _sparse()
    rng = random.Random(0xBEEFF00D)
    c = _rand_comp(rng)
    z, _, _ = O.block_witness(c)
    idx = 1153 + 250 * 10 + 31 * 4 + 5
    z[idx >> 6] ^= np.uint64(1 << (idx & 63))
    assert not np.array_equal(_apply(A, z) & _apply(B, z), z)


def test_matrix_rows_layout():
    A, B = _sparse()
    z = np.zeros(256, np.uint64)
    # Constant row: z[512] * z[512]; input row s: z[s] * z[512]; padding rows empty.
    z[512 >> 6] = np.uint64(1 << (512 & 63))
    a = _apply(A, z)
    b = _apply(B, z)
    assert (int(a[8]) >> 0) & 1 == 1 and (int(b[8]) >> 0) & 1 == 1
    # all-zero z except the constant: every input row has a = 0, b = 1
    assert _bits(b, 0, 32) == 0xFFFFFFFF and _bits(a, 0, 32) == 0
    assert _bits(a, 15409, 32) == 0 and _bits(b, 15409, 32) == 0


def _f8_ntt_reference(values_s):
    """S -> Lambda by the reference's own algorithm (ntt.rs:26-120 restated):
    inverse LCH NTT on beta = 0, forward on beta = 64, k = 6."""
    def next_s(s, root):
        return pyref.gf8_mul(s, s) ^ pyref.gf8_mul(root, s)

    def twiddles(k, beta):
        n = 1 << k
        tw = [0] * (n - 1)
        ln = 1 << (k - 1)
        layer = [beta ^ (2 * i) for i in range(ln)]
        s_at_root = 1
        write_at = ln
        for i in range(ln):
            tw[write_at - 1 + i] = layer[i]
        for _ in range(1, k):
            write_at >>= 1
            nsr = next_s(layer[1] ^ layer[0], s_at_root)
            for i in range(write_at):
                layer[i] = next_s(layer[2 * i], s_at_root)
            s_at_root = nsr
            inv = pyref.gf_pow(s_at_root, 254, 0x11B, 8)
            for j in range(write_at):
                tw[write_at - 1 + j] = pyref.gf8_mul(inv, layer[j])
        return tw

    def fft(v, tw, idx):
        n = len(v)
        if n == 1:
            return v
        half = n // 2
        lam = tw[idx - 1]
        for i in range(half):
            w = v[half + i]
            v[i] ^= pyref.gf8_mul(lam, w)
            v[half + i] = w ^ v[i]
        return fft(v[:half], tw, 2 * idx) + fft(v[half:], tw, 2 * idx + 1)

    def ifft(v, tw, idx):
        n = len(v)
        if n == 1:
            return v
        half = n // 2
        lo = ifft(v[:half], tw, 2 * idx)
        hi = ifft(v[half:], tw, 2 * idx + 1)
        v = lo + hi
        lam = tw[idx - 1]
        for i in range(half):
            v[half + i] ^= v[i]
            v[i] ^= pyref.gf8_mul(lam, v[half + i])
        return v

    coeffs = ifft(list(values_s), twiddles(6, 0), 1)
    return fft(coeffs, twiddles(6, 64), 1)


def test_f8_extension_matrix_matches_lch_ntt():
    M = np.zeros((64, 64), np.uint8)
    O.lib().o_ext_matrix(O.ptr(M))
    rng = random.Random(9)
    for trial in range(6):
        bits = [rng.getrandbits(1) for _ in range(64)] if trial else [1] + [0] * 63
        want = _f8_ntt_reference(bits)
        got = [0] * 64
        for j in range(64):
            acc = 0
            for s in range(64):
                if bits[s]:
                    acc ^= int(M[j, s])
            got[j] = acc
        assert got == want


def test_lagrange_weights_interpolate():
    # sum_i L_i(z) * p(node_i) == p(z) for a random degree-<64 polynomial on phi8 nodes.
    rng = random.Random(3)
    coeffs = [rng.getrandbits(128) for _ in range(64)]

    def p(x):
        acc = 0
        for c in reversed(coeffs):
            acc = pyref.gf_mul(acc, x) ^ c
        return acc

    z = rng.getrandbits(128)
    phi = [O.i(O.lib().of_phi8(a)) for a in range(128)]
    for fn, nodes in ((O.lib().o_lagrange_s, phi[:64]), (O.lib().o_lagrange_lambda, phi[64:])):
        w = np.zeros(128, np.uint64)
        fn(O.f(z), O.ptr(w))
        ws = O.f128_array_to_ints(w)
        acc = 0
        for wi, node in zip(ws, nodes):
            acc ^= pyref.gf_mul(wi, p(node))
        assert acc == p(z)


class _Csr(ctypes.Structure):
    _fields_ = [("row_ptr", ctypes.POINTER(ctypes.c_uint32)), ("cols", ctypes.POINTER(ctypes.c_uint32)),
                ("nnz", ctypes.c_size_t)]


def _csr_arrays(handle):
    s = ctypes.cast(handle, ctypes.POINTER(_Csr)).contents
    ptr = np.ctypeslib.as_array(s.row_ptr, shape=(K + 1,)).astype(np.uint64)
    cols = np.ctypeslib.as_array(s.cols, shape=(s.nnz,)).astype(np.uint64)
    return ptr, cols


def _framed_matrix(ptr, cols):
    """absorb_matrix (crates/flock-core/src/r1cs.rs:431-440) as one byte
    string: rows, cols, then per row its length and its column indices, all
    u64 little-endian; assembled with numpy, row by row in one vector."""
    lens = np.diff(ptr)
    # each row contributes [len, col...]: place the lengths at the row starts
    out = np.empty(K + cols.size, np.uint64)
    starts = ptr[:-1] + np.arange(K, dtype=np.uint64)
    mask = np.ones(out.size, bool)
    mask[starts.astype(np.int64)] = False
    out[starts.astype(np.int64)] = lens
    out[mask] = cols
    return np.array([K, K], np.uint64).tobytes() + out.astype("<u8").tobytes()


def test_statement_digest_framing():
    """BlockR1cs::statement_digest (r1cs.rs:343-365, 431-459) assembled
    independently from the oracle's matrices and hashed by the library's own
    BLAKE3 (a separate implementation): equal to the oracle's digest for
    m = 22 and 32; different m give different digests.  The matrices
    themselves are pinned by R1CS satisfaction above; the digest value has
    no reference datum (only the canary proof's length depends on it)."""
    import gpu_util  # noqa: F401  (puts the package on sys.path)
    import flockgpu
    a, b = _sparse()
    fa, fb = _framed_matrix(*_csr_arrays(a)), _framed_matrix(*_csr_arrays(b))
    ident = np.empty(2 * K, np.uint64)
    ident[0::2] = 1
    ident[1::2] = np.arange(K, dtype=np.uint64)
    fc = np.array([K, K], np.uint64).tobytes() + ident.astype("<u8").tobytes()
    for m in (22, 32):
        head = b"flock-r1cs-stmt-v1" + np.array([m, 14, 6], "<u8").tobytes() + b"\x00"  # RowMajor
        want = flockgpu.blake3(head + fa + fb + fc)
        assert O.statement_digest(m) == want, m
    assert O.statement_digest(22) != O.statement_digest(23)


@pytest.fixture(scope="module")
def proof22():
    blocks = O.generate_compressions(8, 0x00C0FFEEBEEFD15C)
    return blocks, O.prove(blocks)


def test_prove_verify_roundtrip_m22(proof22):
    blocks, pf = proof22
    assert pf[:7] == b"FLOCK\x04\x02"
    assert O.verify(pf, blocks.size, blocks) == 0


def test_proof_is_deterministic(proof22):
    blocks, pf = proof22
    assert O.prove(blocks) == pf


def test_non_power_of_two_batch_verifies():
    # 200 compressions pad to 256 slots with the all-zero compression (blake3.rs:2954).
    blocks = O.generate_compressions(8, 77)[:200].copy()
    pf = O.prove(blocks)
    assert O.verify(pf, blocks.size, blocks) == 0


def test_too_small_batch_has_no_ligerito_config():
    # m = 14 + 3 has no embedded Ligerito ladder (configs exist for m = 22..35),
    # where the reference panics (prover.rs:206-208); the oracle reports it.
    blocks = O.generate_compressions(3, 77)[:5].copy()
    with pytest.raises(AssertionError):
        O.prove(blocks)


@pytest.mark.parametrize("offset", [7 + 5, 7 + 40, 7 + 64 + 100, 7 + 64 + 1500, 3000, 9000, 40000, -300, -40, -1])
def test_tampered_proof_is_rejected(proof22, offset):
    blocks, pf = proof22
    bad = bytearray(pf)
    bad[offset] ^= 0x01
    assert O.verify(bytes(bad), blocks.size, blocks) != 0


def test_wrong_statement_is_rejected(proof22):
    blocks, pf = proof22
    other = blocks.copy()
    other[3]["m"][0] ^= 1
    assert O.verify(pf, blocks.size, other) != 0
    assert O.verify(pf + b"\x00", blocks.size, blocks) != 0
    assert O.verify(pf[:-1], blocks.size, blocks) != 0


@pytest.mark.slow
@pytest.mark.skipif(not os.environ.get("FLOCK_SLOW_TESTS"), reason="3-minute m=32 oracle proof; set FLOCK_SLOW_TESTS=1")
def test_ranked_canary_length_matches_reference():
    """The reference's ranked canary (blake3.rs:4910-4939) asserts a 437,551-byte
    proof for this input; multiproof sizes depend on every transcript byte before
    the queries, so the length fingerprints the whole transcript."""
    import json
    sys_path = os.path.join(os.path.dirname(__file__), "golden")
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk", os.path.join(sys_path, "make_canary_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    gold = json.load(open(os.path.join(sys_path, "canary_m32.json")))
    blocks = mk.canary_blocks()
    pf = O.prove(blocks)
    assert len(pf) == 437551 == gold["reference_expected_len"]
    assert O.blake3(pf).hex() == gold["proof_blake3"]
    assert O.verify(pf, blocks.size, None) == 0


def test_proof_format_parses_every_byte():
    """The field map covers an oracle proof exactly and names the first
    differing field of a tampered copy."""
    import proof_format as PF
    blocks = O.generate_compressions(8, 0x00C0FFEEBEEFD15C)
    p = O.prove(blocks)
    fs = PF.fields(p)
    assert fs[0] == ("magic", 0, 5) and sum(n for _, _, n in fs) == len(p)
    names = [n for n, _, _ in fs]
    off = dict((n, o) for n, o, _ in fs)["lincheck.z_partial"]
    q = bytearray(p)
    q[off + 3] ^= 1
    assert "lincheck.z_partial" in PF.first_difference(p, bytes(q))
    assert PF.first_difference(p, p) is None
    assert "ligerito.fold_nonces" in names
