"""Pin the oracle's additive NTT, Merkle tree and PCS commit."""
import random

import numpy as np
import pytest

import oracle_lib as O
import pyref


def _W_hat(log_d):
    """Normalized subspace polynomials of the standard basis {x^i}, evaluated
    directly from their definition (no twiddle tables)."""
    basis = [1 << i for i in range(log_d)]

    def W(i, z):
        if i == 0:
            return z
        w = W(i - 1, z)
        return pyref.gf_mul(w, w ^ W(i - 1, basis[i - 1]))

    norms = [pyref.gf_inv(W(i, basis[i])) for i in range(log_d)]
    return lambda i, z: pyref.gf_mul(W(i, z), norms[i])


@pytest.mark.parametrize("log_d", [1, 2, 3, 4])
def test_ntt_is_novel_basis_evaluation(log_d):
    """Forward NTT of coefficients c gives P(w_k) with P = sum_j c_j X_j,
    X_j = prod_{bit i of j} W^_i, and w_k = sum_i bit_i(k) x^i."""
    rng = random.Random(log_d)
    n = 1 << log_d
    coeffs = [rng.getrandbits(128) for _ in range(n)]
    Wh = _W_hat(log_d)
    want = []
    for k in range(n):
        w = k  # span element: bit i of k times x^i
        acc = 0
        for j in range(n):
            x = 1
            for i in range(log_d):
                if (j >> i) & 1:
                    x = pyref.gf_mul(x, Wh(i, w))
            acc ^= pyref.gf_mul(coeffs[j], x)
        want.append(acc)
    t = O.NttHandle(log_d)
    data = O.ints_to_f128_array(coeffs)
    t.forward_interleaved(data, 1, 0)
    assert O.f128_array_to_ints(data) == want


def test_twiddle_structure():
    t = O.NttHandle(6)
    assert t.twiddle(0, 0) == 0
    # deepest layer: twiddle(l-1, b) = span of x^1.. = b * x
    for b in range(32):
        assert t.twiddle(5, b) == b << 1
    # linearity in the block index
    for layer in range(6):
        for b1 in range(1 << layer):
            for b2 in range(1 << layer):
                assert t.twiddle(layer, b1 ^ b2) == t.twiddle(layer, b1) ^ t.twiddle(layer, b2)


def test_ntt_roundtrip_linearity_zero():
    rng = np.random.default_rng(3)
    for log_d in [1, 3, 6, 9]:
        t = O.NttHandle(log_d)
        n = 1 << log_d
        a = rng.integers(0, 2**63, size=2 * n, dtype=np.uint64) * 2 + 1
        b = rng.integers(0, 2**63, size=2 * n, dtype=np.uint64)
        fa, fb, fab = a.copy(), b.copy(), a ^ b
        t.forward_interleaved(fa, 1, 0)
        t.forward_interleaved(fb, 1, 0)
        t.forward_interleaved(fab, 1, 0)
        assert np.array_equal(fab, fa ^ fb)
        back = fa.copy()
        t.inverse(back)
        assert np.array_equal(back, a)
        z = np.zeros(2 * n, np.uint64)
        t.forward_interleaved(z, 1, 0)
        assert not z.any()


def test_interleaved_equals_per_lane_and_from_layer():
    rng = np.random.default_rng(4)
    log_d, lanes = 7, 8
    t = O.NttHandle(log_d)
    data = rng.integers(0, 2**64, size=(1 << log_d) * lanes * 2, dtype=np.uint64)
    inter = data.copy()
    t.forward_interleaved(inter, lanes, 0)
    v = data.reshape(1 << log_d, lanes, 2)
    for lane in range(lanes):
        one = np.ascontiguousarray(v[:, lane, :]).reshape(-1)
        t.forward_interleaved(one, 1, 0)
        assert np.array_equal(one.reshape(-1, 2), inter.reshape(1 << log_d, lanes, 2)[:, lane, :])
    # replicate-then-start-at-layer-r equals the full transform of [msg, 0, ...]
    msg = rng.integers(0, 2**64, size=(1 << (log_d - 2)) * lanes * 2, dtype=np.uint64)
    full = np.concatenate([msg, np.zeros(3 * msg.size, np.uint64)])
    t.forward_interleaved(full, lanes, 0)
    rep = np.concatenate([msg] * 4)
    t.forward_interleaved(rep, lanes, 2)
    assert np.array_equal(full, rep)


def _py_tree(data, n):
    leaf = len(data) // n
    nodes = [pyref.blake3_non_root_cv(data[i * leaf:(i + 1) * leaf]) for i in range(n)]
    tree = list(nodes)
    level = nodes
    while len(level) > 1:
        level = [pyref.blake3_parent_cv(level[2 * i], level[2 * i + 1]) for i in range(len(level) // 2)]
        tree += level
    return b"".join(tree)


@pytest.mark.parametrize("n,leaf", [(1, 64), (2, 128), (8, 1024), (16, 96)])
def test_merkle_tree_matches_independent(n, leaf):
    data = np.frombuffer(bytes((7 * i + 3) % 256 for i in range(n * leaf)), np.uint8).copy()
    assert O.merkle_tree(data, n).tobytes() == _py_tree(data.tobytes(), n)


def test_multiproof_roundtrip_and_tamper():
    n = 64
    data = np.arange(n * 128, dtype=np.uint32).astype(np.uint8)
    tree = O.merkle_tree(data, n)
    root = tree[-32:].copy()
    pos = np.array([3, 4, 5, 17, 17, 63, 0], dtype=np.uint64)
    cap = 6 * 7
    proof = np.zeros(cap * 32, np.uint8)
    k = O.lib().omerkle_multi_proof(O.ptr(tree), n, O.ptr(pos), pos.size, O.ptr(proof), cap)
    su = np.unique(pos)
    leaves = np.concatenate([tree[32 * p:32 * p + 32] for p in su])
    assert O.lib().omerkle_verify_multi(O.ptr(root), n, O.ptr(su), O.ptr(leaves), su.size, O.ptr(proof), k) == 1
    # every sibling counted once: 4/5 share parent
    bad = proof.copy()
    bad[5] ^= 1
    assert O.lib().omerkle_verify_multi(O.ptr(root), n, O.ptr(su), O.ptr(leaves), su.size, O.ptr(bad), k) == 0
    assert O.lib().omerkle_verify_multi(O.ptr(root), n, O.ptr(su), O.ptr(leaves), su.size, O.ptr(proof), k - 1) == 0


def test_commit_composition_small():
    """Tiny commit restated independently: replicate, full NTT per lane from
    layer 0 on [z, 0], then BLAKE3 tree over positions."""
    m, rate, lb = 15, 1, 6  # 2^8 F128 = 4 positions x 64 lanes -> 8 positions
    rng = np.random.default_rng(5)
    z = rng.integers(0, 2**64, size=(1 << (m - 7)) * 2, dtype=np.uint64)
    root, cw, tree = O.pcs_commit(m, rate, lb, z, want_codeword=True)
    npos = 1 << (m - 7 - lb + rate)
    t = O.NttHandle(npos.bit_length() - 1)
    full = np.concatenate([z, np.zeros(z.size, np.uint64)])
    t.forward_interleaved(full, 64, 0)
    assert np.array_equal(full, cw)
    assert tree.tobytes() == _py_tree(cw.tobytes(), npos)
    assert root == tree[-32:].tobytes()
