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


# ---- multiproof known answers: the reference's own structural tests of
# merkle_multi_proof (crates/flock-core/src/merkle.rs:1461-1550), restated
# against the oracle, plus a set-level definition of the emission order.

def _oracle_multiproof(tree, n, positions):
    pos = np.array(positions, dtype=np.uint64)
    cap = max(1, len(positions)) * max(1, n.bit_length())
    out = np.zeros(cap * 32, np.uint8)
    k = O.lib().omerkle_multi_proof(O.ptr(tree), n, O.ptr(pos), pos.size, O.ptr(out), cap)
    return [out[32 * i:32 * i + 32].tobytes() for i in range(k)]


def _node(tree, idx):
    return tree[32 * idx:32 * idx + 32].tobytes()


def _single_path(tree, n, i):
    """The single opening path (merkle_proof, merkle.rs:819-838): the sibling
    at every level, leaves first, in the flat layout (level after level)."""
    out, start, length = [], 0, n
    while length > 1:
        out.append(_node(tree, start + (i ^ 1)))
        start += length
        length >>= 1
        i >>= 1
    return out


def _set_order(tree, n, positions):
    """The multiproof's hashes defined over sets: at each level from the leaves
    up, the siblings of the active nodes that are not active themselves, in
    ascending position order; the parents become the next active set."""
    act, out, start, length = set(positions), [], 0, n
    while length > 1:
        need = sorted({p ^ 1 for p in act} - act)
        out += [_node(tree, start + p) for p in need]
        act = {p >> 1 for p in act}
        start += length
        length >>= 1
    return out


def _tree(n, leaf, seed):
    rng = np.random.default_rng(seed)
    return O.merkle_tree(rng.integers(0, 256, n * leaf, dtype=np.uint8), n)


def test_multiproof_single_position_equals_single_path():
    n = 16
    tree = _tree(n, 8, 42)
    for i in range(n):
        assert _oracle_multiproof(tree, n, [i]) == _single_path(tree, n, i)


def test_multiproof_sibling_pair_emits_no_leaf_level_hash():
    n = 8
    tree = _tree(n, 4, 7)
    got = _oracle_multiproof(tree, n, [0, 1])
    # one sibling per level above the leaves: node 1 of level 1, node 1 of level 2
    assert got == [_node(tree, 8 + 1), _node(tree, 12 + 1)]


def test_multiproof_full_query_set_is_empty():
    n = 16
    tree = _tree(n, 8, 99)
    assert _oracle_multiproof(tree, n, list(range(n))) == []


def test_multiproof_random_subsets_follow_the_set_order():
    rng = np.random.default_rng(2024)
    for n in (2, 8, 64, 256):
        tree = _tree(n, 16, n)
        for _ in range(20):
            k = int(rng.integers(1, n + 1))
            pos = sorted(set(int(p) for p in rng.integers(0, n, k)))
            got = _oracle_multiproof(tree, n, pos)
            assert got == _set_order(tree, n, pos), (n, pos)
            # unsorted input with repeats gives the same hashes
            shuffled = list(pos) * 2
            rng.shuffle(shuffled)
            assert _oracle_multiproof(tree, n, shuffled) == got
