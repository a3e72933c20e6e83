"""GPU PCS commit (NTT + BLAKE3 Merkle) against the oracle, bit-exact."""
import numpy as np
import pytest

import gpu_util
import oracle_lib as O

flockgpu = gpu_util.flockgpu
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = flockgpu.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("m,rate", [(15, 1), (16, 1), (18, 2), (20, 1), (22, 1), (24, 3)])
def test_commit_root_matches_oracle(ctx, m, rate):
    print(f"commit parity m={m} rate={rate}", flush=True)
    z = gpu_util.commit_message(m - 7, seed=0x1234 + m)
    want, _, _ = O.pcs_commit(m, rate, 6, z)
    p = flockgpu.PcsParams(m, rate, 6, 0, 1)
    assert ctx.commit(p, z) == want


def test_commit_codeword_and_tree_match_oracle(ctx):
    m, rate = 19, 1
    z = gpu_util.commit_message(m - 7, seed=77)
    want_root, want_cw, want_tree = O.pcs_commit(m, rate, 6, z, want_codeword=True)
    p = flockgpu.PcsParams(m, rate, 6, 0, 1)
    dz = ctx.upload(z)
    dcw = ctx.alloc(p.codeword_len_f128() * 16)
    dtree = ctx.alloc((2 * p.n_positions() - 1) * 32)
    root = ctx.commit_device(p, dz.ptr, dcw.ptr, dtree.ptr)
    assert root == want_root
    assert np.array_equal(dcw.download(np.uint64), want_cw)
    assert dtree.download().tobytes() == want_tree.tobytes()


def test_commit_config2_shape(ctx):
    """Isolated commit: 2^22 F128 message, rate 1/4, 64 lanes -> 2^18 x 1 KiB leaves."""
    m, rate = 29, 2
    z = gpu_util.commit_message(22)
    print("config-2 oracle commit ...", flush=True)
    want, _, _ = O.pcs_commit(m, rate, 6, z)
    p = flockgpu.PcsParams(m, rate, 6, 1, 1)
    assert ctx.commit(p, z) == want


def test_commit_rejects_bad_shapes(ctx):
    p = flockgpu.PcsParams(20, 1, 6, 0, 1)
    with pytest.raises(flockgpu.FlockError):
        ctx.commit(p, np.zeros(10, np.uint64))
    with pytest.raises(flockgpu.FlockError):
        ctx.commit(flockgpu.PcsParams(20, 1, 6, 0, 0), np.zeros(2 << 13, np.uint64))


# Plans for the transform of m = 24, rate 1/2 (k = 12 from layer 1) that
# exercise every kernel and layer count: (l, F, kernel) passes.
S82, S81, S161, T2, T1, N16, N8, SH, O81, O42 = range(10)
PLANS_K12 = {
    "default": None,
    "16-wave shared, one-table tail": [(1, 3, S161), (4, 4, S161), (8, 4, T1)],
    "8x1 and 8x2 shared, mixed tails": [(1, 4, S81), (5, 3, S82), (8, 1, T1), (9, 2, T2), (11, 1, T1)],
    "tail passes only, first reads the message": [(1, 2, T2), (3, 4, T1), (7, 4, T2), (11, 1, T2)],
    "one layer per pass": [(l, 1, T1 if l % 2 else T2) for l in range(1, 12)],
    "nibble-shared 16 and 8": [(1, 4, N16), (5, 4, N8), (9, 3, T1)],
    "nibble-shared, three layers": [(1, 3, N8), (4, 3, N16), (7, 3, T2), (10, 2, T1)],
    # shifted-XOR tails at every layer count, on the short deepest twiddles
    # and on long ones (first pass from the message, zero block-0 twiddles)
    "shift tails, deepest layers": [(1, 3, N8), (4, 4, S82), (8, 4, SH)],
    "shift tails everywhere": [(1, 4, SH), (5, 3, SH), (8, 2, SH), (10, 1, SH), (11, 1, T1)],
    # one-table passes (two workgroups per CU), from the message and on the
    # codeword, zero block-0 twiddles included
    "one-table 8x1 and 4x2": [(1, 4, O81), (5, 4, O42), (9, 3, T1)],
    "one-table, three layers": [(1, 3, O42), (4, 3, O81), (7, 2, T2), (9, 3, SH)],
}


@pytest.mark.parametrize("name", list(PLANS_K12))
def test_every_plan_gives_the_oracle_root(ctx, name):
    m, rate, k = 24, 1, 12
    print(f"commit plan '{name}' m={m}", flush=True)
    z = gpu_util.commit_message(m - 7, seed=0x9147)
    want, _, _ = O.pcs_commit(m, rate, 6, z)
    prev = flockgpu.ntt_plan_get(k, rate)
    try:
        if PLANS_K12[name] is not None:
            flockgpu.ntt_plan_set(k, rate, PLANS_K12[name], ctx)
            assert flockgpu.ntt_plan_get(k, rate) == PLANS_K12[name]
        assert ctx.commit(flockgpu.PcsParams(m, rate, 6, 0, 1), z) == want
    finally:
        flockgpu.ntt_plan_set(k, rate, prev)


@pytest.mark.parametrize("m,rate", [(24, 1), (25, 2)])
def test_autotuned_plan_is_installed_and_gives_the_oracle_root(ctx, m, rate):
    k = m - 13 + rate
    print(f"ntt autotune m={m} rate={rate}", flush=True)
    z = gpu_util.commit_message(m - 7, seed=0x7111 + m)
    want, _, _ = O.pcs_commit(m, rate, 6, z)
    prev = flockgpu.ntt_plan_get(k, rate)
    try:
        r = ctx.ntt_autotune(flockgpu.PcsParams(m, rate, 6, 0, 1), reps=2)
        assert (r["k"], r["start"]) == (k, rate)
        passes = [tuple(p) for p in r["passes"]]
        assert flockgpu.ntt_plan_get(k, rate) == passes
        # the partition covers the layers in order and is no slower than the
        # default's passes as measured in the same run
        assert passes[0][0] == rate and sum(p[1] for p in passes) == k - rate
        assert r["ms"] <= r["default_ms"] + 1e-9
        measured = {(t[0], t[1], t[2]) for t in r["timings"]}
        assert all(p in measured for p in passes)
        assert ctx.commit(flockgpu.PcsParams(m, rate, 6, 0, 1), z) == want
    finally:
        flockgpu.ntt_plan_set(k, rate, prev)
