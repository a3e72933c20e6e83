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
