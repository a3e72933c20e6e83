"""flock_commit_device_noalloc (the commit entry point that allocates
nothing) against flock_commit_device, through the C ABI: identical root,
codeword and Merkle tree on two calls in a row, the second with a smaller
input, and its size limit reported as a shape error."""
import ctypes

import numpy as np
import pytest

import gpu_util
import flockgpu

pytestmark = pytest.mark.gpu


def _commit_both(ctx, m, rate, seed):
    params = flockgpu.PcsParams(m, rate, 6, flockgpu.PROFILE_SLIM, flockgpu.HASH_BLAKE3)
    n_msg, n_cw, n_tree = params.sizes()
    z = gpu_util.commit_message(m - 7, seed=seed)
    dz = ctx.upload(z)
    out = []
    for fn in (flockgpu.lib().flock_commit_device, flockgpu.lib().flock_commit_device_noalloc):
        dcw, dtree = ctx.alloc(n_cw * 16), ctx.alloc(n_tree * 32)
        root = (ctypes.c_uint8 * 32)()
        rc = fn(ctx.handle, ctypes.byref(params), ctypes.c_void_p(dz.ptr), ctypes.c_void_p(dcw.ptr),
                ctypes.c_void_p(dtree.ptr), root, None)
        assert rc == 0, flockgpu.lib().flock_last_error(ctx.handle)
        out.append((bytes(root), dcw.download(np.uint8), dtree.download(np.uint8)))
        dcw.free()
        dtree.free()
    dz.free()
    return out


def test_noalloc_commit_matches_commit_device_twice():
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    ctx = flockgpu.Context(0)
    try:
        # k_code 20 (the largest the fixed array takes), then 17
        for m, rate, seed in [(32, 1, 0x11A0), (29, 1, 0x11A1)]:
            print(f"no-alloc commit m={m} rate={rate}", flush=True)
            (r0, cw0, t0), (r1, cw1, t1) = _commit_both(ctx, m, rate, seed)
            assert r1 == r0
            assert t1[-32:].tobytes() == r0
            assert np.array_equal(cw1, cw0)
            assert np.array_equal(t1, t0)
        # k_code 21 exceeds the fixed twiddle array
        params = flockgpu.PcsParams(33, 1, 6, 0, 1)
        root = (ctypes.c_uint8 * 32)()
        dummy = ctx.alloc(256)
        rc = flockgpu.lib().flock_commit_device_noalloc(ctx.handle, ctypes.byref(params), ctypes.c_void_p(dummy.ptr),
                                                        ctypes.c_void_p(dummy.ptr), ctypes.c_void_p(dummy.ptr),
                                                        root, None)
        assert rc == -1
        dummy.free()
    finally:
        ctx.close()
