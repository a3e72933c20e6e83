"""The C-ABI library loads and exports every symbol include/flock_gpu.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import gpu_util  # noqa: F401  (puts the package on sys.path)
import flockgpu

HEADER = os.path.join(gpu_util.ROOT, "include", "flock_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flock_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(flockgpu.LIB_PATH)
    names = declared_functions()
    assert names, "header declares no functions?"
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(flockgpu.exported_symbols()) == names


def test_ctx_create_without_gpu_fails_cleanly():
    if gpu_util.gpu_available():
        return
    h = ctypes.c_void_p()
    rc = flockgpu.lib().flock_ctx_create(0, ctypes.byref(h))
    assert rc == -3 and not h.value


def test_commit_params_sizes_and_rejections():
    """The commit's parameter check (shared by flock_commit and
    flock_commit_device, run before any allocation or launch) accepts the
    benchmark shapes with the PcsParams sizes (pcs/commit.rs:54-90) and
    rejects what the kernels cannot take."""
    P = flockgpu.PcsParams
    # config 2 (m = 29, rate 1/4) and the 2^18 proof's commit (m = 32, rate 1/2)
    assert P(29, 2, 6, 1, 1).sizes() == (1 << 22, 1 << 24, (2 << 18) - 1)
    assert P(32, 1, 6, 0, 1).sizes() == (1 << 25, 1 << 26, (2 << 20) - 1)
    for m, rate, lb, h in [(29, 0, 6, 1), (29, 5, 6, 1), (29, 31, 6, 1), (29, 2, 5, 1), (29, 2, 7, 1),
                           (29, 2, 40, 1), (13, 1, 6, 1), (41, 1, 6, 1), (4000, 1, 6, 1), (29, 2, 6, 0)]:
        try:
            P(m, rate, lb, 0, h).sizes()
        except flockgpu.FlockError as e:
            assert e.code == -1, (m, rate, lb, h)
        else:
            raise AssertionError(f"accepted unsupported params {(m, rate, lb, h)}")
    assert flockgpu.lib().flock_pcs_params_sizes(None, None, None, None) == -5
