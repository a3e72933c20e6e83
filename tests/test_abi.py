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
