"""The table-free product by a short constant (flock-challenge_amd/csrc/
gf128_short.h), which the NTT's deepest layers use for twiddles below 2^64,
against the oracle's GF(2^128) product on the CPU: a small C++ driver
(tests/native/shortmul_check.cpp) built with gcc/g++ compiles the same
header the GPU kernels include."""
import os
import shutil
import subprocess

import pytest

import gpu_util

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(gpu_util.PKG, "csrc")
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


def test_short_product_matches_oracle(tmp_path):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("gcc/g++ not available")
    objs = []
    for src in ("field.c", "ntt.c"):
        obj = str(tmp_path / (src + ".o"))
        subprocess.check_call(["gcc", "-O2", "-std=c11", "-I" + ORACLE, "-c", os.path.join(ORACLE, src), "-o", obj])
        objs.append(obj)
    exe = str(tmp_path / "shortmul_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + CSRC, "-I" + ORACLE,
                           os.path.join(HERE, "native", "shortmul_check.cpp")] + objs + ["-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
