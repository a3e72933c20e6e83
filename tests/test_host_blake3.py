"""The library's host BLAKE3 (flock-challenge_amd/csrc/host_blake3.cpp: the
compression function and the incremental hasher) against the oracle's on
random inputs: a small C++ driver (tests/native/blake3_check.cpp) built
with gcc/g++ and run on the CPU."""
import os
import shutil
import subprocess

import pytest

import gpu_util

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(gpu_util.PKG, "csrc")
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


def test_host_blake3_matches_oracle(tmp_path):
    if not shutil.which("g++") or not shutil.which("gcc"):
        pytest.skip("gcc/g++ not available")
    obj = str(tmp_path / "ob3.o")
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-I" + ORACLE, "-c", os.path.join(ORACLE, "blake3.c"),
                           "-o", obj])
    exe = str(tmp_path / "blake3_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-I" + CSRC, "-I" + ORACLE,
                           os.path.join(HERE, "native", "blake3_check.cpp"),
                           os.path.join(CSRC, "host_blake3.cpp"), obj, "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
