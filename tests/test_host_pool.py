"""The prover's host thread pool (flock-challenge_amd/csrc/host_pool.cpp),
used for the lincheck rounds: a small C++ driver (tests/native/
host_pool_check.cpp) built with g++ and run on the CPU."""
import os
import shutil
import subprocess

import pytest

import gpu_util

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(gpu_util.PKG, "csrc")


def test_host_pool_runs_every_chunk_once(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path / "host_pool_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-pthread", "-I" + CSRC,
                           os.path.join(HERE, "native", "host_pool_check.cpp"),
                           os.path.join(CSRC, "host_pool.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
