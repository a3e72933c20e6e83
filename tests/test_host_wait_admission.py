"""Two pieces of the library's host logic on the CPU (tests/native/
wait_admission_check.cpp, built with g++): the bounded waits every host wait
for the device goes through (a flag never set ends at the deadline, a set
flag ends the wait, a failing probe ends it at once), and the proof server's
admission by HBM footprint with uneven reservations (the head request that
fits no worker as reserved makes idle workers free their buffers instead of
stalling the queue)."""
import os
import shutil
import subprocess

import pytest

import gpu_util

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(gpu_util.PKG, "csrc")


def test_bounded_waits_and_admission(tmp_path):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    exe = str(tmp_path / "wait_admission_check")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-pthread", "-I" + CSRC,
                           os.path.join(HERE, "native", "wait_admission_check.cpp"), "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
