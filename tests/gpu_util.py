"""Shared helpers for the GPU tests: the package import and seeded inputs."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "flock-challenge_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import flockgpu  # noqa: E402

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64_stream(seed, n):
    """n consecutive SplitMix64 outputs (full 64-bit) from `seed`, vectorized."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def commit_message(log_len, seed=0xCA5E10CAC0111701):
    """2^log_len F128 {lo = next_u64, hi = next_u64} as 2*2^log_len uint64 words."""
    return splitmix64_stream(seed, 2 << log_len)


def gpu_available():
    """Through the library's own HIP runtime (never torch's: a second HIP
    runtime in the process would contend for the device)."""
    try:
        return flockgpu.device_count() > 0
    except Exception:
        return False
