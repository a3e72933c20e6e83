#!/usr/bin/env python3
"""Benchmark of the MI355X Flock prover hot path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload commit]

Prints ONE JSON line.  Workloads:
  commit  isolated PCS commit: 2^22 F128 (seeded SplitMix64) -> rate-1/4
          interleaved additive NTT to 2^24 F128 -> BLAKE3 Merkle over 2^18
          1 KiB leaves.  One step = one commit, inputs resident in HBM.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "flock-challenge_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def bench_commit(args):
    import numpy as np

    import flockgpu
    import gpu_util

    m, rate, lb = 29, 2, 6
    params = flockgpu.PcsParams(m, rate, lb, flockgpu.PROFILE_SLIM, flockgpu.HASH_BLAKE3)
    z = gpu_util.commit_message(22)
    ctx = flockgpu.Context(0)
    dz = ctx.upload(z)
    dcw = ctx.alloc(params.codeword_len_f128() * 16)
    dtree = ctx.alloc((2 * params.n_positions() - 1) * 32)
    for _ in range(args.warmup):
        root = ctx.commit_device(params, dz.ptr, dcw.ptr, dtree.ptr)
    ctx.synchronize()
    ctx.timing(True)
    wall0 = time.perf_counter()
    for _ in range(args.steps):
        root = ctx.commit_device(params, dz.ptr, dcw.ptr, dtree.ptr)
    ctx.synchronize()
    wall = time.perf_counter() - wall0
    rep = ctx.timing_report()
    ctx.timing(False)
    # HIP events on the launching stream bracket every commit ("commit" phase).
    ms_per_step = rep["commit"][1] / rep["commit"][0]

    # Dominant kernel: one NTT pass reads and writes the whole codeword once.
    cw_bytes = params.codeword_len_f128() * 16
    n_launch, ntt_ms = rep.get("ntt_pass_kernel", (0, 0.0))
    avg_ntt_ms = ntt_ms / max(1, n_launch)
    achieved = 2 * cw_bytes / (avg_ntt_ms * 1e-3) / 1e9

    out = {
        "metric": "isolated PCS commit throughput (config 2: 2^22 F128 -> rate-1/4 NTT -> BLAKE3 Merkle)",
        "value": 1000.0 / ms_per_step,
        "unit": "commits/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "gf2_128 (u32 lanes)",
        "data": "synthetic (SplitMix64 seed 0xCA5E10CAC0111701)",
        "config": {"workload": "commit m=29 log_inv_rate=2 log_batch_size=6 blake3", "root": root.hex(),
                   "phases_ms_per_step": {k: v[1] / args.steps for k, v in rep.items()},
                   "wall_s": wall},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "ntt_pass_kernel", "avg_launch_ms": avg_ntt_ms,
                     "algorithmic_bytes_per_launch": 2 * cw_bytes},
    }
    if not args.no_cpu_baseline:
        import oracle_lib as O
        t = time.perf_counter()
        croot, _, _ = O.pcs_commit(m, rate, lb, z)
        dt = time.perf_counter() - t
        assert croot == root, "GPU root differs from the oracle"
        out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "commits/s", "cores": 1, "kind": "port",
                               "sample": "one full config-2 commit by the C oracle (single thread)"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="commit", choices=["commit"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; --gpus must be 1"}))
        return 2
    out = bench_commit(args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
