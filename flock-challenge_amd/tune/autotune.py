"""Tune the commit transform's pass plans on this GPU and write the plan file
the library loads at first use (tune/ntt_plans_gfx950.json beside lib/).

    python flock-challenge_amd/tune/autotune.py [--m 22-35] [--rates 1,2]
        [--reps 3] [--out PATH] [--timings PATH]

For every first-level commit shape the prover uses -- m = 22..35 variables,
rate 1/2 (the Fast and Secure ladders) or 1/4 (Slim), i.e. a transform of
2^k positions, k = m - 13 + log_inv_rate, from layer log_inv_rate -- the
library times each feasible (layer, layer count, kernel) pass and installs
the fastest partition (flock_ntt_autotune).  The plan file keeps the
partitions with their tuned and default times; --timings keeps every
measured pass as well.  Any valid plan gives the same codeword, so a stale
file costs time, never correctness.
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import flockgpu  # noqa: E402

DEFAULT_OUT = os.path.join(HERE, "ntt_plans_gfx950.json")


def parse_range(text):
    if "-" in text:
        a, b = text.split("-")
        return list(range(int(a), int(b) + 1))
    return [int(x) for x in text.split(",")]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--m", default="22-35")
    ap.add_argument("--rates", default="1,2")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--timings", default=None)
    a = ap.parse_args()
    ctx = flockgpu.Context(0)
    plans, full = [], []
    t0 = time.time()
    try:
        for rate in parse_range(a.rates):
            for m in parse_range(a.m):
                p = flockgpu.PcsParams(m=m, log_inv_rate=rate, log_batch_size=6, profile=0, merkle_hash=1)
                r = ctx.ntt_autotune(p, a.reps)
                full.append(dict(r, m=m))
                plans.append({"k": r["k"], "start": r["start"], "passes": r["passes"], "m": m,
                              "ms": round(r["ms"], 4), "default_ms": round(r["default_ms"], 4)})
                print(f"m={m} rate=1/{1 << rate} k={r['k']}: {r['ms']:.3f} ms (default {r['default_ms']:.3f}) "
                      f"passes {r['passes']}", flush=True)
    finally:
        ctx.close()
    doc = {"arch": "gfx950", "kernels": ["shared 8x2", "shared 8x1", "shared 16x1", "tail nb2", "tail nb1",
                                         "nibble-shared 16", "nibble-shared 8", "tail shift",
                                         "one-table 8x1", "one-table 4x2"],
           "reps": a.reps, "plans": plans}
    with open(a.out, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    if a.timings:
        with open(a.timings, "w") as f:
            json.dump(full, f)
    print(f"{len(plans)} plans in {time.time() - t0:.1f} s -> {a.out}", flush=True)


if __name__ == "__main__":
    main()
