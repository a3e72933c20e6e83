"""Commit NTT plans on the host: the plan table, its validation and the
committed plan file (no GPU; the GPU tests run commits under these plans)."""
import json
import os

import pytest

import gpu_util

flockgpu = gpu_util.flockgpu
S82, S81, S161, T2, T1, N16, N8, SH, O81, O42 = range(10)
PLAN_FILE = os.path.join(os.path.dirname(flockgpu.__file__), "tune", "ntt_plans_gfx950.json")


def groups_log(kernel):
    return {S82: 4, S81: 3, S161: 4, N16: 4, N8: 3, O81: 3, O42: 3}.get(kernel, 0)


def feasible(k, l, F, kernel):
    """A shared-table pass needs F in {3, 4} and 2^(k-l-F) groups per block,
    at least the groups one workgroup takes; a tail pass (with nibble tables
    or shifted XORs) any F in 1..4."""
    if F < 1 or F > 4 or l + F > k:
        return False
    if kernel in (S82, S81, S161, N16, N8, O81, O42):
        return F >= 3 and k - l - F >= groups_log(kernel)
    return kernel in (T2, T1, SH)


def test_default_plans_are_passes_of_four_layers():
    for k in range(4, 26):
        for start in (1, 2):
            prev = flockgpu.ntt_plan_get(k, start)
            flockgpu.ntt_plan_set(k, start, [])
            try:
                plan = flockgpu.ntt_plan_get(k, start)
            finally:
                flockgpu.ntt_plan_set(k, start, prev)
            l = start
            for (pl, F, kn) in plan:
                assert pl == l and F == min(4, k - l) and feasible(k, pl, F, kn)
                s = k - l - F
                assert kn == (S82 if F == 4 and s >= 4 else S81 if F == 4 and s == 3 else T2)
                l += F
            assert l == k


@pytest.mark.parametrize("bad", [
    [(1, 4, S82), (5, 4, S82), (9, 3, T2)],      # layer 5: 2^3 groups per block < 16
    [(1, 4, S161), (5, 4, S161), (9, 3, T2)],
    [(1, 2, S81), (3, 4, S82), (7, 5, T2)],      # shared with F = 2; tail with F = 5
    [(1, 4, T2), (5, 4, T2)],                    # stops at layer 9 of 12
    [(1, 4, T2), (6, 4, T2), (10, 2, T2)],       # skips layer 5
    [(2, 4, T2), (6, 4, T2), (10, 2, T2)],       # does not start at layer 1
    [(1, 4, N16), (5, 4, N16), (9, 3, T2)],     # layer 5: 8 groups per block < 16
    [(1, 4, T2), (5, 4, T2), (9, 3, 8)],         # no kernel 8
    [(1, 4, SH), (5, 4, SH), (9, 4, SH)],        # shift tail with F = 4 past layer 12
])
def test_invalid_plans_are_rejected_and_leave_the_table(bad):
    before = flockgpu.ntt_plan_get(12, 1)
    with pytest.raises(flockgpu.FlockError):
        flockgpu.ntt_plan_set(12, 1, bad)
    assert flockgpu.ntt_plan_get(12, 1) == before


def test_plan_file_round_trip_and_malformed_files(tmp_path):
    good = {"plans": [{"k": 12, "start": 1, "passes": [[1, 3, S161], [4, 4, S161], [8, 4, T1]], "ms": 0.1},
                      {"start": 2, "k": 13, "passes": [[2, 1, T1], [3, 4, S82], [7, 3, S81], [10, 3, T2]]}]}
    prev = {(12, 1): flockgpu.ntt_plan_get(12, 1), (13, 2): flockgpu.ntt_plan_get(13, 2)}
    try:
        p = tmp_path / "plans.json"
        p.write_text(json.dumps(good, indent=1))
        flockgpu.ntt_plans_load(str(p))
        assert flockgpu.ntt_plan_get(12, 1) == [(1, 3, S161), (4, 4, S161), (8, 4, T1)]
        assert flockgpu.ntt_plan_get(13, 2) == [(2, 1, T1), (3, 4, S82), (7, 3, S81), (10, 3, T2)]
        # one bad plan rejects the whole file and changes nothing
        bad = json.loads(json.dumps(good))
        bad["plans"][0]["passes"] = [[1, 4, T2]]
        bad["plans"][1]["passes"] = [[2, 11, T2]]
        q = tmp_path / "bad.json"
        q.write_text(json.dumps(bad))
        with pytest.raises(flockgpu.FlockError):
            flockgpu.ntt_plans_load(str(q))
        assert flockgpu.ntt_plan_get(12, 1) == [(1, 3, S161), (4, 4, S161), (8, 4, T1)]
        for text in ["", "{}", '{"plans": [{"k": 12}]}', '{"plans": [{"k": 12, "start": 1, "passes": [[1, 4]]}]}']:
            q.write_text(text)
            with pytest.raises(flockgpu.FlockError):
                flockgpu.ntt_plans_load(str(q))
        with pytest.raises(flockgpu.FlockError):
            flockgpu.ntt_plans_load(str(tmp_path / "missing.json"))
    finally:
        for (k, s), v in prev.items():
            flockgpu.ntt_plan_set(k, s, v)


@pytest.mark.skipif(not os.path.exists(PLAN_FILE), reason="no tuned plan file")
def test_committed_plan_file_is_valid_and_in_force():
    doc = json.load(open(PLAN_FILE))
    assert doc["arch"] == "gfx950" and doc["plans"]
    flockgpu.ntt_plans_load(PLAN_FILE)
    for p in doc["plans"]:
        k, start = p["k"], p["start"]
        assert k == p["m"] - 13 + start
        passes = [tuple(x) for x in p["passes"]]
        l = start
        for (pl, F, kn) in passes:
            assert pl == l and feasible(k, pl, F, kn), (k, start, passes)
            l += F
        assert l == k
        assert flockgpu.ntt_plan_get(k, start) == passes
        assert p["ms"] <= p["default_ms"] + 1e-9
