"""The batched and served workloads on the GPU, and one proof of every size
the ABI accepts that the other GPU tests leave out.

Batched: 64 instances with log2 = 10 + draw mod 9 and sub-seeds from one
SplitMix64 stream (seed 0xBA7C000000000040), plus one ragged instance of 3000
compressions (padding), all submitted at once to the proof server.  Served:
200 open-loop Poisson requests, log2 uniform in 12..18, seeded.  For both,
every result must equal a solo proof of the same input on a separate
context, pass the library's verifier with its commitment root recomputed on
the GPU, and, for at most 2^12 compressions, equal the oracle's proof.

Every size: one proof at each log2 in 13..17 and 19..21 (m = 27..31,
33..35, covering the 5- and 7-level Ligerito ladders), compared with a proof
of the same input on a fresh context and checked by the verifier; log2 13..16
also against the oracle's proof."""
import random

import numpy as np
import pytest

import gpu_util
import flockgpu
import oracle_lib as O
import proof_format as PF

pytestmark = pytest.mark.gpu

BATCH_SEED = 0xBA7C000000000040
SERVE_SEED = 0x5E7E000000000041


def _say(msg):
    print(msg, flush=True)


def splitmix64(seed):
    """SplitMix64 stream (benchmark-tools/common/src/lib.rs:4-18)."""
    state = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        state = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


@pytest.fixture(scope="module")
def solo_ctx():
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    c = flockgpu.Context(0)
    yield c
    c.close()


class Solo:
    """Solo proofs and verification on one context, with a setup per size."""

    def __init__(self, ctx):
        self.ctx, self.setups = ctx, {}

    def setup(self, n):
        if n not in self.setups:
            self.setups[n] = flockgpu.Blake3Setup(self.ctx, n)
        return self.setups[n]

    def close(self):
        for s in self.setups.values():
            s.close()


def _serve(requests, n_workers=3):
    """Run (t_offset_s, kind, arg) requests through a proof server, each
    submitted at its offset; kind "seeded" takes (log2, seed), "blocks" a
    compression array.  Returns {index: proof bytes}."""
    server = flockgpu.ProofServer(0, n_workers=n_workers, reserve_log2=18)
    out, ids = {}, {}
    try:
        import time
        t0 = time.perf_counter()
        i = 0
        while len(out) < len(requests):
            now = time.perf_counter() - t0
            while i < len(requests) and requests[i][0] <= now:
                _, kind, arg = requests[i]
                rid = server.submit_seeded(*arg) if kind == "seeded" else server.submit(arg)
                ids[rid] = i
                i += 1
            r = server.poll(2 if i < len(requests) else 120000)
            if r is None:
                assert i < len(requests), "server stalled"
                continue
            rid, proof, lat, queued = r
            assert lat >= queued >= 0
            out[ids[rid]] = proof
        assert server.pending() == 0
    finally:
        server.close()
    return out


def _check_all(solo_ctx, requests, got, oracle_max_log2=12):
    """Each result equals a solo proof on another context and passes the
    verifier with the GPU-recomputed root; small ones equal the oracle's."""
    S = Solo(solo_ctx)
    n_oracle = 0
    try:
        for idx, (_, kind, arg) in enumerate(requests):
            if kind == "seeded":
                log2, seed = arg
                n = 1 << log2
                s = S.setup(n)
                solo = s.prove_seeded(log2, seed)
                what = f"request {idx}: 2^{log2} seed {seed:#x}"
            else:
                n = arg.size
                s = S.setup(n)
                solo = s.prove_fast(arg)
                what = f"request {idx}: {n} host compressions"
            diff = PF.first_difference(got[idx], solo)
            assert diff is None, f"{what}: served/batched proof differs from the solo proof: {diff}"
            if kind == "seeded":
                s.verify_seeded(log2, seed, got[idx])
            else:
                s.verify(got[idx], arg)
            if n <= (1 << oracle_max_log2):
                blocks = O.generate_compressions(log2, seed) if kind == "seeded" else arg
                assert got[idx] == O.prove(blocks), f"{what}: differs from the oracle's proof"
                n_oracle += 1
    finally:
        S.close()
    return n_oracle


def batched_requests():
    g = splitmix64(BATCH_SEED)
    reqs = []
    for _ in range(64):
        log2 = 10 + next(g) % 9
        reqs.append((0.0, "seeded", (log2, next(g))))
    ragged = O.generate_compressions(12, 0xBA7C0000000000FF)[:3000].copy()
    reqs.append((0.0, "blocks", ragged))
    return reqs


def served_requests(n=200, rate=150.0):
    rng = random.Random(SERVE_SEED)
    g = splitmix64(SERVE_SEED)
    t, reqs = 0.0, []
    for _ in range(n):
        t += rng.expovariate(rate)
        reqs.append((t, "seeded", (12 + next(g) % 7, next(g))))
    return reqs


def test_batched_64_instances_match_solo_verifier_and_oracle(solo_ctx):
    reqs = batched_requests()
    sizes = sorted({r[2][0] for r in reqs if r[1] == "seeded"})
    _say(f"batched: {len(reqs)} instances, log2 in {sizes} + 3000 ragged, 3 workers")
    got = _serve(reqs)
    _say("batched: checking against solo proofs, verifier, oracle")
    n_oracle = _check_all(solo_ctx, reqs, got)
    assert n_oracle >= 10
    _say(f"batched: {len(reqs)} proofs checked, {n_oracle} against the oracle")


def test_served_poisson_200_requests_match_solo_verifier_and_oracle(solo_ctx):
    reqs = served_requests()
    _say(f"served: {len(reqs)} Poisson requests over {reqs[-1][0]:.2f} s, log2 12..18, 3 workers")
    got = _serve(reqs)
    _say("served: checking against solo proofs, verifier, oracle")
    n_oracle = _check_all(solo_ctx, reqs, got)
    assert n_oracle >= 10
    _say(f"served: {len(reqs)} proofs checked, {n_oracle} against the oracle")


@pytest.mark.parametrize("log2", [13, 14, 15, 16, 17, 19, 20, 21])
def test_every_size_matches_fresh_context_and_verifies(solo_ctx, log2):
    seed = 0x51E0000000000000 | log2
    _say(f"size: 2^{log2} (m = {14 + log2}) seed {seed:#x}")
    s = flockgpu.Blake3Setup(solo_ctx, 1 << log2)
    try:
        got = s.prove_seeded(log2, seed)
        assert len(got) < 500_000
        fresh = flockgpu.Context(0)
        try:
            fs = flockgpu.Blake3Setup(fresh, 1 << log2)
            try:
                d = flockgpu.generate_compressions_device(fresh, log2, seed)
                other = fs.prove_fast_device(d.ptr, 1 << log2)
                d.free()
            finally:
                fs.close()
        finally:
            fresh.close()
        diff = PF.first_difference(got, other)
        assert diff is None, f"2^{log2}: proofs on two contexts differ: {diff}"
        s.verify_seeded(log2, seed, got)
        # a proof of one input does not verify against another's commitment
        with pytest.raises(flockgpu.VerifyError):
            s.verify_seeded(log2, seed ^ 1, got)
    finally:
        s.close()
    if log2 <= 16:
        _say(f"size: 2^{log2} against the oracle")
        want = O.prove(O.generate_compressions(log2, seed))
        diff = PF.first_difference(got, want)
        assert diff is None, f"2^{log2}: differs from the oracle's proof: {diff}"


def _batch_items(ctx, salt, keep):
    """The batched workload as flock_prove_batch items (64 seeded + one
    ragged device item of 3000 compressions); the device buffer goes into
    `keep` so it outlives the call."""
    g = splitmix64(BATCH_SEED)
    items, reqs = [], []
    for _ in range(64):
        log2 = 10 + next(g) % 9
        seed = next(g) ^ salt
        items.append(("seeded", log2, seed))
        reqs.append((0.0, "seeded", (log2, seed)))
    ragged = O.generate_compressions(12, 0xBA7C0000000000FF ^ salt)[:3000].copy()
    d = ctx.upload(ragged)
    keep.append(d)
    items.append(("device", d.ptr, 3000))
    reqs.append((0.0, "blocks", ragged))
    return items, reqs


def test_prove_batch_fused_fronts_match_solo_and_replay_the_graph(solo_ctx):
    ctx = flockgpu.Context(0)
    keep = []
    try:
        items, reqs = _batch_items(ctx, 0, keep)
        _say(f"prove_batch: {len(items)} items, fronts fused, graph captured")
        proofs, done, front_s, replayed = flockgpu.prove_batch(ctx, items, n_workers=3)
        assert not replayed and len(proofs) == len(items)
        _say(f"prove_batch: fronts {1000 * front_s:.2f} ms, last proof at {1000 * max(done):.1f} ms")
        n_oracle = _check_all(solo_ctx, reqs, dict(enumerate(proofs)))
        assert n_oracle >= 10
        # the same shape with other inputs replays the captured graph
        items2, reqs2 = _batch_items(ctx, 0x5A17, keep)
        assert [it[1] for it in items2[:64]] == [it[1] for it in items[:64]]
        _say("prove_batch: same shape, new seeds, graph replayed")
        proofs2, _, _, replayed2 = flockgpu.prove_batch(ctx, items2, n_workers=3)
        assert replayed2
        _check_all(solo_ctx, reqs2, dict(enumerate(proofs2)), oracle_max_log2=0)
        # and without the graph, on one worker, the same bytes
        _say("prove_batch: no graph, one worker")
        proofs3, _, _, replayed3 = flockgpu.prove_batch(ctx, items2[:5] + items2[-1:], n_workers=1, use_graph=False)
        assert not replayed3 and proofs3 == proofs2[:5] + proofs2[-1:]
    finally:
        for d in keep:
            d.free()
        ctx.close()
