"""The proof server (continuous batching, server.cpp): every result is the
same bytes alone and in the mix.  Requests of mixed sizes and both input
kinds (host compressions, device-expanded seeds) run concurrently on three
workers; each result is compared with a solo proof of the same input on a
separate context and, for the small sizes, with the oracle's proof."""
import pytest

import gpu_util
import flockgpu
import oracle_lib as O
import proof_format as PF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    c = flockgpu.Context(0)
    yield c
    c.close()


def _solo(ctx, blocks):
    s = flockgpu.Blake3Setup(ctx, blocks.size)
    try:
        return s.prove_fast(blocks)
    finally:
        s.close()


def test_mixed_requests_match_solo_proofs(ctx):
    print("server: mixed sizes on 3 workers", flush=True)
    server = flockgpu.ProofServer(0, n_workers=3)
    try:
        want = {}
        specs = [(10, 0x51), (8, 0x52), (12, 0x53), (9, 0x54), (8, 0x55), (11, 0x56)]
        for log2, seed in specs:
            rid = server.submit_seeded(log2, seed)
            want[rid] = O.generate_compressions(log2, seed)
        ragged = O.generate_compressions(9, 0x57)[:300].copy()
        want[server.submit(ragged)] = ragged
        got = {}
        while len(got) < len(want):
            r = server.poll(120000)
            assert r is not None, "server stalled"
            rid, proof, lat, queued = r
            assert lat >= queued >= 0
            got[rid] = proof
        assert server.pending() == 0
    finally:
        server.close()
    for rid, blocks in want.items():
        solo = _solo(ctx, blocks)
        diff = PF.first_difference(got[rid], solo)
        assert diff is None, f"request {rid} ({blocks.size} compressions): {diff}"
        if blocks.size <= 1024:
            assert got[rid] == O.prove(blocks)


def test_oversized_request_is_rejected():
    server = flockgpu.ProofServer(0, n_workers=1, hbm_budget_bytes=1 << 30)
    try:
        with pytest.raises(flockgpu.FlockError):
            server.submit_seeded(18, 1)  # needs ~9 GB
    finally:
        server.close()


def test_queued_requests_share_fused_fronts(ctx):
    """Requests queued together have their fronts (witness, codeword NTT,
    Merkle tree) run fused; every result still equals the solo proof."""
    print("server: 12 requests at once, 1 worker", flush=True)
    server = flockgpu.ProofServer(0, n_workers=1)
    try:
        want = {}
        for i in range(12):
            log2, seed = 8 + i % 4, 0x7700 + i
            want[server.submit_seeded(log2, seed)] = (log2, seed)
        got = {}
        while len(got) < len(want):
            r = server.poll(120000)
            assert r is not None, "server stalled"
            got[r[0]] = r[1]
        fused, launches = server.stats()
    finally:
        server.close()
    print(f"server: {fused} requests fused over {launches} front launch sets", flush=True)
    assert fused >= 2 and launches < 12
    for rid, (log2, seed) in want.items():
        s = flockgpu.Blake3Setup(ctx, 1 << log2)
        try:
            assert got[rid] == s.prove_seeded(log2, seed)
        finally:
            s.close()
