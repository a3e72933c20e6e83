"""The benchmark-protocol programs (flock_harness, flock_worker,
flock_verify_proof; benchmark-tools/harness/src/main.rs and
benchmark-tools/worker/src/main.rs).

CPU tests drive the real flock_harness with stand-in workers and a stand-in
verifier (small Python programs written into a temporary copy of the
harness's directory), covering the protocol itself: readiness, the seed line,
publication by rename, the capture checks (size bounds, no symlinks, regular
file), header rejection, the percentile arithmetic (the reference's own test
values, main.rs:477-482) and the score files.  The GPU test runs the real
worker and verifier: the worker's proof equals the library's for the same
seed, and a short harness run verifies every trial."""
import json
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

import gpu_util

LIB = os.path.join(gpu_util.PKG, "lib")
HARNESS = os.path.join(LIB, "flock_harness")
WORKER = os.path.join(LIB, "flock_worker")
VERIFIER = os.path.join(LIB, "flock_verify_proof")


def _script(path, body):
    with open(path, "w") as f:
        f.write("#!" + sys.executable + "\n" + textwrap.dedent(body))
    os.chmod(path, 0o755)
    return path


# A stand-in worker: readiness, one seed line, publication by rename.  PAYLOAD
# selects what it publishes; it records the seed it read in SEEDS_FILE.
FAKE_WORKER = """
import os, sys
log2, ready, proof = sys.argv[1], sys.argv[2], sys.argv[3]
mode = os.environ["FAKE_MODE"]
if mode == "die":
    sys.exit(3)
open(ready, "w").write("ready\\n")
seed = sys.stdin.readline().strip()
with open(os.environ["SEEDS_FILE"], "a") as f:
    f.write(seed + " " + log2 + " " + os.environ.get("RAYON_NUM_THREADS", "") + "\\n")
if mode == "silent":
    sys.exit(0)
tmp = proof + ".tmp"
if mode == "symlink":
    target = os.path.join(os.path.dirname(proof), "elsewhere")
    open(target, "wb").write(b"FLOCK\\x04\\x02" + b"x" * 100)
    os.symlink(target, proof)
    sys.exit(0)
data = {"ok": b"FLOCK\\x04\\x02" + seed.encode() * 10,
        "header": b"FLOCX\\x04\\x02" + b"x" * 100,
        "big": b"FLOCK\\x04\\x02" + b"x" * 500000,
        "empty": b""}[mode]
open(tmp, "wb").write(data)
os.rename(tmp, proof)
"""

# A stand-in verifier: accepts when the captured copy holds the seed it was
# given (so the harness must pass the right seed and bytes).
FAKE_VERIFIER = """
import sys
log2, seed, path = sys.argv[1], sys.argv[2], sys.argv[3]
data = open(path, "rb").read()
sys.exit(0 if data == b"FLOCK\\x04\\x02" + seed.encode() * 10 else 2)
"""


@pytest.fixture()
def rig(tmp_path):
    if not os.path.exists(HARNESS):
        pytest.skip("flock_harness not built")
    bindir = tmp_path / "bin"
    bindir.mkdir()
    shutil.copy(HARNESS, bindir / "flock_harness")
    _script(str(bindir / "flock_verify_proof"), FAKE_VERIFIER)
    worker = _script(str(bindir / "worker"), FAKE_WORKER)
    return {"bin": bindir, "worker": worker, "tmp": tmp_path}


def _run(rig, mode, log2=12, threads=3, warm=1, runs=3):
    env = dict(os.environ, FAKE_MODE=mode, SEEDS_FILE=str(rig["tmp"] / "seeds.txt"))
    args = [str(rig["bin"] / "flock_harness"), rig["worker"], str(rig["tmp"] / "scratch"),
            str(rig["tmp"] / "score.json"), str(rig["tmp"] / "summary.md"), str(log2), str(threads), str(warm),
            str(runs)]
    return subprocess.run(args, env=env, capture_output=True, text=True, timeout=120)


def test_full_protocol_scores_and_writes_results(rig):
    r = _run(rig, "ok", log2=12, threads=3, warm=1, runs=3)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert lines[0].startswith("warmup_trial=1/1 ") and "verified=true included_in_score=false" in lines[0]
    assert [ln.split()[0] for ln in lines[1:4]] == ["measured_trial=1/3", "measured_trial=2/3", "measured_trial=3/3"]
    assert lines[-1].startswith("score=") and lines[-1].endswith(" compressions_per_second")
    seeds = open(rig["tmp"] / "seeds.txt").read().split("\n")[:-1]
    assert len(seeds) == 4
    for ln in seeds:
        seed, log2, threads = ln.split()
        assert 0 <= int(seed) < 2 ** 64 and log2 == "12" and threads == "3"
    assert len({ln.split()[0] for ln in seeds}) == 4  # fresh seed per trial
    score = json.load(open(rig["tmp"] / "score.json"))
    m = score["metrics"]
    assert m["batch_size"] == 4096 and m["warmup_runs"] == 1 and m["measured_runs"] == 3 and m["threads"] == 3
    assert m["verified"] is True and len(m["trial_seconds"]) == 3 and len(m["warmup_trial_seconds"]) == 1
    t = sorted(m["trial_seconds"])
    assert m["median_seconds"] == pytest.approx(t[1])
    assert m["p10_seconds"] == pytest.approx(t[0] + (t[1] - t[0]) * 0.2)
    assert score["score"] == pytest.approx(4096 / t[1])
    assert m["aggregate_compressions_per_second"] == pytest.approx(4096 * 3 / sum(t))
    assert m["proof_bytes"] == max(7 + 10 * len(ln.split()[0]) for ln in seeds)
    summary = open(rig["tmp"] / "summary.md").read()
    assert summary.startswith("# Flock BLAKE3 benchmark") and "Trusted verification: `passed`" in summary
    # scratch is emptied after each verified trial
    assert os.listdir(rig["tmp"] / "scratch") == []


def test_percentiles_interpolate_like_the_reference(rig):
    # trials 4, 1, 3, 2 s -> p10 1.3, p50 2.5, p90 3.7 (main.rs:477-482), via a
    # stand-in worker that sleeps the given times before publishing
    _script(rig["worker"], """
        import os, sys, time
        ready, proof = sys.argv[2], sys.argv[3]
        open(ready, "w").write("ready\\n")
        seed = sys.stdin.readline().strip()
        n = len(open(os.environ["SEEDS_FILE"]).read().splitlines()) if os.path.exists(os.environ["SEEDS_FILE"]) else 0
        open(os.environ["SEEDS_FILE"], "a").write(seed + "\\n")
        time.sleep([0.4, 0.1, 0.3, 0.2][n])
        open(proof + ".tmp", "wb").write(b"FLOCK\\x04\\x02" + seed.encode() * 10)
        os.rename(proof + ".tmp", proof)
        """)
    r = _run(rig, "ok", warm=0, runs=4)
    assert r.returncode == 0, r.stderr
    m = json.load(open(rig["tmp"] / "score.json"))["metrics"]
    # each trial is its sleep plus a few ms of process and polling overhead
    for key, want in (("p10_seconds", 0.13), ("median_seconds", 0.25), ("p90_seconds", None)):
        if want is not None:
            assert want <= m[key] < want + 0.05, (key, m[key])
    t = sorted(m["trial_seconds"])
    assert m["p10_seconds"] == pytest.approx(t[0] + 0.3 * (t[1] - t[0]))
    assert m["median_seconds"] == pytest.approx((t[1] + t[2]) / 2)
    assert m["p90_p10_latency_ratio"] == pytest.approx((t[2] + 0.7 * (t[3] - t[2])) / m["p10_seconds"])


@pytest.mark.parametrize("mode,message", [
    ("die", "before signalling readiness"),
    ("silent", "without publishing a proof"),
    ("big", "500007 bytes (allowed 1..500000)"),
    ("empty", "0 bytes (allowed 1..500000)"),
    ("symlink", "symbolic link"),
    ("header", "not a FLOCK v4 R1CS bundle"),
])
def test_rejections(rig, mode, message):
    r = _run(rig, mode, warm=0, runs=1)
    assert r.returncode == 1
    assert message in r.stderr, r.stderr
    assert not os.path.exists(rig["tmp"] / "score.json")


def test_verifier_rejection_fails_the_run(rig):
    _script(str(rig["bin"] / "flock_verify_proof"), "import sys\nsys.exit(2)\n")
    r = _run(rig, "ok", warm=0, runs=1)
    assert r.returncode == 1 and "verifier did not accept the proof" in r.stderr


def test_usage_errors(rig):
    h = str(rig["bin"] / "flock_harness")
    for args in ([], ["w", "s", "o", "m", "7", "1", "0", "1"], ["w", "s", "o", "m", "21", "1", "0", "1"],
                 ["w", "s", "o", "m", "12", "0", "0", "1"], ["w", "s", "o", "m", "12", "1", "0", "0"],
                 ["w", "/", "o", "m", "12", "1", "0", "x"]):
        r = subprocess.run([h] + args, capture_output=True, text=True, timeout=30)
        assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([h, rig["worker"], "/", "o", "m", "12", "1", "0", "1"], capture_output=True, text=True,
                       timeout=30)
    assert r.returncode == 1 and "is not allowed" in r.stderr


def test_worker_and_verifier_usage_without_gpu():
    for prog, args in ((WORKER, ["7", "r", "p"]), (WORKER, ["21", "r", "p"]), (WORKER, []),
                       (VERIFIER, ["12", "x", "f"]), (VERIFIER, ["30", "1", "f"])):
        if not os.path.exists(prog):
            pytest.skip("programs not built")
        r = subprocess.run([prog] + args, capture_output=True, text=True, timeout=30)
        assert r.returncode == 1 and "usage" in r.stderr


@pytest.mark.gpu
def test_worker_matches_library_and_harness_verifies(tmp_path):
    if not gpu_util.gpu_available():
        pytest.skip("no GPU")
    import flockgpu
    log2, seed = 12, 0x5EED_0001
    ready, proof = tmp_path / "w.ready", tmp_path / "w.proof"
    p = subprocess.Popen([WORKER, str(log2), str(ready), str(proof)], stdin=subprocess.PIPE,
                         stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    try:
        out, err = p.communicate(input=(str(seed) + "\n").encode(), timeout=120)
    finally:
        if p.poll() is None:
            p.kill()
    assert p.returncode == 0, err.decode()
    assert ready.read_bytes() == b"ready\n"
    got = proof.read_bytes()
    assert not (tmp_path / "w.proof.tmp").exists()
    ctx = flockgpu.Context(0)
    try:
        setup = flockgpu.Blake3Setup(ctx, 1 << log2)
        want = setup.prove_seeded(log2, seed)
        setup.close()
    finally:
        ctx.close()
    assert got == want
    # the verifier program accepts it, and rejects it under another seed
    r = subprocess.run([VERIFIER, str(log2), str(seed), str(proof)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "verified", r.stderr
    r = subprocess.run([VERIFIER, str(log2), str(seed + 1), str(proof)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "rejected" in r.stderr
    # a short harness run with the real worker: every trial verified
    r = subprocess.run([HARNESS, WORKER, str(tmp_path / "scratch"), str(tmp_path / "score.json"),
                        str(tmp_path / "summary.md"), str(log2), "1", "1", "2"], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr
    m = json.load(open(tmp_path / "score.json"))["metrics"]
    assert m["verified"] is True and m["measured_runs"] == 2 and m["batch_size"] == 1 << log2


def test_verify_reference_script_refuses_without_cargo(tmp_path):
    """flock-challenge_amd/verify_reference.sh (the pristine verifier over
    this library's worker) parses, and where cargo is absent it says so and
    exits 2 without building or running anything."""
    script = os.path.join(gpu_util.PKG, "verify_reference.sh")
    assert subprocess.run(["bash", "-n", script]).returncode == 0
    if shutil.which("cargo"):
        pytest.skip("cargo is installed: the script would build the reference")
    env = dict(os.environ, FLOCK_REFERENCE=str(tmp_path), FLOCK_VERIFY_OUT=str(tmp_path / "out"))
    r = subprocess.run([script], env=env, capture_output=True, text=True, timeout=30)
    assert r.returncode == 2 and "cargo is not installed" in r.stderr
    assert not (tmp_path / "out").exists()
