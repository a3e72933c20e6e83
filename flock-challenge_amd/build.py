"""Build libflockgpu.so (HIP kernels for gfx950 + host C++) in-tree.

    python flock-challenge_amd/build.py

Outputs flock-challenge_amd/lib/libflockgpu.so and, from csrc/cli/, the
benchmark-protocol programs flock_worker, flock_verify_proof (both linked
against the library) and flock_harness (no GPU code).  Object files go to
flock-challenge_amd/lib/obj/ and are rebuilt when their source (or any
header) is newer.
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
OUT = os.path.join(LIBDIR, "libflockgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]
HIPFLAGS = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def _headers():
    return glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(ROOT, "include", "*.h"))


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in _headers()])
    if os.path.exists(obj) and os.path.getmtime(obj) >= newest and not os.environ.get("FLOCK_FORCE_BUILD"):
        return obj
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIPFLAGS + COMMON + ["-c", src, "-o", obj]
    else:
        # host code: carry-less multiply and SSE4.1 inline everywhere
        cmd = [HIPCC] + COMMON + ["-mpclmul", "-msse4.1", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("compile failed: " + src)
    return obj


def build(force=False):
    os.makedirs(OBJDIR, exist_ok=True)
    if force:
        os.environ["FLOCK_FORCE_BUILD"] = "1"
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", OUT + ".tmp"] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(OUT + ".tmp", OUT)
    build_cli()
    return OUT


# (program, source, links against libflockgpu.so)
CLI = [("flock_worker", "worker.cpp", True), ("flock_verify_proof", "verify_proof.cpp", True),
       ("flock_harness", "harness.cpp", False)]


def build_cli():
    for name, src, with_lib in CLI:
        out = os.path.join(LIBDIR, name)
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(ROOT, "include"),
               os.path.join(CSRC, "cli", src), "-o", out + ".tmp"]
        if with_lib:
            cmd += ["-L" + LIBDIR, "-lflockgpu", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("build failed: " + name)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
