#!/usr/bin/env bash
# Acceptance of this library's proofs by the challenge's pristine verifier.
#
# The challenge judges proofs with its own harness binary
# (benchmark-tools/harness/src/main.rs, built by
# benchmark-tools/build-trusted-verifier.sh): it starts a worker per trial,
# hands it a fresh seed, captures the published proof and verifies it
# in-process with the unmodified Flock verifier before scoring.  This script
# builds that harness from a checkout of the challenge and points it at this
# library's GPU worker (flock_worker speaks the same protocol), so every
# proof it scores has passed the pristine verifier.
#
# It needs what the challenge needs: a Rust toolchain (cargo, the challenge
# pins 1.97.0) with the crates of its Cargo.lock available.  Where cargo is
# missing it says so and exits 2, running nothing.
#
#   FLOCK_REFERENCE=/path/to/flock-challenge \
#     flock-challenge_amd/verify_reference.sh [LOG2=18] [WARMUP=1] [RUNS=5] [THREADS=1]
#
# Build outputs and results go to ${FLOCK_VERIFY_OUT:-./verify-reference-out}.
set -euo pipefail

log2="${1:-18}"
warmup="${2:-1}"
runs="${3:-5}"
threads="${4:-1}"

here="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
ref="${FLOCK_REFERENCE:?set FLOCK_REFERENCE to a checkout of the challenge repository}"
out="${FLOCK_VERIFY_OUT:-$(pwd)/verify-reference-out}"
worker="${here}/lib/flock_worker"

command -v cargo >/dev/null 2>&1 || {
  echo "verify_reference: cargo is not installed; the pristine verifier cannot be built here" >&2
  exit 2
}
[[ -f "${ref}/benchmark-tools/harness/Cargo.toml" ]] || {
  echo "verify_reference: ${ref} is not a checkout of the challenge (no benchmark-tools/harness)" >&2
  exit 2
}
[[ -x "${worker}" ]] || {
  echo "verify_reference: ${worker} is missing; run python flock-challenge_amd/build.py" >&2
  exit 2
}

mkdir -p "${out}"
out="$(cd "${out}" && pwd -P)"

# The pristine harness, built from the checkout's locked workspace.
(
  cd "${ref}"
  CARGO_INCREMENTAL=0 cargo build --locked --release --target-dir "${out}/target" \
    -p flock-benchmark-harness --bin flock_benchmark_harness
)
harness="${out}/target/release/flock_benchmark_harness"

# The harness clears the worker's environment; the GPU runtime's library
# path comes back through a one-line wrapper that then becomes the worker.
wrapper="${out}/flock_worker_env"
cat > "${wrapper}" <<EOF
#!/bin/sh
export LD_LIBRARY_PATH="${LD_LIBRARY_PATH:-}:/opt/rocm/lib"
exec "${worker}" "\$@"
EOF
chmod 755 "${wrapper}"

rm -rf "${out}/scratch" "${out}/score.json" "${out}/summary.md"
mkdir -p "${out}/scratch"
"${harness}" "${wrapper}" "${out}/scratch" "${out}/score.json" "${out}/summary.md" \
  "${log2}" "${threads}" "${warmup}" "${runs}"
# reaching here means every proof passed the pristine verifier
cat "${out}/summary.md"
