// flock_verify_proof LOG2 SEED PROOF_FILE
//
// The acceptance step the benchmark harness applies to a captured proof
// (verify_proof, benchmark-tools/harness/src/main.rs:307-338):
//  - recompute the commitment root of generate_compressions(LOG2, SEED)'s
//    witness (pcs::commit of the packed witness, on the GPU);
//  - check the bundle's commitment root and parameters against it, replay
//    the "flock-bench-v0" BLAKE3 transcript and run every verifier check
//    (flock_verify_blake3, host code).
// Exit status 0 and "verified" on stdout when accepted; 2 and the reason on
// stderr when rejected; 1 on a usage, I/O or device error.  The harness runs
// it as a separate process so that the harness itself never opens the GPU.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "flock_gpu.h"

namespace {

const char kDomain[] = "flock-bench-v0";  // benchmark-tools/common/src/lib.rs:2

int fail(const std::string &msg) {
  fprintf(stderr, "flock_verify_proof: %s\n", msg.c_str());
  return 1;
}

bool parse_u64(const char *s, uint64_t *out) {
  if (!*s) return false;
  uint64_t v = 0;
  for (; *s; ++s) {
    if (*s < '0' || *s > '9') return false;
    const uint64_t d = (uint64_t)(*s - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  uint64_t log2 = 0, seed = 0;
  if (argc != 4 || !parse_u64(argv[1], &log2) || log2 < 8 || log2 > 20 || !parse_u64(argv[2], &seed))
    return fail("usage: flock_verify_proof LOG2 SEED PROOF_FILE (8 <= LOG2 <= 20)");
  FILE *f = fopen(argv[3], "rb");
  if (!f) return fail(std::string("cannot open ") + argv[3]);
  std::vector<uint8_t> bytes;
  uint8_t chunk[65536];
  size_t got;
  while ((got = fread(chunk, 1, sizeof(chunk), f)) > 0) bytes.insert(bytes.end(), chunk, chunk + got);
  const bool read_err = ferror(f);
  fclose(f);
  if (read_err) return fail(std::string("cannot read ") + argv[3]);

  flock_ctx *ctx = nullptr;
  flock_setup *setup = nullptr;
  if (flock_ctx_create(0, &ctx)) return fail("no GPU context");
  int status = 1;
  uint8_t root[32];
  char err[512] = {0};
  if (flock_setup_create(ctx, 1ull << log2, 0, 1, &setup)) {
    fprintf(stderr, "flock_verify_proof: setup failed: %s\n", flock_last_error(ctx));
  } else if (flock_commitment_root_seeded(setup, (uint32_t)log2, seed, root)) {
    fprintf(stderr, "flock_verify_proof: commitment root failed: %s\n", flock_last_error(ctx));
  } else {
    const int rc = flock_verify_blake3(bytes.data(), bytes.size(), 1ull << log2, (const uint8_t *)kDomain,
                                       sizeof(kDomain) - 1, root, err, sizeof(err));
    if (rc == FLOCK_OK) {
      printf("verified\n");
      status = 0;
    } else {
      fprintf(stderr, "flock_verify_proof: rejected: %s\n", err);
      status = rc == FLOCK_E_VERIFY ? 2 : 1;
    }
  }
  if (setup) flock_setup_destroy(setup);
  flock_ctx_destroy(ctx);
  return status;
}
