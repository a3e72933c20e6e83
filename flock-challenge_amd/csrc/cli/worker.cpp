// flock_worker LOG2 READY_PATH PROOF_PATH
//
// The proving worker of the benchmark protocol
// (benchmark-tools/worker/src/main.rs:12-48), on the GPU through the C ABI:
//  1. a prover setup for 2^LOG2 compressions (8 <= LOG2 <= 20) and one
//     untimed warm-up proof of generate_compressions(LOG2, 0x00C0FFEEBEEFD15C)
//     in the same process, which also reserves every device buffer;
//  2. READY_PATH is written ("ready\n");
//  3. one decimal u64 seed is read from stdin;
//  4. the proof bundle of generate_compressions(LOG2, seed) under the
//     "flock-bench-v0" BLAKE3 transcript is written to PROOF_PATH.tmp and
//     renamed to PROOF_PATH, so a reader never sees a partial file.
// Exit status 0 on success; 1 with a message on stderr otherwise.
#include <errno.h>
#include <fcntl.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <string>
#include <vector>

#include "flock_gpu.h"

namespace {

const char kDomain[] = "flock-bench-v0";             // benchmark-tools/common/src/lib.rs:2
const uint64_t kWarmupSeed = 0x00C0FFEEBEEFD15Cull;  // worker/src/main.rs:28

int fail(const std::string &msg) {
  fprintf(stderr, "flock_worker: %s\n", msg.c_str());
  return 1;
}

// Decimal u64 as Rust's str::parse::<u64> accepts it: an optional '+', then
// one or more digits, no overflow.
bool parse_u64(const std::string &s, uint64_t *out) {
  size_t i = 0;
  if (i < s.size() && s[i] == '+') ++i;
  if (i == s.size()) return false;
  uint64_t v = 0;
  for (; i < s.size(); ++i) {
    if (s[i] < '0' || s[i] > '9') return false;
    const uint64_t d = (uint64_t)(s[i] - '0');
    if (v > (UINT64_MAX - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\n' || s[a] == '\r')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\n' || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

bool write_file(const std::string &path, const void *data, size_t len) {
  const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0) return false;
  const char *p = (const char *)data;
  while (len) {
    const ssize_t w = write(fd, p, len);
    if (w < 0) {
      if (errno == EINTR) continue;
      close(fd);
      return false;
    }
    p += w;
    len -= (size_t)w;
  }
  return close(fd) == 0;
}

struct Prover {
  flock_ctx *ctx = nullptr;
  flock_setup *setup = nullptr;
  std::vector<uint8_t> buf = std::vector<uint8_t>(1 << 20);
  ~Prover() {
    if (setup) flock_setup_destroy(setup);
    if (ctx) flock_ctx_destroy(ctx);
  }
  // The proof bytes for generate_compressions(log2, seed), or an error text.
  int prove(uint32_t log2, uint64_t seed, size_t *len, std::string *err) {
    int rc = flock_prove_blake3_seeded(setup, log2, seed, (const uint8_t *)kDomain, sizeof(kDomain) - 1, buf.data(),
                                       buf.size(), len);
    if (rc == FLOCK_E_CAP) {
      buf.resize(*len);
      rc = flock_last_proof(ctx, buf.data(), buf.size(), len);
    }
    if (rc) *err = std::string("prove failed: ") + flock_last_error(ctx);
    return rc;
  }
};

}  // namespace

int main(int argc, char **argv) {
  uint64_t log2 = 0;
  if (argc != 4 || !parse_u64(argv[1], &log2) || log2 < 8 || log2 > 20)
    return fail("usage: flock_worker LOG2 READY_PATH PROOF_PATH (8 <= LOG2 <= 20)");
  const std::string ready_path = argv[2], proof_path = argv[3];

  Prover P;
  if (flock_ctx_create(0, &P.ctx)) return fail("no GPU context");
  if (flock_setup_create(P.ctx, 1ull << log2, 0, 1, &P.setup))
    return fail(std::string("setup failed: ") + flock_last_error(P.ctx));
  size_t len = 0;
  std::string err;
  // untimed warm-up in the same process, before readiness
  if (P.prove((uint32_t)log2, kWarmupSeed, &len, &err)) return fail(err);

  static const char kReady[] = "ready\n";
  if (!write_file(ready_path, kReady, sizeof(kReady) - 1)) return fail("cannot write " + ready_path);

  std::string line;
  int c;
  bool any = false;
  while ((c = fgetc(stdin)) != EOF) {
    any = true;
    if (c == '\n') break;
    line.push_back((char)c);
  }
  if (!any) return fail("missing seed on stdin");
  uint64_t seed = 0;
  if (!parse_u64(trim(line), &seed)) return fail("seed is not a decimal u64: " + line);

  if (P.prove((uint32_t)log2, seed, &len, &err)) return fail(err);
  const std::string tmp = proof_path + ".tmp";
  if (!write_file(tmp, P.buf.data(), len)) return fail("cannot write " + tmp);
  if (rename(tmp.c_str(), proof_path.c_str()) != 0) return fail("cannot rename to " + proof_path);
  return 0;
}
