// The R1CS statement digest (BlockR1cs::statement_digest,
// crates/flock-core/src/r1cs.rs:343-365, 431-459) as setup work computed once
// per process and kept on disk.
//
// The digest is BLAKE3 over S_m = header(m) || body: a 43-byte header
// ("flock-r1cs-stmt-v1", u64 m, u64 k_log, u64 k_skip, the RowMajor layout
// byte) and the framing of A_0, B_0 and C_0 = I (u64 rows, u64 cols, then
// per row its u64 length and u64 column indices), about 170 MB, the same
// for every m.  BLAKE3 hashes 1 KiB chunks and merges them in a
// left-balanced tree, so only chunk 0 depends on m and only the nodes on
// the tree's left edge change with it.  The "spine" is everything else that
// is needed: the number of chunks and, for each level of the tree, the
// chaining value of the second node (the left-edge node's sibling).  With
// it, the digest of any m is chunk 0's chaining value folded with the
// siblings: about 20 compressions.
//
// The spine is computed on the GPU when a context is at hand
// (spine_device, digest.hip) or on host threads otherwise, and persisted as
// a small file keyed by a fingerprint of the matrices, so later processes
// skip the 170 MB hash altogether.
#pragma once
#include <stdint.h>

#include <array>
#include <string>
#include <vector>

namespace flock {
struct Ctx;
namespace r1cs {

struct Spine {
  uint64_t n_chunks = 0;
  std::vector<std::array<uint32_t, 8>> sib;  // level l's node 1
  bool valid() const { return n_chunks >= 2 && !sib.empty(); }
};

constexpr size_t DIGEST_HEADER_BYTES = 43;

// The body as u64 words (little-endian in memory), and its word count.
std::vector<uint64_t> digest_body_words();
uint64_t digest_body_word_count();
// header(m) || body[0 .. 1024 - 43): chunk 0 of S_m.
void digest_chunk0(int m, uint8_t out[1024]);
// The spine from host threads.
Spine spine_host();
// The digest of S_m from a spine.
void digest_from_spine(const Spine &sp, int m, uint8_t out[32]);
// A fingerprint of the matrices (the cache key).
uint64_t matrices_fingerprint();
// The spine of this process: from memory, else the persisted file, else
// computed (on the GPU through `ctx` when given, else on host threads) and
// persisted.  `from` says where it came from ("memory", "file", "gpu",
// "host").  Thread-safe.
const Spine &spine(Ctx *ctx = nullptr, std::string *from = nullptr);
// The persisted spine's path (FLOCK_CACHE_DIR, else a cache directory next
// to the library).
std::string spine_cache_path();

}  // namespace r1cs
}  // namespace flock
