// The BLAKE3 compression R1CS (crates/flock-prover/src/r1cs_hashes/blake3.rs:
// 100-633): slot layout, sparse A_0 / B_0 (C_0 = I), and the statement
// digest (crates/flock-core/src/r1cs.rs:343-365).  Built once per process.
#pragma once
#include <stdint.h>

#include <vector>

namespace flock {
namespace r1cs {

constexpr int K_LOG = 14;
constexpr int K = 1 << K_LOG;
constexpr int K_SKIP = 6;
constexpr int USEFUL_BITS = 15409;
constexpr int Z_CONST = 512;
// slot bases (blake3.rs:178-189)
constexpr int CV_BASE = 0, OUT_LO_BASE = 256, M_BASE = 513, T_LO_BASE = 1025, T_HI_BASE = 1057,
              BLEN_BASE = 1089, FLAGS_BASE = 1121, GS_BASE = 1153, G_STRIDE = 250, OUT_HI_BASE = 15153;

struct Csr {
  std::vector<uint32_t> ptr;   // K + 1
  std::vector<uint32_t> cols;  // column indices, sorted per row
};

struct Matrices {
  Csr a, b;         // rows
  Csr a_t, b_t;     // transposes: per column, the rows holding a 1
};

const Matrices &matrices();
// BLAKE3 digest of the instance for 2^(m - 14) blocks.
void statement_digest(int m, uint8_t out[32]);
// Digests for every m in [m_lo, m_hi], computed in parallel threads (setup
// work for a server that proves many sizes).
void precompute_digests(int m_lo, int m_hi);
// Message word read by G g of round r as operand 0 (mx) or 1 (my).
int msg_index(int r, int g, int which);

}  // namespace r1cs
}  // namespace flock
