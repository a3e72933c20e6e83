// Host-side algebra shared by the prover and the verifier: eq tables,
// Lagrange weights on the GF(2^8) node sets, the zerocheck constants and the
// Ligerito query sampler.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "host_field.h"
#include "transcript.h"

namespace flock {

// eq(r, x) for x in [0, 2^n), bit i of x <-> r[i] (LSB first).
void build_eq(const F128 *r, int n, F128 *t);
std::vector<F128> eq_table(const F128 *r, int n);
// Lagrange weights at z over nodes phi8(first .. first + count) within the
// node set phi8(set_first .. set_first + set_count)
// (zerocheck/multilinear.rs:370-389, 509-563).
void lagrange(F128 z, int first, int count, int set_first, int set_count, F128 *w);
F128 dot(const F128 *a, const F128 *b, size_t n);
uint8_t f8_inv(uint8_t a);
// Medium constants beta_i = g / (1 + g), g = x^(2^(i-1))
// (univariate_skip_optimized.rs:119-148).
F128 medium(int i);
// SMALL_CHAL_F8 (univariate_skip_optimized.rs:103-106).
extern const uint8_t SMALL_CHAL[3];
int ceil_log2(size_t n);
// sample_distinct_queries (pcs/ligerito.rs:5830-5850): v.lo mod block_len,
// distinct, then sorted.
std::vector<size_t> sample_queries(Transcript &t, size_t block_len, size_t count);

}  // namespace flock
