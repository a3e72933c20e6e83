/*
 * CPU oracle for the Flock BLAKE3 prover hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms, used by tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py as the checker.  No product code links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the Layr-Labs/flock-challenge checkout).
 *
 * Pins (tests/test_oracle_*.py): GHASH known answers; the PHI_8_TABLE literal
 * values and the phi8 homomorphism; published BLAKE3 digests and an
 * independent Python BLAKE3 / GF(2^128); SplitMix64 outputs; the novel-basis
 * NTT by brute force on tiny domains; the GF(2^8) S -> Lambda extension
 * against a restatement of the reference's F8 LCH NTT; R1CS satisfaction by
 * honest witnesses and rejection of mutated ones; prove -> verify round
 * trips with tamper rejection; the reference's canary proof length (437,551
 * bytes at m = 32).
 * Parity unpinned (no reference datum reachable here): the statement digest
 * value, Merkle roots, and proof bytes beyond the canary length.
 */
#ifndef FLOCK_ORACLE_H
#define FLOCK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- F128 */
/* crates/flock-core/src/field/gf2_128.rs:22-26: lo = x^0..x^63, hi = x^64..x^127,
 * modulus x^128 + x^7 + x^2 + x + 1. */
typedef struct { uint64_t lo, hi; } of128;

of128 of128_mul(of128 a, of128 b);
of128 of128_mul_portable(of128 a, of128 b);
of128 of128_inv(of128 a);
of128 of128_mul_by_x(of128 a);
static inline of128 of128_add(of128 a, of128 b) { of128 r = {a.lo ^ b.lo, a.hi ^ b.hi}; return r; }
static inline int of128_eq(of128 a, of128 b) { return a.lo == b.lo && a.hi == b.hi; }
/* unreduced 256-bit product (gf2_128/portable.rs:22-36) and reduction (gf2_128.rs:192-214) */
typedef struct { uint64_t r0, r1, r2, r3; } of256;
of256 of128_mul_unreduced(of128 a, of128 b);
of128 of256_reduce(of256 u);

/* GF(2^8), AES polynomial 0x11B (field/gf2_8.rs:11-20) */
uint8_t of8_mul(uint8_t a, uint8_t b);
uint8_t of8_inv(uint8_t a);
/* phi8: GF(2^8) -> GF(2^128) embedding (field/phi8.rs:1047) */
of128 of_phi8(uint8_t a);
extern of128 OF_PHI8_TABLE[256];

/* ---------------------------------------------------------------- BLAKE3 */
/* Public BLAKE3 specification.  Flags per spec. */
enum { OB3_CHUNK_START = 1, OB3_CHUNK_END = 2, OB3_PARENT = 4, OB3_ROOT = 8 };
extern const uint32_t OB3_IV[8];
void ob3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                  uint32_t block_len, uint32_t flags, uint32_t out[16]);

typedef struct {
    uint32_t key[8];
    /* chunk state */
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t block[64];
    uint32_t block_len;
    uint32_t blocks_compressed;
    /* cv stack */
    uint32_t stack[54][8];
    uint32_t stack_len;
} ob3_hasher;

void ob3_init(ob3_hasher *h);
void ob3_update(ob3_hasher *h, const void *data, size_t len);
/* root XOF output from byte 0 (blake3::Hasher::finalize_xof().fill) */
void ob3_finalize_xof(const ob3_hasher *h, uint8_t *out, size_t out_len);
/* non-root chaining value of the whole input (blake3 hazmat finalize_non_root) */
void ob3_finalize_non_root(const ob3_hasher *h, uint8_t out[32]);
void ob3_hash(const void *data, size_t len, uint8_t out[32]);
/* merkle.rs:200-210 */
void ob3_leaf_cv(const void *data, size_t len, uint8_t out[32]);
void ob3_parent_cv(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);

/* ---------------------------------------------------------------- generator */
/* benchmark-tools/common/src/lib.rs:20-33 */
typedef struct {
    uint32_t cv[8];
    uint32_t m[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} ocompression;
void o_generate_compressions(uint32_t log2_size, uint64_t seed, ocompression *out);
uint64_t o_splitmix_next(uint64_t *state); /* full 64-bit output, same mixer */

/* ---------------------------------------------------------------- NTT */
/* ntt/additive_ntt_f128.rs:57-128, 494-557 */
typedef struct {
    int log_d;
    of128 *twiddles; /* breadth-first: layer l at (1<<l)-1 */
} ontt;
int ontt_init(ontt *t, int dim);
void ontt_free(ontt *t);
of128 ontt_twiddle(const ontt *t, int layer, size_t block);
/* ntt/additive_ntt_f128.rs:1281-1320 */
void ontt_forward_interleaved(const ontt *t, of128 *data, size_t n_total, size_t num_ntts,
                              int start_layer);
/* ntt/additive_ntt_f128.rs:2122-2145 */
void ontt_inverse(const ontt *t, of128 *data, size_t n);

/* ---------------------------------------------------------------- Merkle */
/* merkle.rs:706-779, 781-803 (BLAKE3 only) */
void omerkle_tree(const uint8_t *data, size_t data_len, size_t num_leaves, uint8_t *tree /*(2n-1)*32*/);
/* merkle.rs:879-925; returns number of hashes written */
size_t omerkle_multi_proof(const uint8_t *tree, size_t num_leaves, const size_t *positions,
                           size_t n_pos, uint8_t *proof_out, size_t proof_cap);
int omerkle_verify_multi(const uint8_t root[32], size_t num_leaves, const size_t *pos,
                         const uint8_t *leaf_hashes, size_t n, const uint8_t *proof, size_t n_proof);

/* ---------------------------------------------------------------- PCS commit */
/* pcs/commit.rs:250-320, 451-467, 722-784 */
typedef struct {
    int m, log_inv_rate, log_batch_size, profile, merkle_hash;
} opcs_params;
/* codeword (n_positions*num_ntts F128) and tree ((2*n_positions-1)*32 B) may be NULL */
int o_pcs_commit(const opcs_params *p, const of128 *z_packed, of128 *codeword, uint8_t *tree,
                 uint8_t root[32]);

#ifdef __cplusplus
}
#endif
#endif
