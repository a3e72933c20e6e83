/*
 * Oracle: protocol layer (R1CS, transcript, zerocheck, lincheck, ring switch,
 * Ligerito, proof format, verifier).  TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#ifndef FLOCK_ORACLE_PROTO_H
#define FLOCK_ORACLE_PROTO_H

#include "oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define OK_LOG 14
#define OK_K (1 << OK_LOG)
#define OK_SKIP 6
#define OK_USEFUL 15409

/* ---------------------------------------------------------------- R1CS */
typedef struct {
    uint32_t *row_ptr; /* OK_K + 1 */
    uint32_t *cols;
    size_t nnz;
} osparse;

void o_r1cs_matrices(const osparse **a0, const osparse **b0);
void o_statement_digest(int m, uint8_t out[32]);
void o_block_witness(const ocompression *c, uint64_t *z, uint64_t *a, uint64_t *b);
void o_witness_packed(const ocompression *blocks, size_t n, int n_log, uint64_t *z, uint64_t *a, uint64_t *b);
void o_apply_block(const osparse *s, const uint64_t *zblk, uint64_t *out);
int o_min_n_blocks_log(size_t n_blocks);

/* ---------------------------------------------------------------- transcript */
/* FsChallenger over BLAKE3 (crates/flock-core/src/challenger.rs:246-437). */
typedef struct {
    ob3_hasher h;
} ofs;
void ofs_init(ofs *t, const void *domain, size_t len);
void ofs_label(ofs *t, const char *label);
void ofs_observe_f128(ofs *t, of128 v);
void ofs_observe_slice(ofs *t, const of128 *v, size_t n);
void ofs_observe_bytes(ofs *t, const void *b, size_t n);
of128 ofs_sample(ofs *t);
void ofs_sample_vec(ofs *t, of128 *out, size_t n);
uint64_t ofs_grind_pow(ofs *t, uint32_t bits);
int ofs_verify_pow(ofs *t, uint64_t nonce, uint32_t bits);
int o_pow_ok(const uint8_t digest[32], uint64_t nonce, uint32_t bits);

/* ---------------------------------------------------------------- field helpers */
void o_build_eq(const of128 *r, int n, of128 *out); /* out[x] = prod_i (bit_i ? r_i : 1 + r_i) */
void o_lagrange_s(of128 z, of128 *w);              /* 64 weights on phi8(0..63) */
void o_lagrange_lambda(of128 z, of128 *w);         /* 64 weights on phi8(64..127) */
of128 o_interp_combined(const of128 *vals_lambda, of128 z);
void o_ext_matrix(uint8_t M[64][64]); /* Lambda value j = xor_{s: a_s} M[j][s] (F8) */
of128 o_medium_challenge(int i);

/* ---------------------------------------------------------------- proof */
typedef struct {
    of128 round1_ab[64], round1_c[64];
    int n_rounds;
    of128 (*rounds)[2];
    of128 final_a, final_b, final_c;
} ozc_proof;

typedef struct {
    of128 rounds[8][2];
    of128 z_partial[64];
} olc_proof;

typedef struct {
    size_t n_rows, row_len;
    of128 *rows; /* n_rows * row_len */
    size_t n_hashes;
    uint8_t *hashes;
} orec_proof;

typedef struct {
    uint8_t initial_root[32];
    orec_proof initial;
    int n_rec;
    uint8_t (*roots)[32];
    orec_proof *rec;
    size_t n_yr;
    of128 *yr;
    orec_proof final_open;
    size_t n_msgs;
    of128 (*msgs)[2];
    size_t n_nonces;
    uint64_t *nonces;
    size_t n_ood;
    of128 *ood;
    size_t n_fold_nonces;
    uint64_t *fold_nonces;
} olig_proof;

typedef struct {
    uint8_t root[32];
    opcs_params params;
    ozc_proof zc;
    olc_proof lc;
    of128 s_hat_v[2][128];
    olig_proof lig;
} oproof;

/* Prove n compressions; returns serialized bundle bytes (malloc'd). */
int o_prove(const ocompression *blocks, size_t n, const uint8_t *domain, size_t domain_len, uint8_t **out,
            size_t *out_len, oproof *keep /* may be NULL */);
int o_serialize(const oproof *p, uint8_t **out, size_t *out_len);
int o_deserialize(const uint8_t *bytes, size_t len, oproof *p);
/* Verifier (crates/flock-core/src/verifier.rs:53-82 + the harness's commitment
 * check, benchmark-tools/harness/src/main.rs:319-348).  Returns 0 on accept,
 * a negative step code on reject.  If blocks is non-NULL the commitment root
 * must equal the commit of the honest witness of those blocks. */
int o_verify(const uint8_t *bytes, size_t len, size_t n_blocks, const uint8_t *domain, size_t domain_len,
             const ocompression *blocks);
void o_proof_free(oproof *p);

#ifdef __cplusplus
}
#endif
#endif
