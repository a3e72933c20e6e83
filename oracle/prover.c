/*
 * Oracle: whole-proof composition, the FLOCK v4 proof format and the verifier.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_zc.h"

/* ------------------------------------------------------------ serialization
 * Header "FLOCK" | VERSION 4 | flavor 2 (crates/flock-prover/src/proof_io.rs:8-13,
 * 86, 504-508), then bincode-1 fixint little-endian of R1csProofBundleLigerito
 * (proof_io.rs:141-200): usize/u64 -> 8 B, Vec -> u64 length + items, unit
 * enums -> u32 index, structs -> fields in declaration order, F128 -> lo, hi. */
typedef struct {
    uint8_t *p;
    size_t n, cap;
} obuf;
static void put(obuf *b, const void *d, size_t n) {
    if (b->n + n > b->cap) {
        b->cap = (b->n + n) * 2 + 1024;
        b->p = (uint8_t *)realloc(b->p, b->cap);
    }
    memcpy(b->p + b->n, d, n);
    b->n += n;
}
static void put_u64(obuf *b, uint64_t v) { put(b, &v, 8); }
static void put_u32(obuf *b, uint32_t v) { put(b, &v, 4); }
static void put_f(obuf *b, of128 v) { put(b, &v, 16); }
static void put_fvec(obuf *b, const of128 *v, size_t n) {
    put_u64(b, n);
    put(b, v, 16 * n);
}
static void put_rec(obuf *b, const orec_proof *r) {
    put_u64(b, r->n_rows);
    for (size_t i = 0; i < r->n_rows; i++) put_fvec(b, r->rows + i * r->row_len, r->row_len);
    put_u64(b, r->n_hashes);
    put(b, r->hashes, 32 * r->n_hashes);
}

int o_serialize(const oproof *pf, uint8_t **out, size_t *out_len) {
    obuf b = {0, 0, 0};
    put(&b, "FLOCK", 5);
    uint8_t hv[2] = {4, 2};
    put(&b, hv, 2);
    put(&b, pf->root, 32);
    put_u64(&b, (uint64_t)pf->params.m);
    put_u64(&b, (uint64_t)pf->params.log_inv_rate);
    put_u64(&b, (uint64_t)pf->params.log_batch_size);
    put_u32(&b, (uint32_t)pf->params.profile);
    put_u32(&b, (uint32_t)pf->params.merkle_hash);
    /* ZerocheckProof (zerocheck.rs:217-231) */
    put_fvec(&b, pf->zc.round1_ab, 64);
    put_fvec(&b, pf->zc.round1_c, 64);
    put_u64(&b, (uint64_t)pf->zc.n_rounds);
    for (int i = 0; i < pf->zc.n_rounds; i++) {
        put_f(&b, pf->zc.rounds[i][0]);
        put_f(&b, pf->zc.rounds[i][1]);
    }
    put_f(&b, pf->zc.final_a);
    put_f(&b, pf->zc.final_b);
    put_f(&b, pf->zc.final_c);
    /* LincheckProof (lincheck.rs:436-445) */
    put_u64(&b, 8);
    for (int i = 0; i < 8; i++) {
        put_f(&b, pf->lc.rounds[i][0]);
        put_f(&b, pf->lc.rounds[i][1]);
    }
    put_fvec(&b, pf->lc.z_partial, 64);
    /* BatchOpeningProofLigerito (pcs.rs:51-55, ring_switch.rs:2503-2505) */
    put_u64(&b, 2);
    put_fvec(&b, pf->s_hat_v[0], 128);
    put_fvec(&b, pf->s_hat_v[1], 128);
    /* LigeritoProof (pcs/ligerito.rs:1504-1545) */
    const olig_proof *l = &pf->lig;
    put(&b, l->initial_root, 32);
    put_rec(&b, &l->initial);
    put_u64(&b, (uint64_t)l->n_rec);
    put(&b, l->roots, 32 * (size_t)l->n_rec);
    put_u64(&b, (uint64_t)(l->n_rec - 1));
    for (int i = 0; i < l->n_rec - 1; i++) put_rec(&b, &l->rec[i]);
    put_fvec(&b, l->yr, l->n_yr);
    put_rec(&b, &l->final_open);
    put_u64(&b, l->n_msgs);
    for (size_t i = 0; i < l->n_msgs; i++) {
        put_f(&b, l->msgs[i][0]);
        put_f(&b, l->msgs[i][1]);
    }
    put_u64(&b, l->n_nonces);
    put(&b, l->nonces, 8 * l->n_nonces);
    put_fvec(&b, l->ood, l->n_ood);
    put_u64(&b, l->n_fold_nonces);
    put(&b, l->fold_nonces, 8 * l->n_fold_nonces);
    *out = b.p;
    *out_len = b.n;
    return 0;
}

typedef struct {
    const uint8_t *p;
    size_t n, off;
    int err;
} ird;
static void get(ird *r, void *d, size_t n) {
    if (r->err || r->off + n > r->n) {
        r->err = 1;
        memset(d, 0, n);
        return;
    }
    memcpy(d, r->p + r->off, n);
    r->off += n;
}
static uint64_t get_u64(ird *r) {
    uint64_t v;
    get(r, &v, 8);
    return v;
}
static uint32_t get_u32(ird *r) {
    uint32_t v;
    get(r, &v, 4);
    return v;
}
static of128 get_f(ird *r) {
    of128 v;
    get(r, &v, 16);
    return v;
}
static of128 *get_fvec(ird *r, size_t *n, size_t maxn) {
    *n = (size_t)get_u64(r);
    if (*n > maxn) {
        r->err = 1;
        *n = 0;
        return NULL;
    }
    of128 *v = (of128 *)malloc(16 * (*n ? *n : 1));
    get(r, v, 16 * *n);
    return v;
}
static void get_rec(ird *r, orec_proof *p) {
    p->n_rows = (size_t)get_u64(r);
    if (p->n_rows > 4096) {
        r->err = 1;
        p->n_rows = 0;
    }
    p->row_len = 0;
    p->rows = NULL;
    for (size_t i = 0; i < p->n_rows && !r->err; i++) {
        size_t len;
        of128 *row = get_fvec(r, &len, 1 << 12);
        if (i == 0) {
            p->row_len = len;
            p->rows = (of128 *)malloc(16 * p->n_rows * (len ? len : 1));
        } else if (len != p->row_len) {
            r->err = 1;
        }
        if (!r->err && row) memcpy(p->rows + i * p->row_len, row, 16 * len);
        free(row);
    }
    p->n_hashes = (size_t)get_u64(r);
    if (p->n_hashes > (1u << 20)) {
        r->err = 1;
        p->n_hashes = 0;
    }
    p->hashes = (uint8_t *)malloc(32 * (p->n_hashes ? p->n_hashes : 1));
    get(r, p->hashes, 32 * p->n_hashes);
}

int o_deserialize(const uint8_t *bytes, size_t len, oproof *pf) {
    memset(pf, 0, sizeof *pf);
    if (len < 7 || memcmp(bytes, "FLOCK", 5) || bytes[5] != 4 || bytes[6] != 2) return -1;
    ird r = {bytes, len, 7, 0};
    get(&r, pf->root, 32);
    pf->params.m = (int)get_u64(&r);
    pf->params.log_inv_rate = (int)get_u64(&r);
    pf->params.log_batch_size = (int)get_u64(&r);
    pf->params.profile = (int)get_u32(&r);
    pf->params.merkle_hash = (int)get_u32(&r);
    size_t n;
    of128 *v = get_fvec(&r, &n, 64);
    if (n == 64) memcpy(pf->zc.round1_ab, v, 16 * 64); else r.err = 1;
    free(v);
    v = get_fvec(&r, &n, 64);
    if (n == 64) memcpy(pf->zc.round1_c, v, 16 * 64); else r.err = 1;
    free(v);
    n = (size_t)get_u64(&r);
    if (n > 64) r.err = 1;
    pf->zc.n_rounds = (int)n;
    pf->zc.rounds = (of128(*)[2])malloc(32 * (n ? n : 1));
    for (size_t i = 0; i < n && !r.err; i++) {
        pf->zc.rounds[i][0] = get_f(&r);
        pf->zc.rounds[i][1] = get_f(&r);
    }
    pf->zc.final_a = get_f(&r);
    pf->zc.final_b = get_f(&r);
    pf->zc.final_c = get_f(&r);
    if (get_u64(&r) != 8) r.err = 1;
    for (int i = 0; i < 8; i++) {
        pf->lc.rounds[i][0] = get_f(&r);
        pf->lc.rounds[i][1] = get_f(&r);
    }
    v = get_fvec(&r, &n, 64);
    if (n == 64) memcpy(pf->lc.z_partial, v, 16 * 64); else r.err = 1;
    free(v);
    if (get_u64(&r) != 2) r.err = 1;
    for (int k = 0; k < 2; k++) {
        v = get_fvec(&r, &n, 128);
        if (n == 128) memcpy(pf->s_hat_v[k], v, 16 * 128); else r.err = 1;
        free(v);
    }
    olig_proof *l = &pf->lig;
    get(&r, l->initial_root, 32);
    get_rec(&r, &l->initial);
    n = (size_t)get_u64(&r);
    if (n > 32 || n < 1) r.err = 1;
    l->n_rec = (int)n;
    l->roots = (uint8_t(*)[32])malloc(32 * (n ? n : 1));
    get(&r, l->roots, 32 * n);
    size_t nrp = (size_t)get_u64(&r);
    if (r.err || nrp + 1 != n) r.err = 1;
    l->rec = (orec_proof *)calloc(n ? n : 1, sizeof(orec_proof));
    for (size_t i = 0; i < nrp && !r.err; i++) get_rec(&r, &l->rec[i]);
    l->yr = get_fvec(&r, &l->n_yr, 1 << 20);
    get_rec(&r, &l->final_open);
    l->n_msgs = (size_t)get_u64(&r);
    if (l->n_msgs > 4096) r.err = 1;
    l->msgs = (of128(*)[2])malloc(32 * (l->n_msgs && !r.err ? l->n_msgs : 1));
    for (size_t i = 0; i < l->n_msgs && !r.err; i++) {
        l->msgs[i][0] = get_f(&r);
        l->msgs[i][1] = get_f(&r);
    }
    l->n_nonces = (size_t)get_u64(&r);
    if (l->n_nonces > 4096) r.err = 1;
    l->nonces = (uint64_t *)malloc(8 * (l->n_nonces && !r.err ? l->n_nonces : 1));
    if (!r.err) get(&r, l->nonces, 8 * l->n_nonces);
    l->ood = get_fvec(&r, &l->n_ood, 4096);
    l->n_fold_nonces = (size_t)get_u64(&r);
    if (l->n_fold_nonces > 4096) r.err = 1;
    l->fold_nonces = (uint64_t *)malloc(8 * (l->n_fold_nonces && !r.err ? l->n_fold_nonces : 1));
    if (!r.err) get(&r, l->fold_nonces, 8 * l->n_fold_nonces);
    if (r.err) return -2;
    if (r.off != r.n) return -3; /* reject trailing bytes (harness/src/main.rs:361-365) */
    return 0;
}

static void rec_free(orec_proof *r) {
    free(r->rows);
    free(r->hashes);
}

void o_proof_free(oproof *p) {
    free(p->zc.rounds);
    olig_proof *l = &p->lig;
    rec_free(&l->initial);
    for (int i = 0; i < l->n_rec - 1; i++) rec_free(&l->rec[i]);
    free(l->rec);
    free(l->roots);
    free(l->yr);
    rec_free(&l->final_open);
    free(l->msgs);
    free(l->nonces);
    free(l->ood);
    free(l->fold_nonces);
    memset(p, 0, sizeof *p);
}

/* ------------------------------------------------------------ claims */
static void make_claims(int m, const ozc_claim *zc, const ozc_proof *zp, const olc_claim *lc, oclaim cl[2]) {
    int n = m - 6;
    /* ab: (r_inner_skip, r_inner_rest, x_ab.x_outer) (r1cs.rs:286-310) */
    cl[0].z_skip = lc->r_inner_skip;
    memcpy(cl[0].x_full, lc->r_inner_rest, 16 * 8);
    memcpy(cl[0].x_full + 8, zc->rhos + 8, 16 * (size_t)(n - 8));
    cl[0].value = lc->w;
    /* c: (z, r_rest) (r1cs.rs:312-330) */
    cl[1].z_skip = zc->z;
    memcpy(cl[1].x_full, zc->r + 6, 16 * (size_t)n);
    cl[1].value = zp->final_c;
}

/* prove_ligerito (crates/flock-prover/src/prover.rs:194-295), the readable
 * composition that prove_fast reproduces byte for byte. */
int o_prove(const ocompression *blocks, size_t n, const uint8_t *domain, size_t domain_len, uint8_t **out,
            size_t *out_len, oproof *keep) {
    int n_log = o_min_n_blocks_log(n);
    int m = OK_LOG + n_log;
    opcs_params pp = {m, 1, 6, 0, 1};
    if (!o_lig_config(m, 0)) return -1;
    size_t words = (size_t)1 << (m - 6);
    uint64_t *z = (uint64_t *)malloc(8 * words), *a = (uint64_t *)malloc(8 * words), *b = (uint64_t *)malloc(8 * words);
    o_witness_packed(blocks, n, n_log, z, a, b);
    size_t npos = (size_t)1 << (m - 7 - 6 + 1);
    of128 *cw = (of128 *)malloc(sizeof(of128) * npos * 64);
    uint8_t *tree = (uint8_t *)malloc(32 * (2 * npos - 1));
    oproof *pf = (oproof *)calloc(1, sizeof(oproof));
    pf->params = pp;
    o_pcs_commit(&pp, (const of128 *)z, cw, tree, pf->root);
    ofs t;
    ofs_init(&t, domain, domain_len);
    /* bind_statement (crates/flock-core/src/proof.rs:51-59) */
    uint8_t dig[32];
    o_statement_digest(m, dig);
    ofs_label(&t, "flock-r1cs-v0");
    ofs_observe_bytes(&t, dig, 32);
    ofs_observe_bytes(&t, pf->root, 32);
    ozc_claim zc;
    o_zerocheck_prove(a, b, z, m, &t, &pf->zc, &zc);
    free(a);
    free(b);
    olc_claim lc;
    o_lincheck_prove(z, m, zc.z, zc.rhos, zc.rhos + 8, &t, &pf->lc, &lc);
    oclaim cl[2];
    make_claims(m, &zc, &pf->zc, &lc, cl);
    o_pcs_open(z, &pp, cw, tree, cl, 2, &t, pf);
    free(z);
    free(cw);
    free(tree);
    free(zc.r);
    free(zc.rhos);
    o_serialize(pf, out, out_len);
    if (keep)
        *keep = *pf;
    else
        o_proof_free(pf);
    free(pf);
    return 0;
}

/* Blake3Setup::verify (blake3.rs:3698) preceded by the harness's checks
 * (harness/src/main.rs:319-366): params, and optionally the commitment root
 * recomputed from the honest witness of `blocks`. */
int o_verify(const uint8_t *bytes, size_t len, size_t n_blocks, const uint8_t *domain, size_t domain_len,
             const ocompression *blocks) {
    oproof pf;
    int rc = o_deserialize(bytes, len, &pf);
    if (rc) {
        o_proof_free(&pf);
        return -10 + rc;
    }
    int n_log = o_min_n_blocks_log(n_blocks);
    int m = OK_LOG + n_log;
    if (pf.params.m != m || pf.params.log_inv_rate != 1 || pf.params.log_batch_size != 6 || pf.params.profile != 0 ||
        pf.params.merkle_hash != 1 || pf.zc.n_rounds != m - 6) {
        o_proof_free(&pf);
        return -20;
    }
    if (blocks) {
        size_t words = (size_t)1 << (m - 6);
        uint64_t *z = (uint64_t *)malloc(8 * words), *a = (uint64_t *)malloc(8 * words),
                 *b = (uint64_t *)malloc(8 * words);
        o_witness_packed(blocks, n_blocks, n_log, z, a, b);
        uint8_t root[32];
        o_pcs_commit(&pf.params, (const of128 *)z, NULL, NULL, root);
        free(z);
        free(a);
        free(b);
        if (memcmp(root, pf.root, 32)) {
            o_proof_free(&pf);
            return -21;
        }
    }
    ofs t;
    ofs_init(&t, domain, domain_len);
    uint8_t dig[32];
    o_statement_digest(m, dig);
    ofs_label(&t, "flock-r1cs-v0");
    ofs_observe_bytes(&t, dig, 32);
    ofs_observe_bytes(&t, pf.root, 32);
    ozc_claim zc;
    rc = o_zerocheck_verify(&pf.zc, m, &t, &zc);
    if (rc) {
        o_proof_free(&pf);
        return -30 + rc;
    }
    olc_claim lc;
    rc = o_lincheck_verify(&pf.lc, m, zc.z, zc.rhos, zc.rhos + 8, pf.zc.final_a, pf.zc.final_b, &t, &lc);
    if (rc) {
        o_proof_free(&pf);
        free(zc.r);
        free(zc.rhos);
        return -40;
    }
    oclaim cl[2];
    make_claims(m, &zc, &pf.zc, &lc, cl);
    rc = o_pcs_verify(&pf, &pf.params, cl, 2, &t);
    free(zc.r);
    free(zc.rhos);
    o_proof_free(&pf);
    return rc ? -1000 + rc : 0;
}
