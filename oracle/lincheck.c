/*
 * Oracle: lincheck (alpha-batched A/B consistency with the constant-wire pin).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_zc.h"

/* build_quirky_eq_table (lincheck.rs:1127-1142): index i_skip + 64 i_rest. */
static void quirky_eq(of128 z_skip, const of128 *x_inner_rest, of128 *out) {
    of128 lam[64], er[256];
    o_lagrange_s(z_skip, lam);
    o_build_eq(x_inner_rest, 8, er);
    for (int ir = 0; ir < 256; ir++)
        for (int is = 0; is < 64; is++) out[is + 64 * ir] = of128_mul(lam[is], er[ir]);
}

/* fold_alpha_batched over the sparse matrices (lincheck.rs:1158-1240):
 * comb[col] = alpha (A_0^T eq)[col] + (B_0^T eq)[col]. */
void o_comb_vector(of128 alpha, const of128 *eq, of128 *comb) {
    const osparse *A, *B;
    o_r1cs_matrices(&A, &B);
    of128 *ta = (of128 *)calloc(OK_K, sizeof(of128)), *tb = (of128 *)calloc(OK_K, sizeof(of128));
    for (int r = 0; r < OK_K; r++) {
        for (uint32_t k = A->row_ptr[r]; k < A->row_ptr[r + 1]; k++) ta[A->cols[k]] = of128_add(ta[A->cols[k]], eq[r]);
        for (uint32_t k = B->row_ptr[r]; k < B->row_ptr[r + 1]; k++) tb[B->cols[k]] = of128_add(tb[B->cols[k]], eq[r]);
    }
    for (int c = 0; c < OK_K; c++) comb[c] = of128_add(of128_mul(alpha, ta[c]), tb[c]);
    free(ta);
    free(tb);
}

/* z_vec[i] = sum_blk eq(x_outer, blk) z[blk * K + i]
 * (partial_fold_packed_z, lincheck.rs:660-845), 8 blocks per byte-table. */
static void z_fold(const uint64_t *z, int m, const of128 *x_outer, of128 *zv) {
    int nl = m - OK_LOG;
    size_t nb = (size_t)1 << nl;
    of128 *eq = (of128 *)malloc(sizeof(of128) * nb);
    o_build_eq(x_outer, nl, eq);
    memset(zv, 0, sizeof(of128) * OK_K);
    const size_t wpb = OK_K / 64;
    of128 tab[256];
    for (size_t g = 0; g < nb; g += 8) {
        tab[0].lo = tab[0].hi = 0;
        for (int t = 0; t < 8; t++)
            for (int v = 0; v < (1 << t); v++) tab[(1 << t) + v] = of128_add(tab[v], eq[g + t]);
        for (size_t w = 0; w < wpb; w++) {
            uint64_t words[8];
            for (int t = 0; t < 8; t++) words[t] = z[(g + t) * wpb + w];
            for (int bit = 0; bit < 64; bit++) {
                unsigned idx = 0;
                for (int t = 0; t < 8; t++) idx |= (unsigned)((words[t] >> bit) & 1) << t;
                if (idx) zv[w * 64 + bit] = of128_add(zv[w * 64 + bit], tab[idx]);
            }
        }
    }
    free(eq);
}

static void bind_top(of128 *v, size_t half, of128 r) {
    for (size_t i = 0; i < half; i++) v[i] = of128_add(v[i], of128_mul(r, of128_add(v[i + half], v[i])));
}

/* prove_padded_inner (lincheck.rs:1480-1664) */
int o_lincheck_prove(const uint64_t *z, int m, of128 z_skip, const of128 *x_inner_rest, const of128 *x_outer,
                     ofs *t, olc_proof *pf, olc_claim *cl) {
    ofs_label(t, "flock-lincheck-v0");
    of128 alpha = ofs_sample(t);
    of128 beta = ofs_sample(t);
    of128 *eq = (of128 *)malloc(sizeof(of128) * OK_K);
    of128 *comb = (of128 *)malloc(sizeof(of128) * OK_K);
    of128 *zv = (of128 *)malloc(sizeof(of128) * OK_K);
    quirky_eq(z_skip, x_inner_rest, eq);
    o_comb_vector(alpha, eq, comb);
    comb[512] = of128_add(comb[512], beta);
    z_fold(z, m, x_outer, zv);
    size_t len = OK_K;
    of128 rr[8];
    for (int rnd = 0; rnd < 8; rnd++) {
        size_t half = len / 2;
        of128 e1 = {0, 0}, einf = {0, 0};
        for (size_t i = 0; i < half; i++) {
            e1 = of128_add(e1, of128_mul(comb[i + half], zv[i + half]));
            einf = of128_add(einf, of128_mul(of128_add(comb[i], comb[i + half]), of128_add(zv[i], zv[i + half])));
        }
        pf->rounds[rnd][0] = e1;
        pf->rounds[rnd][1] = einf;
        ofs_observe_f128(t, e1);
        ofs_observe_f128(t, einf);
        of128 r = ofs_sample(t);
        rr[rnd] = r;
        bind_top(comb, half, r);
        bind_top(zv, half, r);
        len = half;
    }
    memcpy(pf->z_partial, zv, sizeof(of128) * 64);
    ofs_observe_slice(t, pf->z_partial, 64);
    cl->r_inner_skip = ofs_sample(t);
    of128 lam[64], w = {0, 0};
    o_lagrange_s(cl->r_inner_skip, lam);
    for (int i = 0; i < 64; i++) w = of128_add(w, of128_mul(lam[i], pf->z_partial[i]));
    cl->w = w;
    for (int i = 0; i < 8; i++) cl->r_inner_rest[i] = rr[7 - i];
    free(eq);
    free(comb);
    free(zv);
    return 0;
}

/* verify (lincheck.rs:1666-1830) */
int o_lincheck_verify(const olc_proof *pf, int m, of128 z_skip, const of128 *x_inner_rest, const of128 *x_outer,
                      of128 v_a, of128 v_b, ofs *t, olc_claim *cl) {
    (void)m;
    (void)x_outer;
    ofs_label(t, "flock-lincheck-v0");
    of128 alpha = ofs_sample(t);
    of128 *eq = (of128 *)malloc(sizeof(of128) * OK_K);
    of128 *comb = (of128 *)malloc(sizeof(of128) * OK_K);
    quirky_eq(z_skip, x_inner_rest, eq);
    o_comb_vector(alpha, eq, comb);
    of128 target = of128_add(of128_mul(alpha, v_a), v_b);
    of128 beta = ofs_sample(t);
    comb[512] = of128_add(comb[512], beta);
    target = of128_add(target, beta);
    of128 running = target, rr[8];
    size_t len = OK_K;
    for (int rnd = 0; rnd < 8; rnd++) {
        of128 e1 = pf->rounds[rnd][0], einf = pf->rounds[rnd][1];
        ofs_observe_f128(t, e1);
        ofs_observe_f128(t, einf);
        of128 r = ofs_sample(t);
        of128 e0 = of128_add(running, e1);
        of128 c1 = of128_add(of128_add(e0, e1), einf);
        running = of128_add(of128_add(of128_mul(einf, of128_mul(r, r)), of128_mul(c1, r)), e0);
        bind_top(comb, len / 2, r);
        len /= 2;
        rr[rnd] = r;
    }
    ofs_observe_slice(t, pf->z_partial, 64);
    of128 fs = {0, 0};
    for (int i = 0; i < 64; i++) fs = of128_add(fs, of128_mul(comb[i], pf->z_partial[i]));
    free(eq);
    free(comb);
    if (!of128_eq(fs, running)) return -1;
    cl->r_inner_skip = ofs_sample(t);
    of128 lam[64], w = {0, 0};
    o_lagrange_s(cl->r_inner_skip, lam);
    for (int i = 0; i < 64; i++) w = of128_add(w, of128_mul(lam[i], pf->z_partial[i]));
    cl->w = w;
    for (int i = 0; i < 8; i++) cl->r_inner_rest[i] = rr[7 - i];
    return 0;
}
