/*
 * Oracle: zerocheck with univariate skip (k_skip = 6) and its helpers.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_proto.h"
#include "oracle_zc.h"

static const of128 ONE = {1, 0};

/* build_eq (zerocheck/univariate_skip.rs:31-58): bit i of x <-> r_i. */
void o_build_eq(const of128 *r, int n, of128 *t) {
    t[0] = ONE;
    for (int i = 0; i < n; i++) {
        size_t half = (size_t)1 << i;
        for (size_t x = 0; x < half; x++) {
            of128 v = t[x];
            of128 vr = of128_mul(v, r[i]);
            t[x + half] = vr;
            t[x] = of128_add(v, vr);
        }
    }
}

static void lagrange_nodes(of128 z, int first, int count, int total_first, int total_count, of128 *w) {
    /* weights for nodes phi8(first..first+count) over the node set
     * phi8(total_first..total_first+total_count) */
    for (int i = 0; i < count; i++) {
        of128 si = OF_PHI8_TABLE[first + i];
        of128 num = ONE, den = ONE;
        for (int j = total_first; j < total_first + total_count; j++) {
            if (j == first + i) continue;
            of128 sj = OF_PHI8_TABLE[j];
            num = of128_mul(num, of128_add(z, sj));
            den = of128_mul(den, of128_add(si, sj));
        }
        w[i] = of128_mul(num, of128_inv(den));
    }
}

/* lagrange_weights_naive_reference (zerocheck/multilinear.rs:370-389) */
void o_lagrange_s(of128 z, of128 *w) { lagrange_nodes(z, 0, 64, 0, 64, w); }
/* lagrange weights on Lambda = phi8(64..127) (multilinear.rs:509-520) */
void o_lagrange_lambda(of128 z, of128 *w) { lagrange_nodes(z, 64, 64, 64, 64, w); }

/* interpolate_at_z_combined (multilinear.rs:537-563): degree < 128 poly,
 * zero on S, given its values on Lambda. */
of128 o_interp_combined(const of128 *vals, of128 z) {
    of128 w[64], acc = {0, 0};
    lagrange_nodes(z, 64, 64, 0, 128, w);
    for (int i = 0; i < 64; i++) acc = of128_add(acc, of128_mul(w[i], vals[i]));
    return acc;
}

static of128 dot64(const of128 *a, const of128 *b) {
    of128 acc = {0, 0};
    for (int i = 0; i < 64; i++) acc = of128_add(acc, of128_mul(a[i], b[i]));
    return acc;
}

/* Extension S -> Lambda of the degree-<64 polynomial through (s, a_s),
 * s = 0..63 as GF(2^8) points, evaluated at 64 + j.  The reference reaches
 * the same map by an inverse LCH NTT on S and a forward one on Lambda
 * (zerocheck/univariate_skip.rs:106-120, ntt.rs). */
void o_ext_matrix(uint8_t M[64][64]) {
    for (int j = 0; j < 64; j++) {
        uint8_t lam = (uint8_t)(64 + j);
        for (int s = 0; s < 64; s++) {
            uint8_t num = 1, den = 1;
            for (int t = 0; t < 64; t++) {
                if (t == s) continue;
                num = of8_mul(num, (uint8_t)(lam ^ t));
                den = of8_mul(den, (uint8_t)(s ^ t));
            }
            M[j][s] = of8_mul(num, of8_inv(den));
        }
    }
}

/* medium_challenges_ghash (univariate_skip_optimized.rs:119-148):
 * beta_i = g / (1 + g), g = x^(2^(i-1)). */
of128 o_medium_challenge(int i) {
    of128 g = {1ULL << (1u << i), 0};
    return of128_mul(g, of128_inv(of128_add(ONE, g)));
}

static const uint8_t SMALL_CHAL[3] = {0xF7, 0x53, 0xB5};

/* Row extension tables: E[p][v][j] = xor over bits t of v of M[j][8p+t]. */
static uint8_t (*ext_tables(void))[256][64] {
    static uint8_t E[8][256][64];
    static int built = 0;
    if (!built) {
        static uint8_t M[64][64];
        o_ext_matrix(M);
        for (int p = 0; p < 8; p++)
            for (int v = 0; v < 256; v++)
                for (int j = 0; j < 64; j++) {
                    uint8_t acc = 0;
                    for (int t = 0; t < 8; t++)
                        if ((v >> t) & 1) acc ^= M[j][8 * p + t];
                    E[p][v][j] = acc;
                }
        built = 1;
    }
    return E;
}

static void extend_row(uint64_t bits, uint8_t out[64]) {
    uint8_t(*E)[256][64] = ext_tables();
    memset(out, 0, 64);
    for (int p = 0; p < 8; p++) {
        const uint8_t *e = E[p][(bits >> (8 * p)) & 255];
        for (int j = 0; j < 64; j++) out[j] ^= e[j];
    }
}

/* Round 1 (round1_naive, univariate_skip.rs:84-138), bucketed by F8 value:
 * P^AB(j) = sum_x eq(x) phi8(ahat_j(x) bhat_j(x)), P^C(j) likewise with chat. */
static void round1(const uint64_t *a, const uint64_t *b, const uint64_t *c, int m, const of128 *r_rest, of128 *pab,
                   of128 *pc) {
    int n = m - 6;
    int nlo = n / 2, nhi = n - nlo;
    of128 *eq_lo = (of128 *)malloc(sizeof(of128) << nlo), *eq_hi = (of128 *)malloc(sizeof(of128) << nhi);
    o_build_eq(r_rest, nlo, eq_lo);
    o_build_eq(r_rest + nlo, nhi, eq_hi);
    of128(*bab)[256] = (of128(*)[256])calloc(64 * 256, sizeof(of128));
    of128(*bc)[256] = (of128(*)[256])calloc(64 * 256, sizeof(of128));
    static uint8_t logt[256], expt[512];
    static int tbl = 0;
    if (!tbl) {
        uint8_t x = 1;
        for (int i = 0; i < 255; i++) {
            expt[i] = x;
            logt[x] = (uint8_t)i;
            x = of8_mul(x, 3);
        }
        for (int i = 255; i < 512; i++) expt[i] = expt[i - 255];
        tbl = 1;
    }
    size_t rows = (size_t)1 << n;
    for (size_t x = 0; x < rows; x++) {
        of128 e = of128_mul(eq_lo[x & (((size_t)1 << nlo) - 1)], eq_hi[x >> nlo]);
        uint8_t ea[64], eb[64], ec[64];
        extend_row(a[x], ea);
        extend_row(b[x], eb);
        extend_row(c[x], ec);
        for (int j = 0; j < 64; j++) {
            uint8_t y = (ea[j] && eb[j]) ? expt[logt[ea[j]] + logt[eb[j]]] : 0;
            bab[j][y] = of128_add(bab[j][y], e);
            bc[j][ec[j]] = of128_add(bc[j][ec[j]], e);
        }
    }
    for (int j = 0; j < 64; j++) {
        of128 s1 = {0, 0}, s2 = {0, 0};
        for (int v = 1; v < 256; v++) {
            s1 = of128_add(s1, of128_mul(OF_PHI8_TABLE[v], bab[j][v]));
            s2 = of128_add(s2, of128_mul(OF_PHI8_TABLE[v], bc[j][v]));
        }
        pab[j] = s1;
        pc[j] = s2;
    }
    free(bab);
    free(bc);
    free(eq_lo);
    free(eq_hi);
}

/* fold_at_z_naive (multilinear.rs:598-625) with byte tables of weights. */
static void fold_at_z(const uint64_t *w, size_t rows, const of128 *lw, of128 *out) {
    static of128 T[8][256];
    for (int p = 0; p < 8; p++)
        for (int v = 0; v < 256; v++) {
            of128 acc = {0, 0};
            for (int t = 0; t < 8; t++)
                if ((v >> t) & 1) acc = of128_add(acc, lw[8 * p + t]);
            T[p][v] = acc;
        }
    for (size_t x = 0; x < rows; x++) {
        of128 acc = {0, 0};
        for (int p = 0; p < 8; p++) acc = of128_add(acc, T[p][(w[x] >> (8 * p)) & 255]);
        out[x] = acc;
    }
}

void o_zc_r_vector(ofs *t, int m, of128 *r) {
    ofs_sample_vec(t, r, 6);
    for (int i = 0; i < 3; i++) r[6 + i] = OF_PHI8_TABLE[SMALL_CHAL[i]];
    for (int i = 0; i < 4; i++) r[9 + i] = o_medium_challenge(i);
    ofs_sample_vec(t, r + 13, (size_t)(m - 13));
}

/* Prover (zerocheck.rs:482-1364, transcript order per verify :1365-1520). */
int o_zerocheck_prove(const uint64_t *a, const uint64_t *b, const uint64_t *c, int m, ofs *t, ozc_proof *pf,
                      ozc_claim *cl) {
    int n = m - 6;
    ofs_label(t, "flock-zerocheck-v0");
    of128 *r = (of128 *)malloc(sizeof(of128) * (size_t)m);
    o_zc_r_vector(t, m, r);
    round1(a, b, c, m, r + 6, pf->round1_ab, pf->round1_c);
    ofs_observe_slice(t, pf->round1_ab, 64);
    ofs_observe_slice(t, pf->round1_c, 64);
    of128 z = ofs_sample(t);
    of128 wl[64];
    o_lagrange_lambda(z, wl);
    pf->final_c = dot64(wl, pf->round1_c);

    size_t rows = (size_t)1 << n;
    of128 *A = (of128 *)malloc(sizeof(of128) * rows), *B = (of128 *)malloc(sizeof(of128) * rows);
    of128 ws[64];
    o_lagrange_s(z, ws);
    fold_at_z(a, rows, ws, A);
    fold_at_z(b, rows, ws, B);
    pf->n_rounds = n;
    pf->rounds = (of128(*)[2])malloc(sizeof(of128) * 2 * (size_t)n);
    cl->rhos = (of128 *)malloc(sizeof(of128) * (size_t)n);
    of128 *eq = (of128 *)malloc(sizeof(of128) * (rows / 2 > 0 ? rows / 2 : 1));
    size_t len = rows;
    for (int i = 0; i < n; i++) {
        size_t half = len / 2;
        /* G(X) = sum_x' eq(r[6+i+1..], x') A(X, x') B(X, x'), low bit bound
         * (round_pair_naive, multilinear.rs:639-680) */
        o_build_eq(r + 6 + i + 1, n - i - 1, eq);
        of128 g1 = {0, 0}, ginf = {0, 0};
        for (size_t x = 0; x < half; x++) {
            of128 a0 = A[2 * x], a1 = A[2 * x + 1], b0 = B[2 * x], b1 = B[2 * x + 1];
            g1 = of128_add(g1, of128_mul(eq[x], of128_mul(a1, b1)));
            ginf = of128_add(ginf, of128_mul(eq[x], of128_mul(of128_add(a0, a1), of128_add(b0, b1))));
        }
        pf->rounds[i][0] = g1;
        pf->rounds[i][1] = ginf;
        ofs_observe_f128(t, g1);
        ofs_observe_f128(t, ginf);
        of128 rho = ofs_sample(t);
        cl->rhos[i] = rho;
        /* fold_in_place_pair (multilinear.rs:3209-3230) */
        for (size_t x = 0; x < half; x++) {
            A[x] = of128_add(A[2 * x], of128_mul(rho, of128_add(A[2 * x], A[2 * x + 1])));
            B[x] = of128_add(B[2 * x], of128_mul(rho, of128_add(B[2 * x], B[2 * x + 1])));
        }
        len = half;
    }
    pf->final_a = A[0];
    pf->final_b = B[0];
    ofs_observe_f128(t, pf->final_a);
    ofs_observe_f128(t, pf->final_b);
    cl->z = z;
    cl->r = r;
    free(A);
    free(B);
    free(eq);
    return 0;
}

/* Verifier (zerocheck.rs:1365-1520). */
int o_zerocheck_verify(const ozc_proof *pf, int m, ofs *t, ozc_claim *cl) {
    int n = m - 6;
    if (pf->n_rounds != n) return -1;
    ofs_label(t, "flock-zerocheck-v0");
    of128 *r = (of128 *)malloc(sizeof(of128) * (size_t)m);
    o_zc_r_vector(t, m, r);
    ofs_observe_slice(t, pf->round1_ab, 64);
    ofs_observe_slice(t, pf->round1_c, 64);
    of128 z = ofs_sample(t);
    of128 wl[64];
    o_lagrange_lambda(z, wl);
    of128 pcz = dot64(wl, pf->round1_c);
    if (!of128_eq(pcz, pf->final_c)) {
        free(r);
        return -2;
    }
    of128 comb[64];
    for (int i = 0; i < 64; i++) comb[i] = of128_add(pf->round1_ab[i], pf->round1_c[i]);
    of128 running = of128_add(o_interp_combined(comb, z), pcz);
    cl->rhos = (of128 *)malloc(sizeof(of128) * (size_t)n);
    for (int i = 0; i < n; i++) {
        of128 req = r[6 + i];
        of128 g1 = pf->rounds[i][0], ginf = pf->rounds[i][1];
        of128 g0 = of128_mul(of128_add(running, of128_mul(req, g1)), of128_inv(of128_add(ONE, req)));
        ofs_observe_f128(t, g1);
        ofs_observe_f128(t, ginf);
        of128 rho = ofs_sample(t);
        cl->rhos[i] = rho;
        of128 opr = of128_add(ONE, rho);
        running = of128_add(of128_add(of128_mul(g0, opr), of128_mul(g1, rho)), of128_mul(ginf, of128_mul(rho, opr)));
    }
    if (!of128_eq(running, of128_mul(pf->final_a, pf->final_b))) {
        free(r);
        free(cl->rhos);
        cl->rhos = NULL;
        return -3;
    }
    ofs_observe_f128(t, pf->final_a);
    ofs_observe_f128(t, pf->final_b);
    cl->z = z;
    cl->r = r;
    return 0;
}
