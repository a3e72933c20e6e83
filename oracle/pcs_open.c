/*
 * Oracle: PCS opening = ring switch (per claim) + recursive Ligerito with an
 * explicit basis, and the matching verifier.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_zc.h"

static const of128 ONE = {1, 0};
static const of128 ZERO = {0, 0};

const olig_config *o_lig_config(int m, int profile) {
    for (int i = 0; i < OLIG_N_CONFIGS; i++)
        if (OLIG_CONFIGS[i].m == m && OLIG_CONFIGS[i].profile == profile) return &OLIG_CONFIGS[i];
    return NULL;
}

void o_build_eq_lsb(const of128 *r, int n, of128 *out) { o_build_eq(r, n, out); }

static int ceil_log2(size_t n) {
    int l = 0;
    if (n <= 1) return 0;
    n -= 1;
    while (n) {
        l++;
        n >>= 1;
    }
    return l;
}

/* ------------------------------------------------------------ ring switch */

/* fold_1b_rows_naive (ring_switch.rs:1568-1618): s[r] = sum_i bit_r(z[i]) eq[i] */
static void fold_1b_rows(const uint64_t *z, size_t L, const of128 *eq, of128 *s) {
    for (int r = 0; r < 128; r++) s[r] = ZERO;
    of128 tab[256];
    for (size_t g = 0; g < L; g += 8) {
        size_t cnt = L - g < 8 ? L - g : 8;
        tab[0] = ZERO;
        for (size_t t = 0; t < 8; t++)
            for (int v = 0; v < (1 << t); v++) tab[(1 << t) + v] = t < cnt ? of128_add(tab[v], eq[g + t]) : tab[v];
        for (int half = 0; half < 2; half++)
            for (int bit = 0; bit < 64; bit++) {
                unsigned idx = 0;
                for (size_t t = 0; t < cnt; t++) idx |= (unsigned)((z[2 * (g + t) + half] >> bit) & 1) << t;
                if (idx) s[64 * half + bit] = of128_add(s[64 * half + bit], tab[idx]);
            }
    }
}

/* tensor_algebra_transpose (ring_switch.rs:1654-1718) */
static void transpose128(const of128 *v, of128 *u) {
    for (int b = 0; b < 128; b++) u[b] = ZERO;
    for (int i = 0; i < 128; i++)
        for (int b = 0; b < 128; b++) {
            uint64_t bit = b < 64 ? (v[i].lo >> b) & 1 : (v[i].hi >> (b - 64)) & 1;
            if (!bit) continue;
            if (i < 64)
                u[b].lo |= 1ULL << i;
            else
                u[b].hi |= 1ULL << (i - 64);
        }
}

/* fold_b128_elems (ring_switch.rs:1720-1760), accumulated with weight g:
 * out[i] += g * sum_b bit_b(t[i]) e[b]. */
static void fold_b128_acc(const of128 *t, size_t L, const of128 *e, of128 g, of128 *out) {
    of128 ge[128];
    for (int b = 0; b < 128; b++) ge[b] = of128_mul(g, e[b]);
    static of128 tab[16][256];
    for (int p = 0; p < 16; p++)
        for (int v = 0; v < 256; v++) {
            of128 acc = ZERO;
            for (int k = 0; k < 8; k++)
                if ((v >> k) & 1) acc = of128_add(acc, ge[8 * p + k]);
            tab[p][v] = acc;
        }
    for (size_t i = 0; i < L; i++) {
        of128 acc = out[i];
        for (int p = 0; p < 8; p++) acc = of128_add(acc, tab[p][(t[i].lo >> (8 * p)) & 255]);
        for (int p = 0; p < 8; p++) acc = of128_add(acc, tab[8 + p][(t[i].hi >> (8 * p)) & 255]);
        out[i] = acc;
    }
}

/* ------------------------------------------------------------ sumcheck (Ligerito) */
static void round_msg(const of128 *f, const of128 *b, size_t n, of128 *u0, of128 *u2) {
    of128 a0 = ZERO, a2 = ZERO;
    for (size_t j = 0; j < n / 2; j++) {
        of128 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        a0 = of128_add(a0, of128_mul(f0, b0));
        a2 = of128_add(a2, of128_mul(of128_add(f0, f1), of128_add(b0, b1)));
    }
    *u0 = a0;
    *u2 = a2;
}

static void fold_lsb(of128 *v, size_t n, of128 r) {
    for (size_t j = 0; j < n / 2; j++) v[j] = of128_add(v[2 * j], of128_mul(r, of128_add(v[2 * j], v[2 * j + 1])));
}

typedef struct {
    of128 *f, *cb;
    size_t n;
    of128 tr;
    of128 (*msgs)[2];
    size_t n_msgs, cap_msgs;
} osc;

static void push_msg(osc *s, of128 u0, of128 u2, ofs *t) {
    if (s->n_msgs == s->cap_msgs) {
        s->cap_msgs = s->cap_msgs ? 2 * s->cap_msgs : 64;
        s->msgs = (of128(*)[2])realloc(s->msgs, sizeof(of128) * 2 * s->cap_msgs);
    }
    s->msgs[s->n_msgs][0] = u0;
    s->msgs[s->n_msgs][1] = u2;
    s->n_msgs++;
    ofs_observe_f128(t, u0);
    ofs_observe_f128(t, u2);
}

/* SumcheckProver::fold (ligerito.rs:5517-5570): fold f and the combined basis
 * at r (LSB), then the next round message. */
static void sc_fold(osc *s, of128 r, ofs *t) {
    fold_lsb(s->f, s->n, r);
    fold_lsb(s->cb, s->n, r);
    s->n /= 2;
    of128 u0, u2;
    round_msg(s->f, s->cb, s->n, &u0, &u2);
    push_msg(s, u0, u2, t);
}

/* ------------------------------------------------------------ novel basis */

/* eval_sk_at_vks (ligerito.rs:1747-1774): s_k(v_k), standard basis v_i = x^i */
static void sks_vks(int n, of128 *out) {
    of128 *layer = (of128 *)malloc(sizeof(of128) * (size_t)(n + 1));
    out[0] = ONE;
    for (int i = 1; i <= n; i++) {
        layer[i - 1].lo = 1ULL << i;
        layer[i - 1].hi = 0;
    }
    int cur = n;
    for (int i = 0; i < n; i++) {
        for (int j = 0; j < cur; j++) {
            of128 s = of128_add(of128_mul(layer[j], layer[j]), of128_mul(out[i], layer[j]));
            if (j == 0)
                out[i + 1] = s;
            else
                layer[j - 1] = s;
        }
        cur--;
    }
    free(layer);
}

/* normalized W^_k(x) for k < n (ligerito.rs:1776-1824) */
static void w_hat_at(int n, const of128 *svk, of128 x, of128 *w) {
    if (n <= 0) return;
    w[0] = x;
    for (int i = 1; i < n; i++)
        w[i] = of128_add(of128_mul(w[i - 1], w[i - 1]), of128_mul(svk[i - 1], w[i - 1]));
    for (int i = 0; i < n; i++) w[i] = of128_mul(w[i], of128_inv(svk[i]));
}

/* induce_sumcheck_poly (ligerito.rs:1989-2160):
 * basis[j] = sum_i eq(alpha, i) X^_j(q_i), X^_j = prod_{bit k of j} W^_k;
 * enforced = sum_i eq(alpha, i) <row_i, eq(v)>. */
static void induce(int n, const of128 *rows, size_t row_len, const of128 *v, int nv, const size_t *q, size_t nq,
                   const of128 *alpha, int na, of128 *basis, of128 *enforced) {
    of128 *svk = (of128 *)malloc(sizeof(of128) * (size_t)(n + 1));
    sks_vks(n, svk);
    of128 *aw = (of128 *)malloc(sizeof(of128) * ((size_t)1 << na));
    o_build_eq(alpha, na, aw);
    of128 *ev = (of128 *)malloc(sizeof(of128) * ((size_t)1 << nv));
    o_build_eq(v, nv, ev);
    size_t N = (size_t)1 << n;
    memset(basis, 0, sizeof(of128) * N);
    of128 *w = (of128 *)malloc(sizeof(of128) * (size_t)(n > 0 ? n : 1));
    of128 *tmp = (of128 *)malloc(sizeof(of128) * N);
    of128 es = ZERO;
    for (size_t i = 0; i < nq; i++) {
        of128 dot = ZERO;
        for (size_t k = 0; k < row_len; k++) dot = of128_add(dot, of128_mul(rows[i * row_len + k], ev[k]));
        es = of128_add(es, of128_mul(aw[i], dot));
        of128 x = {(uint64_t)q[i], 0};
        w_hat_at(n, svk, x, w);
        tmp[0] = aw[i];
        for (int k = 0; k < n; k++) {
            size_t cl = (size_t)1 << k;
            for (size_t j = 0; j < cl; j++) tmp[j + cl] = of128_mul(w[k], tmp[j]);
        }
        for (size_t j = 0; j < N; j++) basis[j] = of128_add(basis[j], tmp[j]);
    }
    *enforced = es;
    free(svk);
    free(aw);
    free(ev);
    free(w);
    free(tmp);
}

/* closed form of MLE(induced basis)(p) (ligerito.rs:1879-1987):
 * sum_i eq(alpha,i) prod_k (1 + p_k (1 + W^_k(q_i))) */
static of128 induced_mle(int n, const size_t *q, size_t nq, const of128 *alpha, int na, const of128 *p) {
    of128 *svk = (of128 *)malloc(sizeof(of128) * (size_t)(n + 1));
    sks_vks(n, svk);
    of128 *aw = (of128 *)malloc(sizeof(of128) * ((size_t)1 << na));
    o_build_eq(alpha, na, aw);
    of128 *w = (of128 *)malloc(sizeof(of128) * (size_t)(n > 0 ? n : 1));
    of128 sum = ZERO;
    for (size_t i = 0; i < nq; i++) {
        of128 x = {(uint64_t)q[i], 0};
        w_hat_at(n, svk, x, w);
        of128 prod = ONE;
        for (int k = 0; k < n; k++) prod = of128_mul(prod, of128_add(ONE, of128_mul(p[k], of128_add(ONE, w[k]))));
        sum = of128_add(sum, of128_mul(aw[i], prod));
    }
    free(svk);
    free(aw);
    free(w);
    return sum;
}

/* ------------------------------------------------------------ Ligero commit */
typedef struct {
    of128 *cw;
    uint8_t *tree;
    size_t block_len, lanes;
} owit;

/* ligero_commit (ligerito.rs:3340-3560): f as [pos][lane] (lane = low k bits),
 * replicate 2^rate times, NTT of dimension log_msg_cols + rate from layer
 * rate, Merkle over positions. */
static void ligero_commit(const of128 *f, int log_msg_cols, int k, int rate, owit *w) {
    size_t lanes = (size_t)1 << k, msg = (size_t)1 << (log_msg_cols + k);
    int dim = log_msg_cols + rate;
    w->block_len = (size_t)1 << dim;
    w->lanes = lanes;
    w->cw = (of128 *)malloc(sizeof(of128) * w->block_len * lanes);
    for (size_t off = 0; off < w->block_len * lanes; off += msg) memcpy(w->cw + off, f, sizeof(of128) * msg);
    ontt t;
    ontt_init(&t, dim);
    ontt_forward_interleaved(&t, w->cw, w->block_len * lanes, lanes, rate);
    ontt_free(&t);
    w->tree = (uint8_t *)malloc(32 * (2 * w->block_len - 1));
    omerkle_tree((const uint8_t *)w->cw, w->block_len * lanes * 16, w->block_len, w->tree);
}

/* sample_distinct_queries (ligerito.rs:5830-5851) */
static void sample_queries(ofs *t, size_t block_len, size_t count, size_t *out) {
    size_t got = 0;
    while (got < count) {
        of128 v = ofs_sample(t);
        size_t q = (size_t)(v.lo % block_len);
        int dup = 0;
        for (size_t i = 0; i < got; i++)
            if (out[i] == q) dup = 1;
        if (!dup) out[got++] = q;
    }
    for (size_t i = 1; i < count; i++)
        for (size_t j = i; j > 0 && out[j - 1] > out[j]; j--) {
            size_t x = out[j];
            out[j] = out[j - 1];
            out[j - 1] = x;
        }
}

static void open_rows(const of128 *cw, const uint8_t *tree, size_t block_len, size_t lanes, const size_t *q, size_t nq,
                      orec_proof *p) {
    p->n_rows = nq;
    p->row_len = lanes;
    p->rows = (of128 *)malloc(sizeof(of128) * nq * lanes);
    for (size_t i = 0; i < nq; i++) memcpy(p->rows + i * lanes, cw + q[i] * lanes, sizeof(of128) * lanes);
    size_t cap = nq * 64 + 1;
    p->hashes = (uint8_t *)malloc(32 * cap);
    p->n_hashes = omerkle_multi_proof(tree, block_len, q, nq, p->hashes, cap);
}

static void u64_push(uint64_t **arr, size_t *n, uint64_t v) {
    *arr = (uint64_t *)realloc(*arr, sizeof(uint64_t) * (*n + 1));
    (*arr)[(*n)++] = v;
}
static void f128_push(of128 **arr, size_t *n, of128 v) {
    *arr = (of128 *)realloc(*arr, sizeof(of128) * (*n + 1));
    (*arr)[(*n)++] = v;
}

/* ------------------------------------------------------------ prover */
int o_pcs_open(const uint64_t *z, const opcs_params *pp, const of128 *codeword, const uint8_t *tree,
               const oclaim *claims, int n_claims, ofs *t, oproof *pf) {
    const olig_config *cfg = o_lig_config(pp->m, pp->profile);
    if (!cfg) return -1;
    int L = pp->m - 7;
    size_t N = (size_t)1 << L;
    ofs_label(t, "flock-pcs-open-batch-v0");
    /* ring switch per claim (ring_switch.rs:2708-2780, 3040-3117) */
    of128 *eqs = (of128 *)malloc(sizeof(of128) * N);
    of128 eq_rd[2][128], sclaim[2];
    for (int k = 0; k < n_claims; k++) {
        ofs_label(t, "flock-ring-switch-v0");
        o_build_eq(claims[k].x_full + 1, L, eqs);
        fold_1b_rows(z, N, eqs, pf->s_hat_v[k]);
        ofs_observe_slice(t, pf->s_hat_v[k], 128);
        of128 rd[7], u[128];
        ofs_sample_vec(t, rd, 7);
        o_build_eq(rd, 7, eq_rd[k]);
        transpose128(pf->s_hat_v[k], u);
        sclaim[k] = ZERO;
        for (int b = 0; b < 128; b++) sclaim[k] = of128_add(sclaim[k], of128_mul(u[b], eq_rd[k][b]));
    }
    of128 gam[2];
    for (int k = 0; k < n_claims; k++) gam[k] = ofs_sample(t);
    osc s;
    memset(&s, 0, sizeof s);
    s.n = N;
    s.f = (of128 *)malloc(sizeof(of128) * N);
    memcpy(s.f, z, sizeof(of128) * N);
    s.cb = (of128 *)calloc(N, sizeof(of128));
    of128 target = ZERO;
    for (int k = 0; k < n_claims; k++) {
        o_build_eq(claims[k].x_full + 1, L, eqs);
        fold_b128_acc(eqs, N, eq_rd[k], gam[k], s.cb);
        target = of128_add(target, of128_mul(gam[k], sclaim[k]));
    }
    free(eqs);
    s.tr = target;

    /* recursive_prover_with_basis_impl (ligerito.rs:6222-6980) */
    olig_proof *lp = &pf->lig;
    memset(lp, 0, sizeof *lp);
    int initial_k = cfg->initial_k, r = cfg->n_levels - 1;
    ofs_label(t, "flock-ligerito-basis-v0");
    ofs_observe_f128(t, target);
    size_t block_len_0 = (size_t)1 << (L - initial_k + cfg->levels[0].log_inv_rate);
    memcpy(lp->initial_root, tree + 32 * (2 * block_len_0 - 2), 32);
    ofs_observe_bytes(t, lp->initial_root, 32);
    of128 u0, u2;
    round_msg(s.f, s.cb, s.n, &u0, &u2);
    push_msg(&s, u0, u2, t);
    of128 ris[64];
    int nris = 0;
    for (int j = 0; j < initial_k; j++) {
        int bits = cfg->levels[0].fold_grinding_bits - j;
        if (bits > 0) u64_push(&lp->fold_nonces, &lp->n_fold_nonces, ofs_grind_pow(t, (uint32_t)bits));
        of128 rj = ofs_sample(t);
        ris[nris++] = rj;
        sc_fold(&s, rj, t);
    }
    int n1 = L - initial_k;
    lp->roots = (uint8_t(*)[32])malloc(32 * (size_t)(r + 1));
    lp->rec = (orec_proof *)calloc((size_t)(r + 1), sizeof(orec_proof));
    owit prev;
    {
        int k1 = cfg->levels[1].k_recursive;
        ligero_commit(s.f, n1 - k1, k1, cfg->levels[1].log_inv_rate, &prev);
        memcpy(lp->roots[0], prev.tree + 32 * (2 * prev.block_len - 2), 32);
        lp->n_rec = 1;
        ofs_observe_bytes(t, lp->roots[0], 32);
    }
    of128 *eqz = (of128 *)malloc(sizeof(of128) * s.n);
    for (int o = 0; o < cfg->levels[1].ood_samples; o++) {
        of128 zz[64];
        ofs_sample_vec(t, zz, (size_t)n1);
        o_build_eq(zz, n1, eqz);
        of128 y = ZERO;
        for (size_t x = 0; x < s.n; x++) y = of128_add(y, of128_mul(s.f[x], eqz[x]));
        round_msg(s.f, eqz, s.n, &u0, &u2);
        ofs_observe_f128(t, y);
        f128_push(&lp->ood, &lp->n_ood, y);
        push_msg(&s, u0, u2, t);
        of128 be = ofs_sample(t);
        for (size_t x = 0; x < s.n; x++) s.cb[x] = of128_add(s.cb[x], of128_mul(be, eqz[x]));
        s.tr = of128_add(s.tr, of128_mul(be, y));
    }
    u64_push(&lp->nonces, &lp->n_nonces, ofs_grind_pow(t, (uint32_t)cfg->levels[0].grinding_bits));
    size_t nq = (size_t)cfg->levels[0].queries;
    size_t *q = (size_t *)malloc(sizeof(size_t) * nq);
    sample_queries(t, block_len_0, nq, q);
    int na = ceil_log2(nq);
    of128 alpha[16];
    ofs_sample_vec(t, alpha, (size_t)na);
    open_rows(codeword, tree, block_len_0, (size_t)1 << initial_k, q, nq, &lp->initial);
    of128 *basis = (of128 *)malloc(sizeof(of128) * s.n);
    of128 es;
    induce(n1, lp->initial.rows, lp->initial.row_len, ris, initial_k, q, nq, alpha, na, basis, &es);
    free(q);
    round_msg(s.f, basis, s.n, &u0, &u2);
    push_msg(&s, u0, u2, t);
    of128 be = ofs_sample(t);
    for (size_t x = 0; x < s.n; x++) s.cb[x] = of128_add(s.cb[x], of128_mul(be, basis[x]));
    s.tr = of128_add(s.tr, of128_mul(be, es));
    free(basis);

    for (int i = 0; i < r; i++) {
        int ki = cfg->levels[i + 1].k_recursive;
        of128 lr[16];
        for (int j = 0; j < ki; j++) {
            int bits = cfg->levels[i + 1].fold_grinding_bits - j;
            if (bits > 0) u64_push(&lp->fold_nonces, &lp->n_fold_nonces, ofs_grind_pow(t, (uint32_t)bits));
            of128 rj = ofs_sample(t);
            ris[nris++] = rj;
            lr[j] = rj;
            sc_fold(&s, rj, t);
        }
        if (i == r - 1) {
            lp->n_yr = s.n;
            lp->yr = (of128 *)malloc(sizeof(of128) * s.n);
            memcpy(lp->yr, s.f, sizeof(of128) * s.n);
            for (size_t x = 0; x < s.n; x++) ofs_observe_f128(t, s.f[x]);
            u64_push(&lp->nonces, &lp->n_nonces, ofs_grind_pow(t, (uint32_t)cfg->levels[i + 1].grinding_bits));
            nq = (size_t)cfg->levels[i + 1].queries;
            q = (size_t *)malloc(sizeof(size_t) * nq);
            sample_queries(t, prev.block_len, nq, q);
            open_rows(prev.cw, prev.tree, prev.block_len, prev.lanes, q, nq, &lp->final_open);
            free(q);
            break;
        }
        int n_next = 0;
        while (((size_t)1 << n_next) < s.n) n_next++;
        int kn = cfg->levels[i + 2].k_recursive;
        owit next;
        ligero_commit(s.f, n_next - kn, kn, cfg->levels[i + 2].log_inv_rate, &next);
        memcpy(lp->roots[lp->n_rec], next.tree + 32 * (2 * next.block_len - 2), 32);
        ofs_observe_bytes(t, lp->roots[lp->n_rec], 32);
        lp->n_rec++;
        for (int o = 0; o < cfg->levels[i + 2].ood_samples; o++) {
            of128 zz[64];
            ofs_sample_vec(t, zz, (size_t)n_next);
            o_build_eq(zz, n_next, eqz);
            of128 y = ZERO;
            for (size_t x = 0; x < s.n; x++) y = of128_add(y, of128_mul(s.f[x], eqz[x]));
            round_msg(s.f, eqz, s.n, &u0, &u2);
            ofs_observe_f128(t, y);
            f128_push(&lp->ood, &lp->n_ood, y);
            push_msg(&s, u0, u2, t);
            of128 b2 = ofs_sample(t);
            for (size_t x = 0; x < s.n; x++) s.cb[x] = of128_add(s.cb[x], of128_mul(b2, eqz[x]));
            s.tr = of128_add(s.tr, of128_mul(b2, y));
        }
        u64_push(&lp->nonces, &lp->n_nonces, ofs_grind_pow(t, (uint32_t)cfg->levels[i + 1].grinding_bits));
        nq = (size_t)cfg->levels[i + 1].queries;
        q = (size_t *)malloc(sizeof(size_t) * nq);
        sample_queries(t, prev.block_len, nq, q);
        na = ceil_log2(nq);
        ofs_sample_vec(t, alpha, (size_t)na);
        orec_proof *rp = &lp->rec[i];
        open_rows(prev.cw, prev.tree, prev.block_len, prev.lanes, q, nq, rp);
        basis = (of128 *)malloc(sizeof(of128) * s.n);
        induce(n_next, rp->rows, rp->row_len, lr, ki, q, nq, alpha, na, basis, &es);
        free(q);
        round_msg(s.f, basis, s.n, &u0, &u2);
        push_msg(&s, u0, u2, t);
        be = ofs_sample(t);
        for (size_t x = 0; x < s.n; x++) s.cb[x] = of128_add(s.cb[x], of128_mul(be, basis[x]));
        s.tr = of128_add(s.tr, of128_mul(be, es));
        free(basis);
        free(prev.cw);
        free(prev.tree);
        prev = next;
    }
    free(prev.cw);
    free(prev.tree);
    free(eqz);
    lp->msgs = s.msgs;
    lp->n_msgs = s.n_msgs;
    free(s.f);
    free(s.cb);
    return 0;
}

/* ------------------------------------------------------------ verifier */
typedef struct {
    of128 c, b, a;
} quad;
static quad q_from(of128 u0, of128 u2, of128 tr) {
    quad q = {u0, of128_add(tr, u2), u2};
    return q;
}
static of128 q_eval(quad q, of128 r) {
    return of128_add(of128_add(q.c, of128_mul(r, q.b)), of128_mul(of128_mul(r, r), q.a));
}
static quad q_fold(quad p, quad q, of128 al) {
    quad o = {of128_add(p.c, of128_mul(al, q.c)), of128_add(p.b, of128_mul(al, q.b)), of128_add(p.a, of128_mul(al, q.a))};
    return o;
}

static of128 enforced_sum(const orec_proof *p, const of128 *v, int nv, const of128 *alpha, int na) {
    of128 *ev = (of128 *)malloc(sizeof(of128) * ((size_t)1 << nv));
    of128 *aw = (of128 *)malloc(sizeof(of128) * ((size_t)1 << na));
    o_build_eq(v, nv, ev);
    o_build_eq(alpha, na, aw);
    of128 s = ZERO;
    for (size_t i = 0; i < p->n_rows; i++) {
        of128 d = ZERO;
        for (size_t k = 0; k < p->row_len; k++) d = of128_add(d, of128_mul(p->rows[i * p->row_len + k], ev[k]));
        s = of128_add(s, of128_mul(aw[i], d));
    }
    free(ev);
    free(aw);
    return s;
}

static int verify_opens(const uint8_t root[32], size_t block_len, const size_t *q, size_t nq, const orec_proof *p,
                        size_t lanes) {
    if (p->n_rows != nq || p->row_len != lanes) return 0;
    uint8_t *lh = (uint8_t *)malloc(32 * (nq ? nq : 1));
    for (size_t i = 0; i < nq; i++) ob3_leaf_cv(p->rows + i * lanes, 16 * lanes, lh + 32 * i);
    int ok = omerkle_verify_multi(root, block_len, q, lh, nq, p->hashes, p->n_hashes);
    free(lh);
    return ok;
}

typedef struct {
    int log_msg_cols;
    size_t *q;
    size_t nq;
    of128 alpha[16];
    int na;
    int ris_start;
    of128 beta;
} lvlctx;

typedef struct {
    of128 z[64];
    int nz, ris_start;
    of128 beta;
} oodctx;

int o_pcs_verify(const oproof *pf, const opcs_params *pp, const oclaim *claims, int n_claims, ofs *t) {
    const olig_config *cfg = o_lig_config(pp->m, pp->profile);
    if (!cfg) return -100;
    int L = pp->m - 7;
    ofs_label(t, "flock-pcs-open-batch-v0");
    of128 eq_rd[2][128], sclaim[2];
    for (int k = 0; k < n_claims; k++) {
        ofs_label(t, "flock-ring-switch-v0");
        ofs_observe_slice(t, pf->s_hat_v[k], 128);
        /* build_claim_weights (ring_switch.rs:138-160) */
        of128 lam[64], acc = ZERO;
        o_lagrange_s(claims[k].z_skip, lam);
        of128 x0 = claims[k].x_full[0];
        for (int i = 0; i < 128; i++) {
            of128 e = (i >> 6) ? x0 : of128_add(ONE, x0);
            acc = of128_add(acc, of128_mul(of128_mul(lam[i & 63], e), pf->s_hat_v[k][i]));
        }
        if (!of128_eq(acc, claims[k].value)) return -101 - k;
        of128 rd[7], u[128];
        ofs_sample_vec(t, rd, 7);
        o_build_eq(rd, 7, eq_rd[k]);
        transpose128(pf->s_hat_v[k], u);
        sclaim[k] = ZERO;
        for (int b = 0; b < 128; b++) sclaim[k] = of128_add(sclaim[k], of128_mul(u[b], eq_rd[k][b]));
    }
    of128 gam[2], tr = ZERO;
    for (int k = 0; k < n_claims; k++) gam[k] = ofs_sample(t);
    for (int k = 0; k < n_claims; k++) tr = of128_add(tr, of128_mul(gam[k], sclaim[k]));

    /* recursive_verifier_with_basis_succinct (ligerito.rs:6981-7546) */
    const olig_proof *lp = &pf->lig;
    int initial_k = cfg->initial_k, r = cfg->n_levels - 1;
    if (memcmp(lp->initial_root, pf->root, 32)) return -110;
    ofs_label(t, "flock-ligerito-basis-v0");
    ofs_observe_f128(t, tr);
    ofs_observe_bytes(t, lp->initial_root, 32);
    size_t block_len_0 = (size_t)1 << (L - initial_k + cfg->levels[0].log_inv_rate);
    size_t tx = 0, fni = 0, oi = 0, ni = 0;
    if (tx >= lp->n_msgs) return -111;
    ofs_observe_f128(t, lp->msgs[tx][0]);
    ofs_observe_f128(t, lp->msgs[tx][1]);
    quad run = q_from(lp->msgs[tx][0], lp->msgs[tx][1], tr);
    tx++;
    of128 ris[64];
    int nris = 0;
    lvlctx lv[OLIG_MAX_LEVELS + 1];
    int nlv = 0;
    oodctx od[OLIG_MAX_LEVELS + 1];
    int nod = 0;
    for (int j = 0; j < initial_k; j++) {
        int bits = cfg->levels[0].fold_grinding_bits - j;
        if (bits > 0) {
            if (fni >= lp->n_fold_nonces) return -112;
            if (!ofs_verify_pow(t, lp->fold_nonces[fni++], (uint32_t)bits)) return -113;
        }
        of128 rj = ofs_sample(t);
        ris[nris++] = rj;
        tr = q_eval(run, rj);
        if (tx >= lp->n_msgs) return -114;
        ofs_observe_f128(t, lp->msgs[tx][0]);
        ofs_observe_f128(t, lp->msgs[tx][1]);
        run = q_from(lp->msgs[tx][0], lp->msgs[tx][1], tr);
        tx++;
    }
    if (lp->n_rec < 1) return -115;
    ofs_observe_bytes(t, lp->roots[0], 32);
    int n1 = L - initial_k;
    for (int o = 0; o < cfg->levels[1].ood_samples; o++) {
        oodctx *c = &od[nod++];
        c->nz = n1;
        ofs_sample_vec(t, c->z, (size_t)n1);
        if (oi >= lp->n_ood) return -116;
        of128 y = lp->ood[oi++];
        ofs_observe_f128(t, y);
        if (tx >= lp->n_msgs) return -117;
        ofs_observe_f128(t, lp->msgs[tx][0]);
        ofs_observe_f128(t, lp->msgs[tx][1]);
        quad iq = q_from(lp->msgs[tx][0], lp->msgs[tx][1], y);
        tx++;
        c->beta = ofs_sample(t);
        run = q_fold(run, iq, c->beta);
        tr = of128_add(tr, of128_mul(c->beta, y));
        c->ris_start = initial_k;
    }
    if (ni >= lp->n_nonces) return -118;
    if (!ofs_verify_pow(t, lp->nonces[ni++], (uint32_t)cfg->levels[0].grinding_bits)) return -119;
    {
        lvlctx *c = &lv[nlv++];
        c->nq = (size_t)cfg->levels[0].queries;
        c->q = (size_t *)malloc(sizeof(size_t) * c->nq);
        sample_queries(t, block_len_0, c->nq, c->q);
        c->na = ceil_log2(c->nq);
        ofs_sample_vec(t, c->alpha, (size_t)c->na);
        if (!verify_opens(lp->initial_root, block_len_0, c->q, c->nq, &lp->initial, (size_t)1 << initial_k)) return -120;
        of128 es = enforced_sum(&lp->initial, ris, initial_k, c->alpha, c->na);
        if (tx >= lp->n_msgs) return -121;
        ofs_observe_f128(t, lp->msgs[tx][0]);
        ofs_observe_f128(t, lp->msgs[tx][1]);
        quad iq = q_from(lp->msgs[tx][0], lp->msgs[tx][1], es);
        tx++;
        c->beta = ofs_sample(t);
        run = q_fold(run, iq, c->beta);
        tr = of128_add(tr, of128_mul(c->beta, es));
        c->log_msg_cols = n1;
        c->ris_start = initial_k;
    }
    const uint8_t *prev_root = lp->roots[0];
    int prev_k = cfg->levels[1].k_recursive;
    int prev_cols = n1 - prev_k;
    int prev_rate = cfg->levels[1].log_inv_rate;
    int next_root = 1, rec_i = 0, ncur = n1;
    for (int i = 0; i < r; i++) {
        int ki = cfg->levels[i + 1].k_recursive;
        of128 lr[16];
        if (ncur < ki) return -130;
        for (int j = 0; j < ki; j++) {
            int bits = cfg->levels[i + 1].fold_grinding_bits - j;
            if (bits > 0) {
                if (fni >= lp->n_fold_nonces) return -131;
                if (!ofs_verify_pow(t, lp->fold_nonces[fni++], (uint32_t)bits)) return -132;
            }
            of128 rj = ofs_sample(t);
            ris[nris++] = rj;
            lr[j] = rj;
            tr = q_eval(run, rj);
            if (tx >= lp->n_msgs) return -133;
            ofs_observe_f128(t, lp->msgs[tx][0]);
            ofs_observe_f128(t, lp->msgs[tx][1]);
            run = q_from(lp->msgs[tx][0], lp->msgs[tx][1], tr);
            tx++;
        }
        ncur -= ki;
        size_t prev_block = (size_t)1 << (prev_cols + prev_rate);
        size_t prev_lanes = (size_t)1 << prev_k;
        if (i == r - 1) {
            if (tx != lp->n_msgs || oi != lp->n_ood || fni != lp->n_fold_nonces) return -140;
            if (lp->n_yr != ((size_t)1 << ncur)) return -141;
            for (size_t x = 0; x < lp->n_yr; x++) ofs_observe_f128(t, lp->yr[x]);
            if (ni >= lp->n_nonces) return -142;
            if (!ofs_verify_pow(t, lp->nonces[ni], (uint32_t)cfg->levels[i + 1].grinding_bits)) return -143;
            lvlctx *c = &lv[nlv++];
            c->nq = (size_t)cfg->levels[i + 1].queries;
            c->q = (size_t *)malloc(sizeof(size_t) * c->nq);
            sample_queries(t, prev_block, c->nq, c->q);
            c->na = ceil_log2(c->nq);
            ofs_sample_vec(t, c->alpha, (size_t)c->na);
            if (!verify_opens(prev_root, prev_block, c->q, c->nq, &lp->final_open, prev_lanes)) return -144;
            of128 es = enforced_sum(&lp->final_open, lr, ki, c->alpha, c->na);
            c->beta = ofs_sample(t);
            tr = of128_add(tr, of128_mul(c->beta, es));
            c->log_msg_cols = ncur;
            c->ris_start = nris;
            /* residual check: sum_y yr[y] (b(ris, y) + sum beta_k basis_k + ood) */
            int yl = ncur;
            size_t yn = (size_t)1 << yl;
            /* b(ris ++ y): dense MLE of the initial combined basis */
            size_t N = (size_t)1 << L;
            of128 *bb = (of128 *)calloc(N, sizeof(of128));
            of128 *eqs = (of128 *)malloc(sizeof(of128) * N);
            for (int k = 0; k < n_claims; k++) {
                o_build_eq(claims[k].x_full + 1, L, eqs);
                fold_b128_acc(eqs, N, eq_rd[k], gam[k], bb);
            }
            free(eqs);
            size_t cur = N;
            for (int k = 0; k < nris; k++) {
                fold_lsb(bb, cur, ris[k]);
                cur /= 2;
            }
            if (cur != yn) return -145;
            of128 inner = ZERO;
            for (size_t y = 0; y < yn; y++) {
                of128 comb = bb[y];
                for (int kk = 0; kk < nlv; kk++) {
                    of128 p[64];
                    int nfold = lv[kk].log_msg_cols - yl;
                    for (int e = 0; e < nfold; e++) p[e] = ris[lv[kk].ris_start + e];
                    for (int e = 0; e < yl; e++) p[nfold + e] = ((y >> e) & 1) ? ONE : ZERO;
                    of128 v = induced_mle(lv[kk].log_msg_cols, lv[kk].q, lv[kk].nq, lv[kk].alpha, lv[kk].na, p);
                    comb = of128_add(comb, of128_mul(lv[kk].beta, v));
                }
                for (int kk = 0; kk < nod; kk++) {
                    int folded = od[kk].nz - yl;
                    of128 sc = od[kk].beta;
                    for (int e = 0; e < folded; e++)
                        sc = of128_mul(sc, of128_add(of128_add(ONE, od[kk].z[e]), ris[od[kk].ris_start + e]));
                    for (int e = 0; e < yl; e++) {
                        of128 ze = od[kk].z[folded + e];
                        sc = of128_mul(sc, ((y >> e) & 1) ? ze : of128_add(ONE, ze));
                    }
                    comb = of128_add(comb, sc);
                }
                inner = of128_add(inner, of128_mul(lp->yr[y], comb));
            }
            free(bb);
            for (int kk = 0; kk < nlv; kk++) free(lv[kk].q);
            return of128_eq(inner, tr) ? 0 : -150;
        }
        if (next_root >= lp->n_rec) return -160;
        const uint8_t *root_next = lp->roots[next_root++];
        ofs_observe_bytes(t, root_next, 32);
        for (int o = 0; o < cfg->levels[i + 2].ood_samples; o++) {
            oodctx *c = &od[nod++];
            c->nz = ncur;
            ofs_sample_vec(t, c->z, (size_t)ncur);
            if (oi >= lp->n_ood) return -161;
            of128 y = lp->ood[oi++];
            ofs_observe_f128(t, y);
            if (tx >= lp->n_msgs) return -162;
            ofs_observe_f128(t, lp->msgs[tx][0]);
            ofs_observe_f128(t, lp->msgs[tx][1]);
            quad iq = q_from(lp->msgs[tx][0], lp->msgs[tx][1], y);
            tx++;
            c->beta = ofs_sample(t);
            run = q_fold(run, iq, c->beta);
            tr = of128_add(tr, of128_mul(c->beta, y));
            c->ris_start = nris;
        }
        if (ni >= lp->n_nonces) return -163;
        if (!ofs_verify_pow(t, lp->nonces[ni++], (uint32_t)cfg->levels[i + 1].grinding_bits)) return -164;
        lvlctx *c = &lv[nlv++];
        c->nq = (size_t)cfg->levels[i + 1].queries;
        c->q = (size_t *)malloc(sizeof(size_t) * c->nq);
        sample_queries(t, prev_block, c->nq, c->q);
        c->na = ceil_log2(c->nq);
        ofs_sample_vec(t, c->alpha, (size_t)c->na);
        if (rec_i >= lp->n_rec - 1) return -165;
        const orec_proof *rp = &lp->rec[rec_i++];
        if (!verify_opens(prev_root, prev_block, c->q, c->nq, rp, prev_lanes)) return -166;
        of128 es = enforced_sum(rp, lr, ki, c->alpha, c->na);
        if (tx >= lp->n_msgs) return -167;
        ofs_observe_f128(t, lp->msgs[tx][0]);
        ofs_observe_f128(t, lp->msgs[tx][1]);
        quad iq = q_from(lp->msgs[tx][0], lp->msgs[tx][1], es);
        tx++;
        c->beta = ofs_sample(t);
        run = q_fold(run, iq, c->beta);
        tr = of128_add(tr, of128_mul(c->beta, es));
        c->log_msg_cols = ncur;
        c->ris_start = nris;
        prev_root = root_next;
        int kn = cfg->levels[i + 2].k_recursive;
        if (ncur < kn) return -168;
        prev_k = kn;
        prev_cols = ncur - kn;
        prev_rate = cfg->levels[i + 2].log_inv_rate;
    }
    return -199;
}
