/*
 * Oracle: additive NTT over GF(2^128) in the LCH novel polynomial basis with
 * the standard subspace basis {1, x, ..., x^(d-1)}.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>

#include "oracle.h"

/* Normalized subspace-polynomial table (ntt/additive_ntt_f128.rs:57-89):
 * evals[i][k] = W_i(beta_{i+k}) / W_i(beta_i), with
 * W_i(z) = W_{i-1}(z) (W_{i-1}(z) + W_{i-1}(beta_{i-1})). */
static of128 **build_evals(int l) {
    of128 **ev = (of128 **)calloc((size_t)l, sizeof(of128 *));
    ev[0] = (of128 *)malloc(sizeof(of128) * (size_t)l);
    for (int i = 0; i < l; i++) {
        ev[0][i].lo = i < 64 ? (1ULL << i) : 0;
        ev[0][i].hi = i < 64 ? 0 : (1ULL << (i - 64));
    }
    for (int i = 1; i < l; i++) {
        int len = l - i;
        ev[i] = (of128 *)malloc(sizeof(of128) * (size_t)len);
        for (int k = 1; k <= len; k++) {
            of128 p = ev[i - 1][k];
            ev[i][k - 1] = of128_mul(p, of128_add(p, ev[i - 1][0]));
        }
    }
    for (int i = 0; i < l; i++) {
        of128 inv = of128_inv(ev[i][0]);
        for (int k = 0; k < l - i; k++) ev[i][k] = of128_mul(ev[i][k], inv);
    }
    return ev;
}

/* Breadth-first twiddle table (ntt/additive_ntt_f128.rs:106-128):
 * layer l, block b -> twiddle = span of evals[d-l-1][1..] indexed by b. */
int ontt_init(ontt *t, int dim) {
    t->log_d = dim;
    t->twiddles = (of128 *)malloc(sizeof(of128) * (((size_t)1 << dim) > 1 ? ((size_t)1 << dim) - 1 : 1));
    if (!t->twiddles) return -1;
    if (dim == 0) return 0;
    of128 **ev = build_evals(dim);
    for (int layer = 0; layer < dim; layer++) {
        size_t start = ((size_t)1 << layer) - 1;
        const of128 *row = ev[dim - layer - 1];
        t->twiddles[start].lo = 0;
        t->twiddles[start].hi = 0;
        for (int bit = 0; bit < layer; bit++) {
            size_t half = (size_t)1 << bit;
            for (size_t b = 0; b < half; b++) t->twiddles[start + half + b] = of128_add(t->twiddles[start + b], row[1 + bit]);
        }
    }
    for (int i = 0; i < dim; i++) free(ev[i]);
    free(ev);
    return 0;
}

void ontt_free(ontt *t) {
    free(t->twiddles);
    t->twiddles = NULL;
}

of128 ontt_twiddle(const ontt *t, int layer, size_t block) { return t->twiddles[((size_t)1 << layer) - 1 + block]; }

/* Forward interleaved transform, neighbors-last, from start_layer
 * (ntt/additive_ntt_f128.rs:1281-1320): u += tw*v; v += u. */
void ontt_forward_interleaved(const ontt *t, of128 *data, size_t n_total, size_t num_ntts, int start_layer) {
    size_t npos = n_total / num_ntts;
    int log_d = 0;
    while (((size_t)1 << log_d) < npos) log_d++;
    for (int layer = start_layer; layer < log_d; layer++) {
        size_t nblocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t blk = 0; blk < nblocks; blk++) {
            of128 tw = ontt_twiddle(t, layer, blk);
            size_t base = blk * 2 * half * num_ntts;
            for (size_t row = 0; row < half; row++) {
                of128 *top = data + base + row * num_ntts;
                of128 *bot = top + half * num_ntts;
                for (size_t lane = 0; lane < num_ntts; lane++) {
                    of128 u = of128_add(top[lane], of128_mul(bot[lane], tw));
                    top[lane] = u;
                    bot[lane] = of128_add(bot[lane], u);
                }
            }
        }
    }
}

/* Inverse (ntt/additive_ntt_f128.rs:2122-2145): v += u; u += v*tw. */
void ontt_inverse(const ontt *t, of128 *data, size_t n) {
    int log_d = 0;
    while (((size_t)1 << log_d) < n) log_d++;
    for (int layer = log_d - 1; layer >= 0; layer--) {
        size_t nblocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t blk = 0; blk < nblocks; blk++) {
            of128 tw = ontt_twiddle(t, layer, blk);
            size_t start = blk << (log_d - layer);
            for (size_t i0 = start; i0 < start + half; i0++) {
                size_t i1 = i0 | half;
                of128 u = data[i0];
                of128 nv = of128_add(data[i1], u);
                data[i1] = nv;
                data[i0] = of128_add(u, of128_mul(nv, tw));
            }
        }
    }
}
