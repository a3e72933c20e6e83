/*
 * Oracle: flat BLAKE3 Merkle tree, paths and multi-proofs.
 * Layout (merkle.rs:15-21): tree[0..n) = leaves, each level follows, root last.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* merkle_tree_sequential (merkle.rs:781-803) with BLAKE3 leaf/parent CVs. */
void omerkle_tree(const uint8_t *data, size_t data_len, size_t num_leaves, uint8_t *tree) {
    size_t leaf = data_len / num_leaves;
    for (size_t i = 0; i < num_leaves; i++) ob3_leaf_cv(data + i * leaf, leaf, tree + 32 * i);
    size_t rs = 0, rl = num_leaves;
    while (rl > 1) {
        size_t nl = rl >> 1;
        for (size_t i = 0; i < nl; i++)
            ob3_parent_cv(tree + 32 * (rs + 2 * i), tree + 32 * (rs + 2 * i + 1), tree + 32 * (rs + rl + i));
        rs += rl;
        rl = nl;
    }
}

static int cmp_size(const void *a, const void *b) {
    size_t x = *(const size_t *)a, y = *(const size_t *)b;
    return x < y ? -1 : x > y;
}

/* merkle_multi_proof (merkle.rs:879-925): sort + dedup, then bottom-up emit
 * each sibling not itself active. */
size_t omerkle_multi_proof(const uint8_t *tree, size_t num_leaves, const size_t *positions, size_t n_pos,
                           uint8_t *proof_out, size_t proof_cap) {
    if (n_pos == 0 || num_leaves == 1) return 0;
    size_t *act = (size_t *)malloc(sizeof(size_t) * n_pos);
    memcpy(act, positions, sizeof(size_t) * n_pos);
    qsort(act, n_pos, sizeof(size_t), cmp_size);
    size_t n = 0;
    for (size_t i = 0; i < n_pos; i++)
        if (n == 0 || act[n - 1] != act[i]) act[n++] = act[i];
    size_t out = 0, ls = 0, ll = num_leaves;
    while (ll > 1) {
        size_t nn = 0;
        for (size_t i = 0; i < n;) {
            size_t p = act[i];
            if (i + 1 < n && act[i + 1] == (p ^ 1)) {
                i += 2;
            } else {
                if (out < proof_cap) memcpy(proof_out + 32 * out, tree + 32 * (ls + (p ^ 1)), 32);
                out++;
                i += 1;
            }
            act[nn++] = p >> 1;
        }
        n = nn;
        ls += ll;
        ll >>= 1;
    }
    free(act);
    return out;
}

/* verify_merkle_multi_proof (merkle.rs:930-1005). */
int omerkle_verify_multi(const uint8_t root[32], size_t num_leaves, const size_t *pos, const uint8_t *leaf_hashes,
                         size_t n, const uint8_t *proof, size_t n_proof) {
    if (num_leaves == 0 || (num_leaves & (num_leaves - 1))) return 0;
    if (n == 0) return n_proof == 0;
    for (size_t i = 0; i < n; i++) {
        if (pos[i] >= num_leaves) return 0;
        if (i > 0 && pos[i - 1] >= pos[i]) return 0;
    }
    if (num_leaves == 1) return n_proof == 0 && memcmp(leaf_hashes, root, 32) == 0;
    size_t *p = (size_t *)malloc(sizeof(size_t) * n);
    uint8_t *h = (uint8_t *)malloc(32 * n);
    memcpy(p, pos, sizeof(size_t) * n);
    memcpy(h, leaf_hashes, 32 * n);
    size_t used = 0, ll = num_leaves;
    int ok = 1;
    while (ll > 1 && ok) {
        size_t nn = 0;
        for (size_t i = 0; i < n;) {
            uint8_t l[32], r[32];
            size_t q = p[i];
            if (i + 1 < n && p[i + 1] == (q ^ 1)) {
                memcpy(l, h + 32 * i, 32);
                memcpy(r, h + 32 * (i + 1), 32);
                i += 2;
            } else {
                if (used >= n_proof) {
                    ok = 0;
                    break;
                }
                if (q & 1) {
                    memcpy(l, proof + 32 * used, 32);
                    memcpy(r, h + 32 * i, 32);
                } else {
                    memcpy(l, h + 32 * i, 32);
                    memcpy(r, proof + 32 * used, 32);
                }
                used++;
                i += 1;
            }
            p[nn] = q >> 1;
            ob3_parent_cv(l, r, h + 32 * nn);
            nn++;
        }
        n = nn;
        ll >>= 1;
    }
    if (ok) ok = used == n_proof && n == 1 && memcmp(h, root, 32) == 0;
    free(p);
    free(h);
    return ok;
}
