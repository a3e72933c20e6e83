/*
 * Oracle: the BLAKE3 compression R1CS (A_0, B_0, C_0 = I), its statement
 * digest, and the witness generator (z, a = A z, b = B z), F128-packed.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_proto.h"

/* Slot layout (crates/flock-prover/src/r1cs_hashes/blake3.rs:178-189). */
enum {
    WB = 32,
    CV_BASE = 0,
    OUT_LO_BASE = 256,
    Z_CONST = 512,
    M_BASE = 513,
    T_LO_BASE = 1025,
    T_HI_BASE = 1057,
    BLEN_BASE = 1089,
    FLAGS_BASE = 1121,
    GS_BASE = 1153,
    G_STRIDE = 250,
    OUT_HI_BASE = 15153,
};
static int g_carry(int g, int add, int b) { return GS_BASE + G_STRIDE * g + 31 * add + b; }
static int g_lin(int g, int which, int b) { return GS_BASE + G_STRIDE * g + 186 + WB * which + b; }

static const int G_LANES[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                  {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};
static const int MSG_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

/* Message word index read by G `g` of round r for operand mx (0) or my (1):
 * PERM^r applied to the per-round schedule (blake3.rs:312-345). */
static int msg_index(int r, int g, int which) {
    int perm[16], nxt[16];
    for (int i = 0; i < 16; i++) perm[i] = i;
    for (int k = 0; k < r; k++) {
        for (int i = 0; i < 16; i++) nxt[i] = perm[MSG_PERM[i]];
        memcpy(perm, nxt, sizeof perm);
    }
    return perm[2 * g + which];
}

/* ------------------------------------------------------------ symbolic words
 * A Word is 32 linear forms over the K witness slots, each a K-bit set
 * (GF(2) coefficients).  XOR of sets is exactly the reference's xor_dedup of
 * concatenated index lists (common.rs:768-784), so every row comes out as the
 * same sorted, square-free index list. */
#define KW (OK_K / 64)
typedef struct { uint64_t b[WB][KW]; } word_t;

static word_t *w_new(void) { return (word_t *)calloc(1, sizeof(word_t)); }
static void w_slot(word_t *w, int base) {
    memset(w, 0, sizeof *w);
    for (int i = 0; i < WB; i++) w->b[i][(base + i) >> 6] |= 1ULL << ((base + i) & 63);
}
static void w_const(word_t *w, uint32_t v) {
    memset(w, 0, sizeof *w);
    for (int i = 0; i < WB; i++)
        if ((v >> i) & 1) w->b[i][Z_CONST >> 6] |= 1ULL << (Z_CONST & 63);
}
static void w_xor(word_t *out, const word_t *x, const word_t *y) {
    for (int i = 0; i < WB; i++)
        for (int k = 0; k < KW; k++) out->b[i][k] = x->b[i][k] ^ y->b[i][k];
}
static void w_rotr(word_t *out, const word_t *x, int n) {
    static word_t tmp;
    for (int i = 0; i < WB; i++) memcpy(tmp.b[i], x->b[(i + n) % WB], sizeof tmp.b[i]);
    memcpy(out, &tmp, sizeof tmp);
}
static void set_toggle(uint64_t *s, int idx) { s[idx >> 6] ^= 1ULL << (idx & 63); }

/* One ADD x + y with carry_aux slots at carry_base (blake3.rs:430-447):
 * A row = x[i] ^ cin[i], B row = y[i] ^ cin[i] with cin[i] = xor_{j<i} carry[j];
 * returns the sum form x[i] ^ y[i] ^ cin[i]. */
static void add_rows(uint64_t *A, uint64_t *B, const word_t *x, const word_t *y, int cb, word_t *sum) {
    static uint64_t cin[KW];
    memset(cin, 0, sizeof cin);
    for (int i = 0; i < WB; i++) {
        if (i < 31) {
            uint64_t *ar = A + (size_t)(cb + i) * KW, *br = B + (size_t)(cb + i) * KW;
            for (int k = 0; k < KW; k++) {
                ar[k] = x->b[i][k] ^ cin[k];
                br[k] = y->b[i][k] ^ cin[k];
            }
        }
        for (int k = 0; k < KW; k++) sum->b[i][k] = x->b[i][k] ^ y->b[i][k] ^ cin[k];
        if (i < 31) set_toggle(cin, cb + i);
    }
}

static void lin_rows(uint64_t *A, uint64_t *B, int base, const word_t *v) {
    for (int i = 0; i < WB; i++) {
        memcpy(A + (size_t)(base + i) * KW, v->b[i], KW * 8);
        memset(B + (size_t)(base + i) * KW, 0, KW * 8);
        set_toggle(B + (size_t)(base + i) * KW, Z_CONST);
    }
}

/* build_matrices (blake3.rs:485-633): rows as K-bit sets, row-major. */
static void build_rowsets(uint64_t *A, uint64_t *B) {
    memset(A, 0, (size_t)OK_K * KW * 8);
    memset(B, 0, (size_t)OK_K * KW * 8);
    set_toggle(A + (size_t)Z_CONST * KW, Z_CONST);
    set_toggle(B + (size_t)Z_CONST * KW, Z_CONST);
    const int in_base[6] = {CV_BASE, M_BASE, T_LO_BASE, T_HI_BASE, BLEN_BASE, FLAGS_BASE};
    const int in_len[6] = {256, 512, 32, 32, 32, 32};
    for (int r = 0; r < 6; r++)
        for (int j = 0; j < in_len[r]; j++) {
            int s = in_base[r] + j;
            set_toggle(A + (size_t)s * KW, s);
            set_toggle(B + (size_t)s * KW, Z_CONST);
        }
    word_t *st[16];
    for (int i = 0; i < 16; i++) st[i] = w_new();
    for (int w = 0; w < 8; w++) w_slot(st[w], CV_BASE + WB * w);
    for (int i = 0; i < 4; i++) w_const(st[8 + i], OB3_IV[i]);
    w_slot(st[12], T_LO_BASE);
    w_slot(st[13], T_HI_BASE);
    w_slot(st[14], BLEN_BASE);
    w_slot(st[15], FLAGS_BASE);
    word_t *a = w_new(), *b = w_new(), *c = w_new(), *d = w_new(), *mx = w_new(), *my = w_new();
    word_t *tmp0 = w_new(), *a1 = w_new(), *d1 = w_new(), *c1 = w_new(), *b1 = w_new(), *tmp1 = w_new();
    word_t *a2 = w_new(), *d2 = w_new(), *c2 = w_new(), *bn = w_new(), *t = w_new();
    for (int r = 0; r < 7; r++) {
        for (int gi = 0; gi < 8; gi++) {
            int g = r * 8 + gi;
            int la = G_LANES[gi][0], lb = G_LANES[gi][1], lc = G_LANES[gi][2], ld = G_LANES[gi][3];
            memcpy(a, st[la], sizeof *a);
            memcpy(b, st[lb], sizeof *b);
            memcpy(c, st[lc], sizeof *c);
            memcpy(d, st[ld], sizeof *d);
            w_slot(mx, M_BASE + WB * msg_index(r, gi, 0));
            w_slot(my, M_BASE + WB * msg_index(r, gi, 1));
            add_rows(A, B, a, b, g_carry(g, 0, 0), tmp0);
            add_rows(A, B, tmp0, mx, g_carry(g, 1, 0), a1);
            w_xor(t, d, a1);
            w_rotr(d1, t, 16);
            add_rows(A, B, c, d1, g_carry(g, 2, 0), c1);
            w_xor(t, b, c1);
            w_rotr(b1, t, 12);
            add_rows(A, B, a1, b1, g_carry(g, 3, 0), tmp1);
            add_rows(A, B, tmp1, my, g_carry(g, 4, 0), a2);
            w_xor(t, d1, a2);
            w_rotr(d2, t, 8);
            add_rows(A, B, c1, d2, g_carry(g, 5, 0), c2);
            w_xor(t, b1, c2);
            w_rotr(bn, t, 7);
            lin_rows(A, B, g_lin(g, 0, 0), bn);
            lin_rows(A, B, g_lin(g, 1, 0), d2);
            memcpy(st[la], a2, sizeof *a2);
            w_slot(st[lb], g_lin(g, 0, 0));
            memcpy(st[lc], c2, sizeof *c2);
            w_slot(st[ld], g_lin(g, 1, 0));
        }
    }
    for (int w = 0; w < 8; w++) {
        w_xor(t, st[w], st[w + 8]);
        lin_rows(A, B, OUT_LO_BASE + WB * w, t);
        word_t *cvw = w_new();
        w_slot(cvw, CV_BASE + WB * w);
        w_xor(t, st[w + 8], cvw);
        lin_rows(A, B, OUT_HI_BASE + WB * w, t);
        free(cvw);
    }
    for (int i = 0; i < 16; i++) free(st[i]);
    free(a); free(b); free(c); free(d); free(mx); free(my); free(tmp0); free(a1); free(d1);
    free(c1); free(b1); free(tmp1); free(a2); free(d2); free(c2); free(bn); free(t);
}

static void rowsets_to_csr(const uint64_t *S, osparse *out) {
    size_t nnz = 0;
    for (size_t i = 0; i < (size_t)OK_K * KW; i++) nnz += (size_t)__builtin_popcountll(S[i]);
    out->row_ptr = (uint32_t *)malloc(sizeof(uint32_t) * (OK_K + 1));
    out->cols = (uint32_t *)malloc(sizeof(uint32_t) * (nnz ? nnz : 1));
    out->nnz = nnz;
    size_t p = 0;
    for (int r = 0; r < OK_K; r++) {
        out->row_ptr[r] = (uint32_t)p;
        const uint64_t *row = S + (size_t)r * KW;
        for (int k = 0; k < KW; k++) {
            uint64_t v = row[k];
            while (v) {
                int bit = __builtin_ctzll(v);
                out->cols[p++] = (uint32_t)(k * 64 + bit);
                v &= v - 1;
            }
        }
    }
    out->row_ptr[OK_K] = (uint32_t)p;
}

static osparse G_A, G_B;
static int g_built = 0;

void o_r1cs_matrices(const osparse **a0, const osparse **b0) {
    if (!g_built) {
        uint64_t *A = (uint64_t *)malloc((size_t)OK_K * KW * 8);
        uint64_t *B = (uint64_t *)malloc((size_t)OK_K * KW * 8);
        build_rowsets(A, B);
        rowsets_to_csr(A, &G_A);
        rowsets_to_csr(B, &G_B);
        free(A);
        free(B);
        g_built = 1;
    }
    *a0 = &G_A;
    *b0 = &G_B;
}

/* statement_digest (crates/flock-core/src/r1cs.rs:343-365, 431-460):
 * "flock-r1cs-stmt-v1" | u64 m | u64 k_log | u64 k_skip | u8 layout(0) |
 * matrix(A_0) | matrix(B_0) | matrix(C_0 = I_K); matrix = u64 rows | u64 cols
 * | per row (u64 len | u64 col ...). */
static void absorb_u64(ob3_hasher *h, uint64_t v) { ob3_update(h, &v, 8); }

static void absorb_csr(ob3_hasher *h, const osparse *s) {
    absorb_u64(h, OK_K);
    absorb_u64(h, OK_K);
    uint64_t buf[4096];
    for (int r = 0; r < OK_K; r++) {
        uint32_t b = s->row_ptr[r], e = s->row_ptr[r + 1];
        absorb_u64(h, e - b);
        while (b < e) {
            uint32_t n = e - b > 4096 ? 4096 : e - b;
            for (uint32_t i = 0; i < n; i++) buf[i] = s->cols[b + i];
            ob3_update(h, buf, 8 * n);
            b += n;
        }
    }
}

void o_statement_digest(int m, uint8_t out[32]) {
    const osparse *A, *B;
    o_r1cs_matrices(&A, &B);
    ob3_hasher h;
    ob3_init(&h);
    ob3_update(&h, "flock-r1cs-stmt-v1", 18);
    absorb_u64(&h, (uint64_t)m);
    absorb_u64(&h, 14);
    absorb_u64(&h, 6);
    uint8_t layout = 0;
    ob3_update(&h, &layout, 1);
    absorb_csr(&h, A);
    absorb_csr(&h, B);
    absorb_u64(&h, OK_K);
    absorb_u64(&h, OK_K);
    for (uint64_t i = 0; i < OK_K; i++) {
        absorb_u64(&h, 1);
        absorb_u64(&h, i);
    }
    ob3_finalize_xof(&h, out, 32);
}

/* ------------------------------------------------------------ witness
 * build_block_witness (blake3.rs:989-1078) for z; the row witnesses a, b per
 * the row semantics of build_matrices (blake3.rs:1116-1125): carry rows get
 * (X ^ cin, Y ^ cin), input / lin-id rows (value, 1), the constant (1, 1),
 * padding rows 0.  Written into one 2 KiB block of each packed buffer. */
static void put_bits(uint64_t *w, int base, uint32_t val, int nbits) {
    for (int i = 0; i < nbits; i++)
        if ((val >> i) & 1) w[(base + i) >> 6] |= 1ULL << ((base + i) & 63);
}

static uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static uint32_t add_rec(uint32_t x, uint32_t y, int cb, uint64_t *z, uint64_t *a, uint64_t *b) {
    uint32_t sum = x + y;
    uint32_t cin = sum ^ x ^ y;
    uint32_t left = x ^ cin, right = y ^ cin;
    put_bits(z, cb, left & right, 31);
    put_bits(a, cb, left, 31);
    put_bits(b, cb, right, 31);
    return sum;
}

static void lin_rec(int base, uint32_t v, uint64_t *z, uint64_t *a, uint64_t *b) {
    put_bits(z, base, v, 32);
    put_bits(a, base, v, 32);
    put_bits(b, base, 0xFFFFFFFFu, 32);
}

void o_block_witness(const ocompression *c, uint64_t *z, uint64_t *a, uint64_t *b) {
    memset(z, 0, OK_K / 8);
    memset(a, 0, OK_K / 8);
    memset(b, 0, OK_K / 8);
    put_bits(z, Z_CONST, 1, 1);
    put_bits(a, Z_CONST, 1, 1);
    put_bits(b, Z_CONST, 1, 1);
    for (int w = 0; w < 8; w++) lin_rec(CV_BASE + WB * w, c->cv[w], z, a, b);
    for (int i = 0; i < 16; i++) lin_rec(M_BASE + WB * i, c->m[i], z, a, b);
    uint32_t tlo = (uint32_t)c->counter, thi = (uint32_t)(c->counter >> 32);
    lin_rec(T_LO_BASE, tlo, z, a, b);
    lin_rec(T_HI_BASE, thi, z, a, b);
    lin_rec(BLEN_BASE, c->block_len, z, a, b);
    lin_rec(FLAGS_BASE, c->flags, z, a, b);
    uint32_t s[16] = {c->cv[0], c->cv[1], c->cv[2], c->cv[3], c->cv[4], c->cv[5], c->cv[6], c->cv[7],
                      OB3_IV[0], OB3_IV[1], OB3_IV[2], OB3_IV[3], tlo, thi, c->block_len, c->flags};
    for (int r = 0; r < 7; r++)
        for (int gi = 0; gi < 8; gi++) {
            int g = r * 8 + gi;
            int la = G_LANES[gi][0], lb = G_LANES[gi][1], lc = G_LANES[gi][2], ld = G_LANES[gi][3];
            uint32_t mx = c->m[msg_index(r, gi, 0)], my = c->m[msg_index(r, gi, 1)];
            uint32_t A0 = s[la], B0 = s[lb], C0 = s[lc], D0 = s[ld];
            uint32_t t0 = add_rec(A0, B0, g_carry(g, 0, 0), z, a, b);
            uint32_t a1 = add_rec(t0, mx, g_carry(g, 1, 0), z, a, b);
            uint32_t d1 = rotr32(D0 ^ a1, 16);
            uint32_t c1 = add_rec(C0, d1, g_carry(g, 2, 0), z, a, b);
            uint32_t b1 = rotr32(B0 ^ c1, 12);
            uint32_t t1 = add_rec(a1, b1, g_carry(g, 3, 0), z, a, b);
            uint32_t a2 = add_rec(t1, my, g_carry(g, 4, 0), z, a, b);
            uint32_t d2 = rotr32(d1 ^ a2, 8);
            uint32_t c2 = add_rec(c1, d2, g_carry(g, 5, 0), z, a, b);
            uint32_t bn = rotr32(b1 ^ c2, 7);
            lin_rec(g_lin(g, 0, 0), bn, z, a, b);
            lin_rec(g_lin(g, 1, 0), d2, z, a, b);
            s[la] = a2;
            s[lb] = bn;
            s[lc] = c2;
            s[ld] = d2;
        }
    for (int w = 0; w < 8; w++) {
        lin_rec(OUT_LO_BASE + WB * w, s[w] ^ s[w + 8], z, a, b);
        lin_rec(OUT_HI_BASE + WB * w, s[w + 8] ^ c->cv[w], z, a, b);
    }
}

/* generate_witness_with_ab_packed (blake3.rs:2928-2985): n blocks, padded to
 * 2^n_log with the compression of the all-zero input. */
void o_witness_packed(const ocompression *blocks, size_t n, int n_log, uint64_t *z, uint64_t *a, uint64_t *b) {
    size_t total = (size_t)1 << n_log;
    ocompression zero;
    memset(&zero, 0, sizeof zero);
    const size_t wpb = OK_K / 64;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < total; i++)
        o_block_witness(i < n ? &blocks[i] : &zero, z + i * wpb, a + i * wpb, b + i * wpb);
}

/* Row parity of a sparse matrix against one block of z (for A z, B z checks). */
void o_apply_block(const osparse *s, const uint64_t *zblk, uint64_t *out) {
    memset(out, 0, OK_K / 8);
    for (int r = 0; r < OK_K; r++) {
        uint32_t p = 0;
        for (uint32_t k = s->row_ptr[r]; k < s->row_ptr[r + 1]; k++) p ^= (uint32_t)(zblk[s->cols[k] >> 6] >> (s->cols[k] & 63)) & 1u;
        if (p) out[r >> 6] |= 1ULL << (r & 63);
    }
}

int o_min_n_blocks_log(size_t n_blocks) {
    size_t n = n_blocks < 8 ? 8 : n_blocks;
    int l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}
