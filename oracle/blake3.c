/*
 * Oracle: BLAKE3, restated from the public BLAKE3 specification
 * (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, "BLAKE3: one function, fast
 * everywhere", 2020), hash mode only.  The reference calls the `blake3`
 * crate 1.8.5 (Cargo.lock) through merkle.rs:200-210, challenger.rs:271-355
 * and r1cs.rs:343-365; its own in-source restatement of the compression
 * function is crates/flock-prover/src/r1cs_hashes/blake3.rs:239-310.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <string.h>

#include "oracle.h"

const uint32_t OB3_IV[8] = {0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A,
                            0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19};
static const int PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx;
    s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my;
    s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];
    s[b] = rotr(s[b] ^ s[c], 7);
}

void ob3_compress(const uint32_t cv[8], const uint32_t block[16], uint64_t counter,
                  uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      OB3_IV[0], OB3_IV[1], OB3_IV[2], OB3_IV[3],
                      (uint32_t)counter, (uint32_t)(counter >> 32), block_len, flags};
    uint32_t m[16], t[16];
    memcpy(m, block, sizeof m);
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);
        g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);
        g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);
        g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]);
        g(s, 3, 4, 9, 14, m[14], m[15]);
        for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
        memcpy(m, t, sizeof m);
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

static void words_from_le(const uint8_t *b, uint32_t *w, int n) {
    for (int i = 0; i < n; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}
static void words_to_le(const uint32_t *w, uint8_t *b, int n) {
    for (int i = 0; i < n; i++) {
        b[4 * i] = (uint8_t)w[i];
        b[4 * i + 1] = (uint8_t)(w[i] >> 8);
        b[4 * i + 2] = (uint8_t)(w[i] >> 16);
        b[4 * i + 3] = (uint8_t)(w[i] >> 24);
    }
}

/* An "output" in the spec: the inputs of the last compression of a node. */
typedef struct {
    uint32_t cv[8];
    uint32_t block[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} ob3_output;

static void output_cv(const ob3_output *o, uint32_t cv[8]) {
    uint32_t out[16];
    ob3_compress(o->cv, o->block, o->counter, o->block_len, o->flags, out);
    memcpy(cv, out, 32);
}

static void chunk_reset(ob3_hasher *h, uint64_t counter) {
    memcpy(h->cv, h->key, 32);
    h->chunk_counter = counter;
    memset(h->block, 0, 64);
    h->block_len = 0;
    h->blocks_compressed = 0;
}

static size_t chunk_len(const ob3_hasher *h) { return 64 * (size_t)h->blocks_compressed + h->block_len; }

static uint32_t chunk_start_flag(const ob3_hasher *h) { return h->blocks_compressed == 0 ? OB3_CHUNK_START : 0; }

static void chunk_update(ob3_hasher *h, const uint8_t *in, size_t len) {
    while (len) {
        if (h->block_len == 64) {
            uint32_t bw[16], out[16];
            words_from_le(h->block, bw, 16);
            ob3_compress(h->cv, bw, h->chunk_counter, 64, chunk_start_flag(h), out);
            memcpy(h->cv, out, 32);
            h->blocks_compressed++;
            memset(h->block, 0, 64);
            h->block_len = 0;
        }
        size_t take = 64 - h->block_len;
        if (take > len) take = len;
        memcpy(h->block + h->block_len, in, take);
        h->block_len += (uint32_t)take;
        in += take;
        len -= take;
    }
}

static ob3_output chunk_output(const ob3_hasher *h) {
    ob3_output o;
    memcpy(o.cv, h->cv, 32);
    words_from_le(h->block, o.block, 16);
    o.counter = h->chunk_counter;
    o.block_len = h->block_len;
    o.flags = chunk_start_flag(h) | OB3_CHUNK_END;
    return o;
}

static ob3_output parent_output(const uint32_t l[8], const uint32_t r[8], const uint32_t key[8]) {
    ob3_output o;
    memcpy(o.cv, key, 32);
    memcpy(o.block, l, 32);
    memcpy(o.block + 8, r, 32);
    o.counter = 0;
    o.block_len = 64;
    o.flags = OB3_PARENT;
    return o;
}

void ob3_init(ob3_hasher *h) {
    memset(h, 0, sizeof *h);
    memcpy(h->key, OB3_IV, 32);
    chunk_reset(h, 0);
}

static void add_chunk_cv(ob3_hasher *h, uint32_t cv[8], uint64_t total_chunks) {
    /* Merge completed subtrees: one merge per trailing zero bit of total_chunks. */
    while ((total_chunks & 1) == 0) {
        ob3_output p = parent_output(h->stack[h->stack_len - 1], cv, h->key);
        output_cv(&p, cv);
        h->stack_len--;
        total_chunks >>= 1;
    }
    memcpy(h->stack[h->stack_len++], cv, 32);
}

void ob3_update(ob3_hasher *h, const void *data, size_t len) {
    const uint8_t *in = (const uint8_t *)data;
    while (len) {
        if (chunk_len(h) == 1024) {
            ob3_output o = chunk_output(h);
            uint32_t cv[8];
            output_cv(&o, cv);
            uint64_t total = h->chunk_counter + 1;
            add_chunk_cv(h, cv, total);
            chunk_reset(h, total);
        }
        size_t want = 1024 - chunk_len(h);
        if (want > len) want = len;
        chunk_update(h, in, want);
        in += want;
        len -= want;
    }
}

static ob3_output final_output(const ob3_hasher *h) {
    ob3_output o = chunk_output(h);
    for (int i = (int)h->stack_len - 1; i >= 0; i--) {
        uint32_t cv[8];
        output_cv(&o, cv);
        o = parent_output(h->stack[i], cv, h->key);
    }
    return o;
}

void ob3_finalize_xof(const ob3_hasher *h, uint8_t *out, size_t out_len) {
    ob3_output o = final_output(h);
    uint64_t blk = 0;
    while (out_len) {
        uint32_t w[16];
        uint8_t b[64];
        ob3_compress(o.cv, o.block, blk++, o.block_len, o.flags | OB3_ROOT, w);
        words_to_le(w, b, 16);
        size_t take = out_len < 64 ? out_len : 64;
        memcpy(out, b, take);
        out += take;
        out_len -= take;
    }
}

void ob3_finalize_non_root(const ob3_hasher *h, uint8_t out[32]) {
    ob3_output o = final_output(h);
    uint32_t cv[8];
    output_cv(&o, cv);
    words_to_le(cv, out, 8);
}

void ob3_hash(const void *data, size_t len, uint8_t out[32]) {
    ob3_hasher h;
    ob3_init(&h);
    ob3_update(&h, data, len);
    ob3_finalize_xof(&h, out, 32);
}

void ob3_leaf_cv(const void *data, size_t len, uint8_t out[32]) {
    ob3_hasher h;
    ob3_init(&h);
    ob3_update(&h, data, len);
    ob3_finalize_non_root(&h, out);
}

void ob3_parent_cv(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]) {
    uint32_t l[8], r[8], cv[8];
    words_from_le(left, l, 8);
    words_from_le(right, r, 8);
    ob3_output o = parent_output(l, r, OB3_IV);
    output_cv(&o, cv);
    words_to_le(cv, out, 8);
}
