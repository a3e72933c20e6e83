/*
 * Oracle: Fiat-Shamir transcript over BLAKE3 and the BLAKE3 proof-of-work.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <string.h>

#include "oracle_proto.h"

enum { OP_DOMAIN = 1, OP_LABEL = 2, OP_OBSERVE = 3, OP_SQUEEZE = 4, OP_BYTES = 5, K_SCALAR = 1, K_SLICE = 2 };

static void absorb(ofs *t, const void *p, size_t n) { ob3_update(&t->h, p, n); }
static void absorb_u8(ofs *t, uint8_t v) { absorb(t, &v, 1); }
static void absorb_u64(ofs *t, uint64_t v) { absorb(t, &v, 8); }

/* FsChallenger::with_hash (challenger.rs:271-283): [01] | u64 len | domain. */
void ofs_init(ofs *t, const void *domain, size_t len) {
    ob3_init(&t->h);
    absorb_u8(t, OP_DOMAIN);
    absorb_u64(t, len);
    absorb(t, domain, len);
}

/* challenger.rs:362-366 */
void ofs_label(ofs *t, const char *label) {
    size_t n = strlen(label);
    absorb_u8(t, OP_LABEL);
    absorb_u64(t, n);
    absorb(t, label, n);
}

/* challenger.rs:372-375: [03 01] | lo | hi */
void ofs_observe_f128(ofs *t, of128 v) {
    uint8_t tag[2] = {OP_OBSERVE, K_SCALAR};
    absorb(t, tag, 2);
    absorb_u64(t, v.lo);
    absorb_u64(t, v.hi);
}

/* challenger.rs:377-387: [03 02] | u64 n | 16n bytes */
void ofs_observe_slice(ofs *t, const of128 *v, size_t n) {
    uint8_t tag[2] = {OP_OBSERVE, K_SLICE};
    absorb(t, tag, 2);
    absorb_u64(t, n);
    absorb(t, v, 16 * n);
}

/* challenger.rs:389-393: [05] | u64 n | bytes */
void ofs_observe_bytes(ofs *t, const void *b, size_t n) {
    absorb_u8(t, OP_BYTES);
    absorb_u64(t, n);
    absorb(t, b, n);
}

/* challenger.rs:395-406: [04 01], squeeze 16 from a clone, re-absorb them. */
of128 ofs_sample(ofs *t) {
    uint8_t tag[2] = {OP_SQUEEZE, K_SCALAR};
    absorb(t, tag, 2);
    uint8_t buf[16];
    ob3_finalize_xof(&t->h, buf, 16);
    absorb(t, buf, 16);
    of128 v;
    memcpy(&v.lo, buf, 8);
    memcpy(&v.hi, buf + 8, 8);
    return v;
}

/* challenger.rs:408-436: [04 02] | u64 n, squeeze 16n, re-absorb. */
void ofs_sample_vec(ofs *t, of128 *out, size_t n) {
    uint8_t tag[2] = {OP_SQUEEZE, K_SLICE};
    absorb(t, tag, 2);
    absorb_u64(t, n);
    ob3_finalize_xof(&t->h, (uint8_t *)out, 16 * n);
    absorb(t, out, 16 * n);
}

/* PoW predicate (challenger.rs:1246-1300): BLAKE3(digest | nonce_le | 0^24)
 * has `bits` leading zero bits (big-endian bit order within bytes). */
int o_pow_ok(const uint8_t digest[32], uint64_t nonce, uint32_t bits) {
    uint8_t pre[64];
    memset(pre, 0, sizeof pre);
    memcpy(pre, digest, 32);
    memcpy(pre + 32, &nonce, 8);
    uint8_t h[32];
    ob3_hash(pre, 64, h);
    uint32_t full = bits / 8, extra = bits % 8;
    for (uint32_t i = 0; i < full; i++)
        if (h[i]) return 0;
    if (extra && (h[full] >> (8 - extra))) return 0;
    return 1;
}

static void state_digest(const ofs *t, uint8_t d[32]) { ob3_finalize_xof(&t->h, d, 32); }

/* grind_pow (challenger.rs:440-705): smallest nonce; bits = 0 -> nonce 0;
 * the nonce is then absorbed with observe_bytes. */
uint64_t ofs_grind_pow(ofs *t, uint32_t bits) {
    uint64_t nonce = 0;
    if (bits) {
        uint8_t d[32];
        state_digest(t, d);
        while (!o_pow_ok(d, nonce, bits)) nonce++;
    }
    ofs_observe_bytes(t, &nonce, 8);
    return nonce;
}

/* verify_pow (challenger.rs:706-730) */
int ofs_verify_pow(ofs *t, uint64_t nonce, uint32_t bits) {
    int ok;
    if (bits == 0) {
        ok = nonce == 0;
    } else {
        uint8_t d[32];
        state_digest(t, d);
        ok = o_pow_ok(d, nonce, bits);
    }
    ofs_observe_bytes(t, &nonce, 8);
    return ok;
}
