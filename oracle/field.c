/*
 * Oracle: GF(2^128) (GHASH polynomial basis) and GF(2^8) (AES polynomial).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

#if defined(__PCLMUL__)
#include <wmmintrin.h>
#include <smmintrin.h>
#endif

/* 64x64 carry-less product, bit by bit (field/gf2_128/portable.rs:4-20). */
static void clmul64(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
    uint64_t l = 0, h = 0;
    for (int i = 0; i < 64; i++) {
        if ((a >> i) & 1) {
            l ^= b << i;
            if (i) h ^= b >> (64 - i);
        }
    }
    *lo = l;
    *hi = h;
}

#if defined(__PCLMUL__)
static void clmul64_hw(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
    __m128i x = _mm_set_epi64x(0, (long long)a), y = _mm_set_epi64x(0, (long long)b);
    __m128i r = _mm_clmulepi64_si128(x, y, 0x00);
    *lo = (uint64_t)_mm_cvtsi128_si64(r);
    *hi = (uint64_t)_mm_extract_epi64(r, 1);
}
#define CLMUL clmul64_hw
#else
#define CLMUL clmul64
#endif

/* Schoolbook 256-bit product (field/gf2_128/portable.rs:22-36). */
static of256 mul_unreduced_with(of128 a, of128 b, void (*cl)(uint64_t, uint64_t, uint64_t *, uint64_t *)) {
    uint64_t ll0, ll1, lh0, lh1, hl0, hl1, hh0, hh1;
    cl(a.lo, b.lo, &ll0, &ll1);
    cl(a.lo, b.hi, &lh0, &lh1);
    cl(a.hi, b.lo, &hl0, &hl1);
    cl(a.hi, b.hi, &hh0, &hh1);
    of256 u;
    u.r0 = ll0;
    u.r1 = ll1 ^ lh0 ^ hl0;
    u.r2 = hh0 ^ lh1 ^ hl1;
    u.r3 = hh1;
    return u;
}

of256 of128_mul_unreduced(of128 a, of128 b) { return mul_unreduced_with(a, b, CLMUL); }

/* Fold bits 128..255 back using x^128 = x^7 + x^2 + x + 1 (field/gf2_128.rs:192-214). */
of128 of256_reduce(of256 u) {
    uint64_t r2 = u.r2, r3 = u.r3;
    uint64_t t_lo = r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
    uint64_t t_hi = r3 ^ ((r3 << 1) | (r2 >> 63)) ^ ((r3 << 2) | (r2 >> 62)) ^ ((r3 << 7) | (r2 >> 57));
    uint64_t ov = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
    uint64_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
    of128 r = {u.r0 ^ t_lo ^ corr, u.r1 ^ t_hi};
    return r;
}

of128 of128_mul(of128 a, of128 b) { return of256_reduce(mul_unreduced_with(a, b, CLMUL)); }
of128 of128_mul_portable(of128 a, of128 b) { return of256_reduce(mul_unreduced_with(a, b, clmul64)); }

/* field/gf2_128.rs:127-135 */
of128 of128_mul_by_x(of128 z) {
    uint64_t carry = z.hi >> 63;
    of128 r = {(z.lo << 1) ^ (0x87 & (0 - carry)), (z.hi << 1) | (z.lo >> 63)};
    return r;
}

/* Fermat inverse x^(2^128-2) (field/gf2_128.rs:59-69). */
of128 of128_inv(of128 a) {
    of128 r = {1, 0};
    of128 cur = of128_mul(a, a);
    for (int i = 1; i < 128; i++) {
        r = of128_mul(r, cur);
        cur = of128_mul(cur, cur);
    }
    return r;
}

/* GF(2^8) with x^8 + x^4 + x^3 + x + 1 (field/gf2_8.rs:11-20, 70-100). */
uint8_t of8_mul(uint8_t a, uint8_t b) {
    uint16_t p = 0;
    for (int i = 0; i < 8; i++)
        if ((b >> i) & 1) p ^= (uint16_t)a << i;
    for (int i = 14; i >= 8; i--)
        if ((p >> i) & 1) p ^= (uint16_t)(0x11B << (i - 8));
    return (uint8_t)p;
}

uint8_t of8_inv(uint8_t a) {
    uint8_t r = 1, sq = a;
    for (int i = 0; i < 8; i++) {
        if ((0xFE >> i) & 1) r = of8_mul(r, sq);
        sq = of8_mul(sq, sq);
    }
    return r;
}

/*
 * phi8 (field/phi8.rs:19,1047): the field embedding GF(2^8) -> GF(2^128).
 * It is fixed by the image g of the AES generator 0x02 (table entry 2 in
 * phi8.rs:27-30); every other entry is the GF(2)-linear extension
 * phi8(a) = sum_i bit_i(a) g^i.  Built once at load time.
 */
static const of128 PHI8_OF_X = {0x6b8330483c2e9849ULL, 0x0dcb364640a222feULL};
of128 OF_PHI8_TABLE[256];

__attribute__((constructor)) static void build_phi8(void) {
    of128 pw[8];
    pw[0].lo = 1;
    pw[0].hi = 0;
    for (int i = 1; i < 8; i++) pw[i] = of128_mul_portable(pw[i - 1], PHI8_OF_X);
    of128 *t = OF_PHI8_TABLE;
    for (int a = 0; a < 256; a++) {
        of128 acc = {0, 0};
        for (int i = 0; i < 8; i++)
            if ((a >> i) & 1) acc = of128_add(acc, pw[i]);
        t[a] = acc;
    }
}

of128 of_phi8(uint8_t a) { return OF_PHI8_TABLE[a]; }
