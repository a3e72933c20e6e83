/*
 * Oracle: the benchmark's seeded input generator.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

/* SplitMix64 step; the generator keeps only the low 32 bits of each output
 * (benchmark-tools/common/src/lib.rs:10-17). */
uint64_t o_splitmix_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ULL;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static uint32_t next_u32(uint64_t *s) { return (uint32_t)o_splitmix_next(s); }

/* benchmark-tools/common/src/lib.rs:20-33: seed ^ rotl(log2, 29); per input
 * 8 cv words, 16 message words, counter = next_u32 widened (upper half 0),
 * block_len 64, flags 11. */
void o_generate_compressions(uint32_t log2_size, uint64_t seed, ocompression *out) {
    uint64_t l = (uint64_t)log2_size;
    uint64_t st = seed ^ ((l << 29) | (l >> 35));
    uint64_t count = 1ULL << log2_size;
    for (uint64_t i = 0; i < count; i++) {
        for (int w = 0; w < 8; w++) out[i].cv[w] = next_u32(&st);
        for (int w = 0; w < 16; w++) out[i].m[w] = next_u32(&st);
        out[i].counter = (uint64_t)next_u32(&st);
        out[i].block_len = 64;
        out[i].flags = 11;
    }
}
