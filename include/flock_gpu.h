/*
 * flock_gpu.h — C ABI of the MI355X (gfx950) Flock BLAKE3 prover hot path.
 *
 * The reference exposes this path as Rust functions, not as an FFI; each
 * entry point below names the reference interface it replaces
 * (paths relative to the Layr-Labs/flock-challenge checkout).  Buffers are
 * owned by the caller; pointers documented as "device" are HIP device
 * pointers on the context's GPU, all others are host memory.  Every device
 * error is returned as a negative code, never aborted; flock_last_error()
 * gives the text.  Calls on distinct contexts are independent; calls on one
 * context must not overlap.
 */
#ifndef FLOCK_GPU_H
#define FLOCK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FLOCK_OK 0
#define FLOCK_E_SHAPE (-1)  /* input shape does not match the setup / params */
#define FLOCK_E_OOM (-2)    /* device or host allocation failed */
#define FLOCK_E_DEVICE (-3) /* HIP runtime / kernel error */
#define FLOCK_E_CAP (-4)    /* output buffer too small; *out_len holds the need */
#define FLOCK_E_ARG (-5)    /* invalid argument */
#define FLOCK_E_VERIFY (-6) /* proof rejected */

typedef struct flock_ctx flock_ctx;

/* PcsParams (crates/flock-core/src/pcs/commit.rs:32-50).  profile: 0 Fast,
 * 1 Slim, 2 Secure (pcs/ligerito.rs:65-70); merkle_hash: 0 Sha256, 1 Blake3
 * (hash.rs:23-27).  Only merkle_hash = 1 is implemented. */
typedef struct {
    uint32_t m;
    uint32_t log_inv_rate;
    uint32_t log_batch_size;
    uint32_t profile;
    uint32_t merkle_hash;
} flock_pcs_params;

/* One BLAKE3 compression input, the reference `Compression` tuple
 * (crates/flock-prover/src/r1cs_hashes/blake3.rs:1087). */
typedef struct {
    uint32_t cv[8];
    uint32_t m[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} flock_compression;

/* Number of visible GPUs (0 when none), without creating a context. */
int flock_device_count(void);

/* Context on GPU `device`.  Replaces the process-global state the reference
 * keeps in rayon pools and scratch caches (crates/flock-core/src/lib.rs:79). */
int flock_ctx_create(int device, flock_ctx **out);
void flock_ctx_destroy(flock_ctx *ctx);
const char *flock_last_error(const flock_ctx *ctx);

/* PCS commit of an F128-packed witness: replicate, interleaved additive NTT
 * from layer log_inv_rate, BLAKE3 Merkle tree over positions.  Replaces
 * `pcs::commit` (crates/flock-core/src/pcs/commit.rs:250) returning
 * `Commitment.root`.  z_packed: host, 2^(m-7) x 16 bytes. */
int flock_commit(flock_ctx *ctx, const flock_pcs_params *params, const void *z_packed, size_t z_bytes,
                 uint8_t root[32]);

/* Same, device-resident: d_z (2^(m-7) x 16 B), d_codeword (2^(k_code) x
 * 2^log_batch_size x 16 B) and d_tree ((2^(k_code+1) - 1) x 32 B, flat layout of
 * crates/flock-core/src/merkle.rs:15-21) are device pointers; root is host.
 * `stream` is a hipStream_t or NULL for the context's stream. */
int flock_commit_device(flock_ctx *ctx, const flock_pcs_params *params, const void *d_z, void *d_codeword,
                        void *d_tree, uint8_t root[32], void *stream);

/* Device memory helpers for callers without their own HIP runtime handle
 * (tests, benchmarks, FFI hosts).  Pointers are device pointers on the
 * context's GPU; copies are synchronous with respect to the context stream. */
int flock_device_alloc(flock_ctx *ctx, size_t bytes, void **d_ptr);
int flock_device_free(flock_ctx *ctx, void *d_ptr);
int flock_memcpy_h2d(flock_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int flock_memcpy_d2h(flock_ctx *ctx, void *dst, const void *d_src, size_t bytes);
int flock_synchronize(flock_ctx *ctx);

/* Device phase timing (HIP events on the launching stream), the counterpart
 * of the reference's FLOCK_PHASE_TIMING / prove_fast_timed
 * (crates/flock-prover/src/prover.rs:1267-1284).  Enabling resets the totals.
 * The report is text, one line per phase: "<name> <count> <total_ms>". */
int flock_timing_enable(flock_ctx *ctx, int on);
int flock_timing_report(flock_ctx *ctx, char *buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
