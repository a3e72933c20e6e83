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

/* Deadline of every host wait for the device on this context (readbacks,
 * stream and event waits), in seconds; <= 0 waits without limit.  Default
 * 120 s.  A wait past it returns FLOCK_E_DEVICE and marks the context timed
 * out: its work may still be running, so every later call on it returns
 * FLOCK_E_DEVICE, and flock_ctx_destroy releases its host state but leaves
 * its device memory to the process's end instead of waiting on that work.
 * No counterpart in the reference's prover; its harness bounds a whole
 * trial from outside (benchmark-tools/harness/src/main.rs:20-21). */
int flock_ctx_set_wait_timeout(flock_ctx *ctx, double seconds);

/* Context options (the proof bytes never depend on them):
 *  "device_transcript" (default 0): 1 runs the Fiat-Shamir steps of the
 *    zerocheck rounds and of the Ligerito fold rounds on the device (the
 *    BLAKE3 transcript state moves to device memory and back; each round's
 *    reduction kernel absorbs the round's message and, in the zerocheck,
 *    samples its challenge; a Ligerito round's proof of work and challenge
 *    run in kernels of their own; each BLAKE3 compression spread over a quad
 *    of lanes) instead of host round trips per round.  The serial hashing
 *    and the launches cost about what the round trips do: 2^18-compression
 *    proofs measure within 1 % either way, small ones up to 5 % slower; it
 *    is for callers that capture or replay a proof's rounds without the
 *    host.
 *  "zc_graph" (default 1): with the device transcript, the zerocheck rounds'
 *    launches are captured once per size and buffer set in a HIP graph and
 *    replayed by later proofs.
 * Returns FLOCK_E_ARG for an unknown name. */
int flock_ctx_set_option(flock_ctx *ctx, const char *name, int64_t value);

/* PCS commit of an F128-packed witness: replicate, interleaved additive NTT
 * from layer log_inv_rate, BLAKE3 Merkle tree over positions.  Replaces
 * `pcs::commit` (crates/flock-core/src/pcs/commit.rs:250) returning
 * `Commitment.root`.  z_packed: host, 2^(m-7) x 16 bytes. */
int flock_commit(flock_ctx *ctx, const flock_pcs_params *params, const void *z_packed, size_t z_bytes,
                 uint8_t root[32]);

/* The buffer sizes flock_commit_device needs for `params` (PcsParams::
 * log_msg_len / codeword_len_f128 / n_leaves, crates/flock-core/src/pcs/
 * commit.rs:54-90): message and codeword in F128 elements, tree in 32-byte
 * nodes (2 n_leaves - 1).  Returns FLOCK_E_SHAPE for parameters the commit
 * kernels do not support (log_batch_size != 6, log_inv_rate outside 1..4,
 * m outside 14..40, a hash other than BLAKE3); both commit entry points
 * reject those before touching the device.  Host only, no context. */
int flock_pcs_params_sizes(const flock_pcs_params *params, uint64_t *msg_f128, uint64_t *codeword_f128,
                           uint64_t *tree_nodes);

/* Same, device-resident: d_z (2^(m-7) x 16 B), d_codeword (2^(k_code) x
 * 2^log_batch_size x 16 B) and d_tree ((2^(k_code+1) - 1) x 32 B, flat layout of
 * crates/flock-core/src/merkle.rs:15-21) are device pointers; root is host.
 * `stream` is a hipStream_t or NULL for the context's stream. */
int flock_commit_device(flock_ctx *ctx, const flock_pcs_params *params, const void *d_z, void *d_codeword,
                        void *d_tree, uint8_t root[32], void *stream);

/* flock_commit_device for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The same kernels and
 * results; the twiddle table, the only device memory flock_commit_device
 * allocates itself, is rebuilt by every call in a fixed array of 16 MiB
 * per GPU (held until the process ends), so k_code = m - 13 + log_inv_rate
 * must be at most 20 (FLOCK_E_SHAPE otherwise).  Calls that overlap on one
 * GPU share that array and race: serialize them. */
int flock_commit_device_noalloc(flock_ctx *ctx, const flock_pcs_params *params, const void *d_z, void *d_codeword,
                                void *d_tree, uint8_t root[32], void *stream);

/* Schedules of the commit's forward transform.  No counterpart in the
 * reference (its CPU transform has no launch shapes to choose).  The
 * transform over 2^k positions from layer `start` (k = m - 13 +
 * log_inv_rate, start = log_inv_rate) runs as passes, each applying F
 * consecutive layers [l, l + F) with one of several kernels:
 *   0, 1, 2: shared full-table passes (F = 3 or 4; 8 wavefronts x 2
 *            groups, 8 x 1, 16 x 1; they need 2^(k - l - F) >= 16, 8, 16);
 *   3, 4:    tail passes (F = 1..4; two or one nibble tables in flight);
 *   5, 6:    shared nibble-table passes (F = 3 or 4; 16 or 8 wavefronts,
 *            needing 2^(k - l - F) >= 16 or 8);
 *   7:       tail passes without tables (F = 1..4; products by shifted
 *            XORs, for the short twiddles of the deepest layers);
 *   8, 9:    shared full-table passes with one table buffer (F = 3 or 4;
 *            8 x 1 and 4 x 2 wavefronts x groups), so that two
 *            workgroups share a CU.
 * Every valid plan gives the same codeword; only the time differs.  Plans
 * live in a process-wide table, loaded at first use from $FLOCK_NTT_PLANS or
 * else tune/ntt_plans_gfx950.json in the directory above the library's
 * (JSON {"plans": [{"k": K, "start": S, "passes": [[l, F, kernel], ...]},
 * ...]}); shapes without a plan use passes of 4 layers.
 *
 * flock_ntt_autotune times every feasible (l, F, kernel) of the transform
 * `params` selects on the context's GPU (reps launches each), installs the
 * fastest partition and writes it, with the timings, as a JSON object into
 * json (FLOCK_E_CAP if cap is too small).  flock_ntt_plan_set installs a
 * plan of n_passes (l, F, kernel) triples (n_passes = 0: back to the
 * default), flock_ntt_plan_get reads the plan in force, flock_ntt_plans_load
 * merges a plan file (NULL: the default path) into the table. */
int flock_ntt_autotune(flock_ctx *ctx, const flock_pcs_params *params, int reps, char *json, size_t cap);
int flock_ntt_plan_set(flock_ctx *ctx, uint32_t k, uint32_t start, const int32_t *passes, uint32_t n_passes);
int flock_ntt_plan_get(flock_ctx *ctx, uint32_t k, uint32_t start, int32_t *passes, uint32_t cap,
                       uint32_t *n_passes);
int flock_ntt_plans_load(flock_ctx *ctx, const char *path);

/* Device memory helpers for callers without their own HIP runtime handle
 * (tests, benchmarks, FFI hosts).  Pointers are device pointers on the
 * context's GPU; copies are synchronous with respect to the context stream. */
int flock_device_alloc(flock_ctx *ctx, size_t bytes, void **d_ptr);
int flock_device_free(flock_ctx *ctx, void *d_ptr);
int flock_memcpy_h2d(flock_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int flock_memcpy_d2h(flock_ctx *ctx, void *dst, const void *d_src, size_t bytes);
int flock_memcpy_d2d(flock_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int flock_synchronize(flock_ctx *ctx);

/* ------------------------------------------------------------------ prover */

/* A prover setup for n_blocks compressions: Blake3Setup::with_profile
 * (crates/flock-prover/src/r1cs_hashes/blake3.rs:3312-3347) with the
 * benchmark parameters (rate 1/2, 64 lanes, BLAKE3 Merkle hash).  Builds the
 * R1CS matrices and statement digest.  Device buffers are reserved in the
 * context by the first prove and reused by later ones of the same size.
 * profile: only 0 (Fast); merkle_hash: only 1 (Blake3).  m = 14 +
 * ceil(log2(max(n_blocks, 8))) must have a Ligerito configuration (22..35). */
typedef struct flock_setup flock_setup;
int flock_setup_create(flock_ctx *ctx, uint64_t n_blocks, uint32_t profile, uint32_t merkle_hash,
                       flock_setup **out);
void flock_setup_destroy(flock_setup *setup);
/* m of the setup and its R1CS statement digest (BlockR1cs::statement_digest,
 * crates/flock-core/src/r1cs.rs:343-365). */
int flock_setup_info(const flock_setup *setup, uint32_t *m, uint8_t digest[32]);

/* The R1CS statement digest for setups of m = 14 + log2(padded blocks)
 * (BlockR1cs::statement_digest, crates/flock-core/src/r1cs.rs:343-365): the
 * digest the transcript binds first.  Host only, no context: it folds the
 * digest of this m from the statement's m-independent spine (the BLAKE3 tree
 * of the 170 MB matrix framing, which only its first chunk makes depend on
 * m), loading the spine from the persisted cache or hashing it on host
 * threads when neither this process nor the cache has it. */
int flock_statement_digest(uint32_t m, uint8_t digest[32]);

/* Setup work for the statement digest on the context's GPU: the spine is
 * taken from this process's memory, else from the persisted cache (a small
 * file under $FLOCK_CACHE_DIR, else lib/cache next to the library, keyed by a
 * fingerprint of the matrices), else hashed on the GPU and persisted.
 * flock_setup_create does this itself.  `from` (may be NULL) receives where
 * it came from: "memory", "file", "gpu" or "host" (the GPU hash failed). */
int flock_statement_setup(flock_ctx *ctx, char *from, size_t from_cap);

/* Blake3Setup::prove_fast (blake3.rs:3405) followed by the worker's
 * R1csProofBundleLigerito::to_bytes (crates/flock-core/src/proof_io.rs:161-208,
 * worker/src/main.rs:42-45): the proof bundle bytes ("FLOCK" 04 02 + bincode)
 * for `n_blocks` compressions under the BLAKE3 Fiat-Shamir transcript with
 * domain separator `domain`: the challenger the benchmark worker builds,
 * FsChallenger::with_hash(domain, HashKind::Blake3) (benchmark-tools/worker/
 * src/main.rs:29,41; crates/flock-core/src/challenger.rs:271-283).  The ABI
 * is BLAKE3-only: the reference's FsChallenger::new(domain) is the SHA-256
 * transcript (challenger.rs:261-263), which gives different bytes and is not
 * implemented here.  n_blocks must equal the setup's (the reference
 * asserts, blake3.rs:3520).  If cap is too small, returns FLOCK_E_CAP with
 * *out_len set; the bytes stay available through flock_last_proof. */
int flock_prove_blake3(flock_setup *setup, const flock_compression *blocks, uint64_t n_blocks,
                       const uint8_t *domain, size_t domain_len, uint8_t *out, size_t cap, size_t *out_len);

/* Same with the compressions already in device memory (d_blocks, n_blocks x
 * 112 B) and the work issued on `stream` (NULL: the context stream). */
int flock_prove_blake3_device(flock_setup *setup, const flock_compression *d_blocks, uint64_t n_blocks,
                              const uint8_t *domain, size_t domain_len, uint8_t *out, size_t cap, size_t *out_len,
                              void *stream);

/* The worker path from a seed (worker/src/main.rs:36-45): expand
 * generate_compressions(log2, seed) (benchmark-tools/common/src/lib.rs:20-33)
 * on the device, then prove.  The setup must be for 2^log2 blocks. */
int flock_prove_blake3_seeded(flock_setup *setup, uint32_t log2, uint64_t seed, const uint8_t *domain,
                              size_t domain_len, uint8_t *out, size_t cap, size_t *out_len);

/* generate_compressions(log2, seed) into device memory d_out (2^log2 x 112 B). */
int flock_generate_compressions_device(flock_ctx *ctx, uint32_t log2, uint64_t seed, flock_compression *d_out);

/* Copy the bytes of the context's last proof (kept after FLOCK_E_CAP). */
int flock_last_proof(flock_ctx *ctx, uint8_t *out, size_t cap, size_t *out_len);

/* ------------------------------------------------------------------ verifying */

/* Blake3Setup::verify (blake3.rs:3698) preceded by the harness's checks of
 * the bundle (benchmark-tools/harness/src/main.rs:319-366): parse the bundle
 * bytes of a proof of n_blocks compressions, check the commitment parameters,
 * replay the BLAKE3 Fiat-Shamir transcript with domain `domain` and check the
 * zerocheck, lincheck, ring-switch and Ligerito steps (sumchecks, proofs of
 * work, Merkle openings, final residual).  expected_root (may be NULL) is the
 * commitment root recomputed from the compressions (flock_commitment_root);
 * the bundle's must equal it.  Host only: no context or GPU.  Returns
 * FLOCK_OK when accepted, FLOCK_E_VERIFY when rejected, FLOCK_E_SHAPE for an
 * unsupported size; the reason goes to err (NUL-terminated, may be NULL). */
int flock_verify_blake3(const uint8_t *proof, size_t len, uint64_t n_blocks, const uint8_t *domain,
                        size_t domain_len, const uint8_t expected_root[32], char *err, size_t err_cap);

/* The commitment root of the witness of the compressions: the first step of
 * prove_fast (pcs::commit of the packed witness, crates/flock-core/src/pcs/
 * commit.rs:250-320), run on the device.  n_blocks must equal the setup's. */
int flock_commitment_root(flock_setup *setup, const flock_compression *blocks, uint64_t n_blocks, uint8_t root[32]);

/* Same for generate_compressions(log2, seed) expanded on the device. */
int flock_commitment_root_seeded(flock_setup *setup, uint32_t log2, uint64_t seed, uint8_t root[32]);

/* BLAKE3 hash (public BLAKE3 specification, 32-byte output) of len bytes on
 * the host; no counterpart in the reference's API (it calls the blake3 crate),
 * here for fingerprinting proof bytes. */
void flock_blake3(const void *data, size_t len, uint8_t out[32]);

/* ------------------------------------------------------------------ batching */

/* One proof of a batch: d_blocks (device, n_blocks compressions), or, when
 * d_blocks is NULL, generate_compressions(log2, seed) expanded on the
 * device. */
typedef struct {
  const flock_compression *d_blocks;
  uint64_t n_blocks;
  uint32_t log2;
  uint64_t seed;
} flock_batch_item;

/* Many independent proofs of uneven sizes (256 .. 2^21 compressions each).
 * No counterpart in the reference (its worker proves one seed per process,
 * worker/src/main.rs:36-45); each proof is the same bundle bytes as
 * flock_prove_blake3 for the same input.  The fronts (witness, codeword NTT,
 * Merkle tree) of the items of at most 2^14 compressions, whose short
 * launches make launch overhead matter, run fused: one launch per step for
 * all of them, the launch sequence captured in a HIP graph per shape (the
 * list of their sizes) and replayed when the same shape comes again, if
 * use_graph.  Then the n_workers (1..16) worker streams take the larger
 * items as whole proofs, largest first, and the rest of the small items'
 * proofs, one worker at a time on small items while large ones remain (their
 * latency-bound proofs overlap the large proofs' kernels).
 * out receives the proofs back to back: proof i is out[offsets[i] ..
 * offsets[i + 1]) (offsets has n_items + 1 entries); if cap is too small,
 * FLOCK_E_CAP with offsets filled and nothing copied.  done_s (n_items, may
 * be NULL) receives each proof's completion time from the call's start;
 * front_s (may be NULL) the fused fronts' time; replayed (may be NULL) 1
 * when the fronts replayed a captured graph. */
int flock_prove_batch(flock_ctx *ctx, const flock_batch_item *items, uint32_t n_items, const uint8_t *domain,
                      size_t domain_len, uint32_t n_workers, int use_graph, uint8_t *out, size_t cap,
                      uint64_t *offsets, double *done_s, double *front_s, int *replayed);

/* ------------------------------------------------------------------ serving */

/* A proof server on GPU `device`: continuous batching of independent proofs.
 * Small requests (at most 2^14 compressions) queued together have their
 * fronts run fused (one launch per step for all of them, flock_prove_batch's
 * kernels) and the rest of each proof on a worker; larger requests run
 * whole on one of n_workers worker streams.
 * No counterpart in the reference (its worker proves one seed per process,
 * worker/src/main.rs:36-45); each proof is the same bundle bytes as
 * flock_prove_blake3 for the same input, whatever else is in flight.
 * n_workers proofs run concurrently, each on its own HIP stream and buffers;
 * a request is admitted (FIFO) when a worker is free and the HBM footprint of
 * the admitted work stays within hbm_budget_bytes (0: 80% of free memory).
 * A worker keeps the buffers of the largest proof it has run; when the head
 * request fits no idle worker, idle workers free theirs until it does, so a
 * request within the budget is always admitted eventually.
 * reserve_log2 > 0 sizes each worker's buffers for 2^reserve_log2
 * compressions at start (one throwaway proof), so serving never reallocates. */
typedef struct flock_server flock_server;
int flock_server_create(int device, int n_workers, uint64_t hbm_budget_bytes, uint32_t reserve_log2,
                        flock_server **out);
void flock_server_destroy(flock_server *server);
const char *flock_server_last_error(const flock_server *server);
/* Enqueue a proof of host compressions (copied) or of
 * generate_compressions(log2, seed); *request_id identifies the result. */
int flock_server_submit(flock_server *server, const flock_compression *blocks, uint64_t n_blocks,
                        const uint8_t *domain, size_t domain_len, uint64_t *request_id);
int flock_server_submit_seeded(flock_server *server, uint32_t log2, uint64_t seed, const uint8_t *domain,
                               size_t domain_len, uint64_t *request_id);
/* Take one finished result, waiting up to wait_ms (-1: forever, 0: no wait).
 * Returns 1 with the proof bytes copied out, 0 if none finished in time, or a
 * negative error code (the request failed, or cap < *out_len).  latency_s is
 * submit -> bytes ready, queue_s submit -> start. */
int flock_server_poll(flock_server *server, int wait_ms, uint64_t *request_id, uint8_t *out, size_t cap,
                      size_t *out_len, double *latency_s, double *queue_s);
/* Requests queued or running. */
int flock_server_pending(flock_server *server, uint64_t *n);
/* Requests whose fronts (witness, codeword NTT, Merkle tree) ran fused with
 * other requests' in one set of launches, and the number of front launches
 * sets so far. */
int flock_server_stats(flock_server *server, uint64_t *fused_requests, uint64_t *front_launches);
/* The wait deadline of every worker's context (flock_ctx_set_wait_timeout),
 * applied from each worker's next request on.  A request whose wait times
 * out fails with FLOCK_E_DEVICE and its worker retires; when no worker is
 * left, queued requests fail and submits return FLOCK_E_DEVICE. */
int flock_server_set_wait_timeout(flock_server *server, double seconds);

/* Device phase timing (HIP events on the launching stream), the counterpart
 * of the reference's FLOCK_PHASE_TIMING / prove_fast_timed
 * (crates/flock-prover/src/prover.rs:1267-1284).  Enabling resets the totals.
 * The report is text, one line per phase: "<name> <count> <total_ms>". */
int flock_timing_enable(flock_ctx *ctx, int on);
/* Restrict timing to the listed phases (comma-separated names; NULL or "" =
 * every phase, plus host-side sections reported as "host:<name>"). */
int flock_timing_select(flock_ctx *ctx, const char *names);
int flock_timing_report(flock_ctx *ctx, char *buf, size_t cap);

#ifdef __cplusplus
}
#endif
#endif
