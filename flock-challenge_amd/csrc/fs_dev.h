// The Fiat-Shamir transcript on the device (transcript_dev.h, fs_dev.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace flock {

// Bytes of the transcript state, the layout of the host's hb3::Hasher.
constexpr size_t FS_STATE_BYTES = 1848;

// Byte tables of rho * ws[q] for the 64 values ws (device): out[v * 8 + p] =
// sum over the set bits t of v of rho ws[8 p + t] (2048 F128).
int rho_bytetab_device(Ctx *ctx, hipStream_t st, const void *d_ws, const void *d_rho, void *d_out);

// The device side of a Ligerito fold round's challenge with the transcript
// state at d_fs: with bits > 0, the proof of work (Transcript::state_digest,
// the smallest nonce whose BLAKE3(digest | nonce_le | 0^24) starts with
// `bits` zero bits, Transcript::observe_bytes of the nonce, into *d_nonce),
// then Transcript::sample into *d_r.  d_work: FLOCK_POW_WORK_BYTES of device
// memory the steps share: the scan's best nonce in its first 8 bytes (~0
// before the first call; every call leaves it so) and the state's digest in
// bytes 32..63 (lig_pow_digest(d_work)); with digest_ready, a proof of work of
// more than FLOCK_POW_WAVE_BITS bits takes the digest already there (written
// by the previous message's transcript step, or uploaded) instead of
// computing it.  A proof of work that finds no nonce in its bounded search
// sets *d_status to 1 (and carries on with nonce 0).
constexpr size_t FLOCK_POW_WORK_BYTES = 64;
// The proof-of-work scan alone, host transcript or device: the smallest
// nonce >= start for the 32-byte digest (host; null: the one at
// lig_pow_digest(d_work)) into the first 8 bytes of d_work (atomicMin: ~0 on
// entry), through an ordered scan of up to 64 windows of 2^max(bits, 10)
// nonces (2^30 at most); *span (if not null) gets the nonces it covers, after
// which nothing found leaves ~0.
int pow_chain_device(Ctx *ctx, hipStream_t st, void *d_work, const uint8_t *digest, uint32_t bits, uint64_t start,
                     uint64_t *span);
constexpr uint32_t FLOCK_POW_WAVE_BITS = 6;
inline void *lig_pow_digest(void *d_work) { return (uint8_t *)d_work + 32; }
int lig_challenge_device(Ctx *ctx, hipStream_t st, void *d_fs, uint32_t bits, void *d_work, bool digest_ready,
                         void *d_nonce, void *d_r, void *d_status);

}  // namespace flock
