// Device-side PCS commit building blocks (commit.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"

namespace flock {

// Forward interleaved additive NTT over 2^k positions x num_lanes lanes, from
// `start_layer`, fusing up to `max_fuse` layers per pass.
// Forward interleaved NTT from start_layer.  When d_msg is not null the
// first pass reads the message (positions masked to 2^(k - start_layer))
// instead of the replicated codeword, if that pass is a shared-table pass.
// d_twiddles (may be null: the context's cached table) is the breadth-first
// twiddle table of dimension k.
int ntt_forward_interleaved_device(Ctx *ctx, hipStream_t st, void *d_data, int k, int num_lanes, int start_layer,
                                   int max_fuse, const void *d_msg = nullptr, const void *d_twiddles = nullptr);
// Flat BLAKE3 Merkle tree of n_leaves leaves of leaf_bytes each.
int merkle_tree_device(Ctx *ctx, hipStream_t st, const void *d_data, uint32_t leaf_bytes, uint64_t n_leaves,
                       void *d_tree);
// Fill out with out_len/msg_len copies of msg (F128 counts, powers of two).
int replicate_device(Ctx *ctx, hipStream_t st, const void *msg, uint64_t msg_len_f128, void *out,
                     uint64_t out_len_f128);
// Why the commit kernels cannot take these parameters, or nullptr when they
// can (64 lanes, log_inv_rate 1..4, m 14..40, BLAKE3 Merkle hash).
const char *commit_params_unsupported(const flock_pcs_params *p);
// Full commit; copies the root to host memory `root` and synchronizes `st`,
// or, when root_ready is given, records it after the copy instead (then
// `root` must be pinned memory, valid once the event has completed).
// d_twiddles as for ntt_forward_interleaved_device.
int commit_device(Ctx *ctx, const flock_pcs_params *p, const void *d_z, void *d_codeword, void *d_tree,
                  uint8_t root[32], hipStream_t st, hipEvent_t root_ready = nullptr,
                  const void *d_twiddles = nullptr);

}  // namespace flock
