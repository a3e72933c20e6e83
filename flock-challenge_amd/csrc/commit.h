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
// One instance of a fused NTT pass launch: codeword, twiddle table of
// dimension k, the pass's first layer l, the message a first pass reads
// (null: the codeword) with its position mask, and the instance's first
// workgroup in the launch.
struct NttJob {
  uint4 *data;  // F128 elements
  const uint4 *tw;
  const uint4 *src;
  uint64_t src_pos_mask;
  int k, l;
  uint64_t wg_begin;
};
// One instance of a fused Merkle launch: leaves (data, leaf_bytes, n
// leaves) or one subtree step (tree, in_off, n nodes in, levels); for the
// root gather, tree and its node count n.
struct MerkleJob {
  const void *data;
  void *tree;
  uint64_t n, in_off;
  uint32_t leaf_bytes;
  int levels;
  uint64_t wg_begin;
};
// Pass variants (ntt_pass_variant) and their fused launches.
enum NttVariant { NTT_SHARED_G2 = 0, NTT_SHARED_G1, NTT_TAIL1, NTT_TAIL2, NTT_TAIL3, NTT_TAIL4, NTT_VARIANTS };
// The variant that runs F layers from layer l of a 2^k-position codeword,
// and its workgroup count.
NttVariant ntt_pass_variant(int k, int l, int F, uint64_t *wgs);
int ntt_batch_device(Ctx *ctx, hipStream_t st, NttVariant v, const NttJob *d_jobs, uint32_t n_jobs, uint64_t wgs);
int merkle_leaf_batch_device(Ctx *ctx, hipStream_t st, const MerkleJob *d_jobs, uint32_t n_jobs, uint64_t wgs);
int merkle_subtree_batch_device(Ctx *ctx, hipStream_t st, const MerkleJob *d_jobs, uint32_t n_jobs, uint64_t wgs);
int gather_roots_device(Ctx *ctx, hipStream_t st, const MerkleJob *d_jobs, uint32_t n_jobs, void *d_roots);

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
