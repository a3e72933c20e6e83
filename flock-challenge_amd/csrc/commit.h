// Device-side PCS commit building blocks (commit.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/flock_gpu.h"
#include "ctx.h"

namespace flock {

// Forward interleaved NTT over 2^k positions x num_lanes lanes (64), from
// start_layer, run as the passes of ntt_plan(k, start_layer).  When d_msg is
// not null the first pass reads the message (positions masked to
// 2^(k - start_layer)) instead of the codeword: the state after the
// replication layers.  d_twiddles (may be null: the context's cached table)
// is the breadth-first twiddle table of dimension k.
int ntt_forward_interleaved_device(Ctx *ctx, hipStream_t st, void *d_data, int k, int num_lanes, int start_layer,
                                   const void *d_msg = nullptr, const void *d_twiddles = nullptr);
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
// The kernels that can run an NTT pass of F layers: shared-table passes
// (F = 3 or 4; W wavefronts per workgroup with G groups each, 2^s >= W G
// groups sharing the pass's twiddles; full byte tables rebuilt per twiddle,
// or all the pass's nibble tables built once; or one table buffer, so that
// two workgroups share a CU), and per-wavefront tail passes
// (F = 1..4, NB nibble tables in flight, or no tables: products by shifted
// XORs, for the short twiddles of the deepest layers).
enum NttKernel : uint8_t {
  NTT_SHARED_W8G2 = 0,
  NTT_SHARED_W8G1 = 1,
  NTT_SHARED_W16G1 = 2,
  NTT_TAIL_NB2 = 3,
  NTT_TAIL_NB1 = 4,
  NTT_NIBSHARED_W16 = 5,
  NTT_NIBSHARED_W8 = 6,
  NTT_TAIL_SHIFT = 7,
  NTT_SHARED1_W8G1 = 8,
  NTT_SHARED1_W4G2 = 9,
  NTT_KERNELS = 10
};
// One pass: layers [l, l + F) with kernel.
struct NttPass {
  int l, F;
  NttKernel kernel;
};
// log2 of the groups a workgroup of the kernel takes.
int ntt_kernel_groups_log(NttKernel kn);
bool ntt_pass_feasible(int k, int l, int F, int kernel);
// Workgroups of the pass on a 2^k-position codeword.
uint64_t ntt_pass_wgs(int k, int F, NttKernel kn);
// Passes of 4 layers (shared while 8 or more groups share the twiddles),
// the rest in one tail pass.
std::vector<NttPass> ntt_default_plan(int k, int start);
// Passes that cover [start, k) in order, each feasible.
bool ntt_plan_valid(int k, int start, const std::vector<NttPass> &p);
// The tuned plan of (k, start) if one is installed, else the default.  The
// table is process-wide; at first use it is loaded from $FLOCK_NTT_PLANS or
// tune/ntt_plans_gfx950.json beside the library's directory.
std::vector<NttPass> ntt_plan(int k, int start);
// Install a plan (empty: back to the default); FLOCK_E_ARG with err set if
// it is invalid.
int ntt_plan_install(int k, int start, const std::vector<NttPass> &p, std::string &err);
// Merge a plan file (null: the default path) into the table; FLOCK_E_ARG
// with err set if it cannot be read or holds an invalid plan (the table is
// then unchanged).
int ntt_plans_load(const char *path, std::string &err);
// Time every feasible (l, F, kernel) of the transform (k, start) on the
// context's stream (reps launches each, after a warm-up), install the
// fastest partition, and describe it as a JSON object in json.
int ntt_autotune(Ctx *ctx, int k, int start, int reps, std::string &json);
// One fused pass launch of F layers with kernel kn over jobs.
int ntt_batch_device(Ctx *ctx, hipStream_t st, int F, NttKernel kn, const NttJob *d_jobs, uint32_t n_jobs,
                     uint64_t wgs);
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
