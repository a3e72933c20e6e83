// PCS commit on gfx950: interleaved additive NTT of the replicated message ->
// BLAKE3 Merkle.
//
// Layout in HBM (the reference's, crates/flock-core/src/pcs/commit.rs:235-245):
// codeword[pos * NL + lane] F128, NL = 2^log_batch_size interleaved RS lanes;
// Merkle leaf p = the NL*16 contiguous bytes of position p.
//
// The forward transform's kernels, plans and tuning are in ntt.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "commit.h"
#include "ctx.h"
#include "fused_job.h"
#include "gf128_dev.h"

namespace flock {

// Replicate the message into every 2^log_inv_rate sub-block: the exact state
// after the first log_inv_rate layers (pcs/commit.rs:451-467).
__global__ void replicate_kernel(const uint4 *__restrict__ msg, uint4 *__restrict__ cw, uint64_t msg_len,
                                 uint64_t cw_len) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < cw_len; i += stride) cw[i] = msg[i & (msg_len - 1)];
}

// Leaf = non-root BLAKE3 chunk chaining value of leaf_bytes (<= 1024) bytes
// (merkle.rs:200-205): counter 0, CHUNK_START on the first block, CHUNK_END on
// the last, final block zero-padded with its true length.
__device__ __forceinline__ void merkle_leaf_body(const uint4 *__restrict__ data, uint32_t leaf_bytes,
                                                 uint64_t n_leaves, uint4 *__restrict__ out, uint64_t wg) {
  uint64_t leaf = wg * blockDim.x + threadIdx.x;
  if (leaf >= n_leaves) return;
  const uint32_t leaf_u4 = leaf_bytes >> 4;
  const uint4 *p = data + leaf * leaf_u4;
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cv[i] = b3::IV[i];
  const uint32_t nblocks = leaf_bytes == 0 ? 1u : (leaf_bytes + 63u) >> 6;
  for (uint32_t b = 0; b < nblocks; ++b) {
    uint32_t m[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t idx = b * 4u + q;
      uint4 w = idx < leaf_u4 ? p[idx] : make_uint4(0u, 0u, 0u, 0u);
      m[4 * q] = w.x; m[4 * q + 1] = w.y; m[4 * q + 2] = w.z; m[4 * q + 3] = w.w;
    }
    uint32_t len = leaf_bytes - 64u * b;
    len = len > 64u ? 64u : len;
    uint32_t flags = (b == 0 ? b3::CHUNK_START : 0u) | (b + 1 == nblocks ? b3::CHUNK_END : 0u);
    b3::compress_cv(cv, m, 0, len, flags);
  }
  out[2 * leaf] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
  out[2 * leaf + 1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

__global__ __launch_bounds__(256) void merkle_leaf_kernel(const uint4 *__restrict__ data, uint32_t leaf_bytes,
                                                          uint64_t n_leaves, uint4 *__restrict__ out) {
  merkle_leaf_body(data, leaf_bytes, n_leaves, out, blockIdx.x);
}

__global__ __launch_bounds__(256) void merkle_leaf_batch_kernel(const MerkleJob *__restrict__ jobs, uint32_t n_jobs) {
  const MerkleJob &J = job_of(jobs, n_jobs);
  merkle_leaf_body((const uint4 *)J.data, J.leaf_bytes, J.n, (uint4 *)J.tree, blockIdx.x - J.wg_begin);
}

// Up to 9 parent levels per launch: workgroup w takes input nodes
// [w * 2^L, (w + 1) * 2^L) of one level and writes the L levels above them
// (every level is kept: multiproofs read them), staging CVs in LDS.
__device__ __forceinline__ void merkle_subtree_body(uint4 *__restrict__ tree, uint64_t in_off, uint64_t n_in, int L,
                                                    uint64_t w, uint32_t (&cvs)[256][8]) {
  const uint32_t t = threadIdx.x;
  uint64_t off = in_off, n = n_in;  // current level's offset and node count
  uint32_t width = 1u << L;         // nodes of the current level in this workgroup
  for (int lev = 0; lev < L; ++lev) {
    const uint32_t outs = width >> 1;
    const uint64_t out_off = off + n;
    const bool active = t < outs;
    uint32_t cv[8];
    if (active) {
      uint32_t m[16];
      if (lev == 0) {
        const uint4 *src = tree + 2 * (off + w * width + 2 * t);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint4 v = src[q];
          m[4 * q] = v.x; m[4 * q + 1] = v.y; m[4 * q + 2] = v.z; m[4 * q + 3] = v.w;
        }
      } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          m[i] = cvs[2 * t][i];
          m[8 + i] = cvs[2 * t + 1][i];
        }
      }
#pragma unroll
      for (int i = 0; i < 8; ++i) cv[i] = b3::IV[i];
      b3::compress_cv(cv, m, 0, 64, b3::PARENT);
      const uint64_t node = out_off + w * outs + t;
      tree[2 * node] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
      tree[2 * node + 1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
    }
    __syncthreads();  // every read of this level's inputs from cvs is done
    if (active) {
#pragma unroll
      for (int i = 0; i < 8; ++i) cvs[t][i] = cv[i];
    }
    __syncthreads();  // the next level's inputs are in cvs
    off = out_off;
    n >>= 1;
    width = outs;
  }
}

__global__ __launch_bounds__(256) void merkle_subtree_kernel(uint4 *__restrict__ tree, uint64_t in_off,
                                                             uint64_t n_in, int L) {
  __shared__ uint32_t cvs[256][8];
  merkle_subtree_body(tree, in_off, n_in, L, blockIdx.x, cvs);
}

__global__ __launch_bounds__(256) void merkle_subtree_batch_kernel(const MerkleJob *__restrict__ jobs,
                                                                   uint32_t n_jobs) {
  __shared__ uint32_t cvs[256][8];
  const MerkleJob &J = job_of(jobs, n_jobs);
  merkle_subtree_body((uint4 *)J.tree, J.in_off, J.n, J.levels, blockIdx.x - J.wg_begin, cvs);
}

// roots[i] = the last node of job i's tree (n = its node count).
__global__ void gather_roots_kernel(const MerkleJob *__restrict__ jobs, uint32_t n_jobs, uint4 *__restrict__ roots) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_jobs) return;
  const uint4 *t = (const uint4 *)jobs[i].tree + 2 * (jobs[i].n - 1);
  roots[2 * i] = t[0];
  roots[2 * i + 1] = t[1];
}

int replicate_device(Ctx *ctx, hipStream_t st, const void *msg, uint64_t msg_len_f128, void *out,
                     uint64_t out_len_f128) {
  hipLaunchKernelGGL(replicate_kernel, dim3(2048), dim3(256), 0, st, (const uint4 *)msg, (uint4 *)out, msg_len_f128,
                     out_len_f128);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int merkle_tree_device(Ctx *ctx, hipStream_t st, const void *d_data, uint32_t leaf_bytes, uint64_t n_leaves,
                       void *d_tree) {
  if (leaf_bytes > 1024 || (leaf_bytes & 15) || n_leaves == 0 || (n_leaves & (n_leaves - 1))) {
    ctx->last_error = "merkle: leaf must be a multiple of 16 bytes, at most 1024, power-of-two leaf count";
    return FLOCK_E_SHAPE;
  }
  uint4 *tree = (uint4 *)d_tree;
  {
  ScopedPhase ph(ctx, "merkle_leaf_kernel", st);
  hipLaunchKernelGGL(merkle_leaf_kernel, dim3((unsigned)((n_leaves + 255) / 256)), dim3(256), 0, st,
                     (const uint4 *)d_data, leaf_bytes, n_leaves, tree);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  }
  ScopedPhase ph(ctx, "merkle_parents", st);
  uint64_t rs = 0, rl = n_leaves;
  while (rl > 1) {
    int L = 0;
    while (L < 9 && (2ull << L) <= rl) ++L;  // 2^L <= rl, at most 9 levels
    const uint64_t wgs = rl >> L;
    hipLaunchKernelGGL(merkle_subtree_kernel, dim3((unsigned)wgs), dim3(256), 0, st, tree, rs, rl, L);
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
    for (int i = 0; i < L; ++i) {
      rs += rl;
      rl >>= 1;
    }
  }
  return FLOCK_OK;
}

int merkle_leaf_batch_device(Ctx *ctx, hipStream_t st, const MerkleJob *d_jobs, uint32_t n_jobs, uint64_t wgs) {
  ScopedPhase ph(ctx, "merkle_leaf_kernel", st);
  hipLaunchKernelGGL(merkle_leaf_batch_kernel, dim3((unsigned)wgs), dim3(256), 0, st, d_jobs, n_jobs);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int merkle_subtree_batch_device(Ctx *ctx, hipStream_t st, const MerkleJob *d_jobs, uint32_t n_jobs, uint64_t wgs) {
  ScopedPhase ph(ctx, "merkle_parents", st);
  hipLaunchKernelGGL(merkle_subtree_batch_kernel, dim3((unsigned)wgs), dim3(256), 0, st, d_jobs, n_jobs);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int gather_roots_device(Ctx *ctx, hipStream_t st, const MerkleJob *d_jobs, uint32_t n_jobs, void *d_roots) {
  hipLaunchKernelGGL(gather_roots_kernel, dim3((n_jobs + 63) / 64), dim3(64), 0, st, d_jobs, n_jobs,
                     (uint4 *)d_roots);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

const char *commit_params_unsupported(const flock_pcs_params *p) {
  // Shapes the kernels support, checked before anything is allocated or
  // launched: 64 interleaved lanes (one wavefront per group), rate 1/2 ..
  // 1/16, at least one NTT layer above the replication, and sizes whose
  // codeword byte count fits 64 bits with room to spare.
  if (p->merkle_hash != 1) return "only the BLAKE3 Merkle hash is implemented";
  if (p->log_batch_size != 6) return "only log_batch_size = 6 (64 interleaved lanes) is implemented";
  if (p->log_inv_rate < 1 || p->log_inv_rate > 4) return "log_inv_rate must be in 1..4";
  if (p->m < 14 || p->m > 40) return "m must be in 14..40";
  return nullptr;
}

int commit_device(Ctx *ctx, const flock_pcs_params *p, const void *d_z, void *d_codeword, void *d_tree,
                  uint8_t root[32], hipStream_t st, hipEvent_t root_ready, const void *d_twiddles) {
  if (p->merkle_hash != 1) {
    ctx->last_error = "only the BLAKE3 Merkle hash is implemented";
    return FLOCK_E_ARG;
  }
  if (const char *why = commit_params_unsupported(p)) {
    ctx->last_error = std::string("commit: ") + why;
    return FLOCK_E_SHAPE;
  }
  const int log_msg = (int)p->m - 7;
  const int log_dim = log_msg - (int)p->log_batch_size;
  const int k = log_dim + (int)p->log_inv_rate;
  ScopedPhase ph(ctx, "commit", st);
  // The first log_inv_rate layers equal replication: the first pass reads the
  // message with its positions masked instead.
  int rc = ntt_forward_interleaved_device(ctx, st, d_codeword, k, 1 << p->log_batch_size, (int)p->log_inv_rate, d_z,
                                          d_twiddles);
  if (rc) return rc;
  rc = merkle_tree_device(ctx, st, d_codeword, 16u << p->log_batch_size, 1ull << k, d_tree);
  if (rc) return rc;
  const uint64_t root_idx = (2ull << k) - 2ull;
  FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(root, (const uint8_t *)d_tree + 32 * root_idx, 32, hipMemcpyDeviceToHost, st));
  ph.end();
  if (root_ready) {
    FLOCK_HIP_CHECK(ctx, hipEventRecord(root_ready, st));
  } else if (int e = ctx->wait_stream(st)) {
    return e;
  }
  return FLOCK_OK;
}

}  // namespace flock
