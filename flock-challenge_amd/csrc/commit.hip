// PCS commit on gfx950: replicate -> interleaved additive NTT -> BLAKE3 Merkle.
//
// Layout in HBM (the reference's, crates/flock-core/src/pcs/commit.rs:235-245):
// codeword[pos * NL + lane] F128, NL = 2^log_batch_size interleaved RS lanes;
// Merkle leaf p = the NL*16 contiguous bytes of position p.
//
// NTT pass kernel.  A pass applies F consecutive layers [l, l+F) of the
// neighbors-last forward transform (ntt/additive_ntt_f128.rs:1281-1320).
// The 2^F positions that exchange data in those layers share every position
// bit except F contiguous ones, so they form a "group"
//     pos(e) = (B << (k-l)) | (e << s) | r,   s = k - l - F,  e in [0, 2^F).
// One wavefront owns one group for NL = 64: lane = RS lane, and each lane
// keeps its 2^F elements in VGPRs, so every butterfly is lane-local and the
// global loads/stores are 1 KiB-contiguous per position (fully coalesced).
// The twiddle of a butterfly depends only on (layer, block), i.e. it is
// wavefront-uniform: its products run through lookup tables of that twiddle
// in LDS, full byte tables shared by a workgroup (fulltab.h) where the
// twiddle meets many products, per-wavefront nibble tables (nibtab.h) in the
// last layers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "commit.h"
#include "ctx.h"
#include "fulltab.h"
#include "gf128_dev.h"
#include "nibtab.h"

namespace flock {

// Tail pass: F layers whose twiddles are shared by too few groups to pay for
// full byte tables (the last layers of the transform, where a twiddle meets
// 64 .. 512 products).  One wavefront per group; it builds the 8 KiB nibble
// tables of its own twiddles, two at a time, in 16 KiB of LDS, so ten
// wavefronts fit on a CU.  Step (J, C) handles twiddles C, C + 1 of layer
// offset J; template recursion keeps every register index constant.
template <int F, int J, int C>
__device__ __forceinline__ void tail_pass_step(f128 (&v)[1 << F], const f128 *__restrict__ tw, int l, uint64_t B,
                                               char *lds, const NibtabLane &NL, uint32_t lane) {
  constexpr int NT = (1 << J) - C >= 2 ? 2 : 1;
  constexpr int nb = 1 << (F - 1 - J);
  const uint64_t tb = (1ull << (l + J)) - 1ull + (B << J);
  __syncthreads();  // the previous step's tables are read
  build_nibtab<0>(tw[tb + C], lds, lane);
  if constexpr (NT == 2) build_nibtab<NIBTAB_BYTES>(tw[tb + C + 1], lds, lane);
  __syncthreads();
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int el = 0; el < nb; ++el) {
    const int e0 = (C << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
    const f128 u0 = f128_xor(v[e0], f128_mul_nibtab<0>(v[e1], lds, NL));
    v[e0] = u0;
    f128_xor_into(v[e1], u0);
    __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (NT == 2) {
#pragma unroll
    for (int el = 0; el < nb; ++el) {
      const int e0 = ((C + 1) << (F - J)) | el, e1 = e0 | (1 << (F - 1 - J));
      const f128 u0 = f128_xor(v[e0], f128_mul_nibtab<NIBTAB_BYTES>(v[e1], lds, NL));
      v[e0] = u0;
      f128_xor_into(v[e1], u0);
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if constexpr (C + NT < (1 << J))
    tail_pass_step<F, J, C + NT>(v, tw, l, B, lds, NL, lane);
  else if constexpr (J + 1 < F)
    tail_pass_step<F, J + 1, 0>(v, tw, l, B, lds, NL, lane);
}

template <int F>
__device__ __forceinline__ void ntt_tail_body(f128 *__restrict__ data, const f128 *__restrict__ tw, int k, int l,
                                              uint64_t g, char *lds) {
  const uint32_t lane = threadIdx.x;
  const NibtabLane NL(lane & 15u);
  const int s = k - l - F;
  const uint64_t r = g & ((1ull << s) - 1ull);
  const uint64_t B = g >> s;
  const uint64_t base = (B << (k - l)) | r;
  f128 v[1 << F];
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) v[e] = data[((base | ((uint64_t)e << s)) << 6) + lane];
  tail_pass_step<F, 0, 0>(v, tw, l, B, lds, NL, lane);
#pragma unroll
  for (int e = 0; e < (1 << F); ++e) data[((base | ((uint64_t)e << s)) << 6) + lane] = v[e];
}

template <int F>
__global__ __launch_bounds__(64) void ntt_tail_kernel(f128 *__restrict__ data, const f128 *__restrict__ tw, int k,
                                                      int l) {
  __shared__ __align__(16) char lds[2 * NIBTAB_BYTES];
  ntt_tail_body<F>(data, tw, k, l, blockIdx.x, lds);
}

// The job a workgroup of a fused launch belongs to: the last one whose first
// workgroup is at most blockIdx.x (jobs sorted by wg_begin).
template <class Job>
__device__ __forceinline__ const Job &job_of(const Job *__restrict__ jobs, uint32_t n_jobs) {
  uint32_t j = 0;
  for (uint32_t i = 1; i < n_jobs; ++i)
    if (jobs[i].wg_begin <= blockIdx.x) j = i;
  return jobs[j];
}

template <int F>
__global__ __launch_bounds__(64) void ntt_tail_batch_kernel(const NttJob *__restrict__ jobs, uint32_t n_jobs) {
  __shared__ __align__(16) char lds[2 * NIBTAB_BYTES];
  const NttJob &J = job_of(jobs, n_jobs);
  ntt_tail_body<F>(J.data, J.tw, J.k, J.l, blockIdx.x - J.wg_begin, lds);
}

// Shared-table pass.  When 2^s >= WAVES groups share the block B of the pass's
// first layer, they share all 2^F - 1 twiddles of the pass.  A workgroup of
// WAVES wavefronts takes WAVES such groups.  Up front its threads compute the
// 128 values t x^i of every twiddle t of the pass; then for each twiddle they
// build the full byte tables T_k[b] = (b x^(8k)) t, k = 0..15 (64 KiB), and a
// product v t is 16 independent lookups, one per byte of v, with no shifts.
// Two table buffers alternate, so one barrier per twiddle suffices: the
// buffer being rebuilt was last read before the previous barrier.
constexpr int SHARED_TWIDDLES = 15;  // 2^F - 1 for F = 4

// Twiddle n = 2^J - 1 + C of a shared-table pass (layer offset J, block C):
// build its tables into buffer n & 1, then every butterfly of it in each of
// the wave's G groups, two products in flight per lane.  Template recursion
// keeps every register index a compile-time constant.
template <int F, int G, int THREADS, int J, int C>
__device__ __forceinline__ void shared_pass_step(f128 (&v)[G][1 << F], f128 *tabs, const f128 *basis,
                                                 const FulltabLane &FL, uint32_t tid) {
  constexpr int N = (1 << J) - 1 + C;
  constexpr bool HI = N & 1;
  const char *tbase = (const char *)tabs;  // buffer 0 at +0, buffer 1 at +64 KiB
  build_fulltab<THREADS>(basis + N * 128, tabs + (HI ? 4096 : 0), tid);
  __syncthreads();
  constexpr int nb = 1 << (F - 1 - J);
  // product slots (group, butterfly) in pairs
  constexpr int SLOTS = G * nb;
#pragma unroll
  for (int sl = 0; sl < SLOTS; sl += 2) {
    const int ga = sl / nb, la = sl % nb, gb = (sl + 1) / nb, lb = (sl + 1) % nb;
    const int a0 = (C << (F - J)) | la, a1 = a0 | (1 << (F - 1 - J));
    if (sl + 1 < SLOTS) {
      const int b0 = (C << (F - J)) | lb, b1 = b0 | (1 << (F - 1 - J));
      f128 p0, p1;
      f128_mul2_fulltab<HI>(v[ga][a1], v[gb][b1], tbase, FL, p0, p1);
      const f128 u0 = f128_xor(v[ga][a0], p0), u1 = f128_xor(v[gb][b0], p1);
      v[ga][a0] = u0;
      f128_xor_into(v[ga][a1], u0);
      v[gb][b0] = u1;
      f128_xor_into(v[gb][b1], u1);
    } else {
      const f128 u0 = f128_xor(v[ga][a0], f128_mul_fulltab<HI>(v[ga][a1], tbase, FL));
      v[ga][a0] = u0;
      f128_xor_into(v[ga][a1], u0);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  if constexpr (C + 1 < (1 << J))
    shared_pass_step<F, G, THREADS, J, C + 1>(v, tabs, basis, FL, tid);
  else if constexpr (J + 1 < F)
    shared_pass_step<F, G, THREADS, J + 1, 0>(v, tabs, basis, FL, tid);
}

// G groups per wavefront: a workgroup takes WAVES * G groups that share
// block B (so 2^s >= WAVES * G), and each table serves all of them.
constexpr int SHARED_LDS_F128 = 2 * 4096 + SHARED_TWIDDLES * 128;

template <int F, int WAVES, int G>
__device__ __forceinline__ void ntt_pass_shared_body(f128 *__restrict__ data, const f128 *__restrict__ tw, int k,
                                                     int l, const f128 *__restrict__ src, uint64_t src_pos_mask,
                                                     uint64_t wg, f128 *lds) {
  constexpr uint32_t THREADS = 64 * WAVES;
  static_assert((1 << F) - 1 <= SHARED_TWIDDLES, "twiddle basis area");
  f128 *tabs = lds;                // 2 x 16 x 256
  f128 *basis = lds + 2 * 4096;    // SHARED_TWIDDLES x 128
  const uint32_t tid = threadIdx.x, lane = tid & 63, rot = lane & 15;
  const FulltabLane FL(rot);
  const int s = k - l - F;
  // wavefront-uniform group indices (scalar registers for the addresses)
  const uint64_t g0 = (wg * WAVES + __builtin_amdgcn_readfirstlane(tid >> 6)) * G;
  const uint64_t B = g0 >> s;  // the same for the whole workgroup
  // basis of every twiddle: entry n * 128 + i * 16 + kb = t_n x^(8 kb + i),
  // t_n = twiddle n of the pass in (layer offset j, block c) order,
  // n = 2^j - 1 + c
  for (uint32_t e = tid; e < ((1u << F) - 1u) * 128u; e += THREADS) {
    const uint32_t n = e >> 7, i = (e >> 4) & 7, kb = e & 15;
    const uint32_t j = 31 - __builtin_clz(n + 1);
    const uint32_t c = n + 1 - (1u << j);
    basis[e] = f128_mul_xpow(tw[(1ull << (l + j)) - 1ull + (B << j) + c], 8 * kb + i);
  }
  // The first pass of a commit reads the message itself: position p of the
  // replicated codeword holds message position p & src_pos_mask.
  f128 v[G][1 << F];
  uint64_t base[G];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) {
    const uint64_t g = g0 + gi;
    base[gi] = (B << (k - l)) | (g & ((1ull << s) - 1ull));
#pragma unroll
    for (int e = 0; e < (1 << F); ++e)
      v[gi][e] = src[(((base[gi] | ((uint64_t)e << s)) & src_pos_mask) << 6) + lane];
  }
  __syncthreads();  // basis visible
  shared_pass_step<F, G, THREADS, 0, 0>(v, tabs, basis, FL, tid);
#pragma unroll
  for (int gi = 0; gi < G; ++gi)
#pragma unroll
    for (int e = 0; e < (1 << F); ++e) data[((base[gi] | ((uint64_t)e << s)) << 6) + lane] = v[gi][e];
}

template <int F, int WAVES, int G>
__global__ __launch_bounds__(64 * WAVES) void ntt_pass_shared_kernel(f128 *__restrict__ data,
                                                                     const f128 *__restrict__ tw, int k, int l,
                                                                     const f128 *__restrict__ src,
                                                                     uint64_t src_pos_mask) {
  __shared__ f128 lds[SHARED_LDS_F128];
  ntt_pass_shared_body<F, WAVES, G>(data, tw, k, l, src, src_pos_mask, blockIdx.x, lds);
}

template <int F, int WAVES, int G>
__global__ __launch_bounds__(64 * WAVES) void ntt_pass_shared_batch_kernel(const NttJob *__restrict__ jobs,
                                                                           uint32_t n_jobs) {
  __shared__ f128 lds[SHARED_LDS_F128];
  const NttJob &J = job_of(jobs, n_jobs);
  ntt_pass_shared_body<F, WAVES, G>(J.data, J.tw, J.k, J.l, J.src ? J.src : J.data, J.src ? J.src_pos_mask : ~0ull,
                                    blockIdx.x - J.wg_begin, lds);
}

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

static bool ntt_pass_is_shared(int k, int l, int F) { return F == 4 && k - l - F >= 3; }

// One pass of F layers from layer l.  src (may be null) is a message that the
// pass reads instead of data, replicated by masking positions with
// src_pos_mask (shared passes only).
static int ntt_pass(Ctx *ctx, hipStream_t st, f128 *data, const f128 *tw, int k, int l, int F,
                    const f128 *src = nullptr, uint64_t src_pos_mask = ~0ull) {
  // 16 wavefronts share one set of tables when the block has that many
  // groups (s >= 4), else 8.
  if (ntt_pass_is_shared(k, l, F)) {
    ScopedPhase ph(ctx, "ntt_pass_shared_kernel", st);
    const uint64_t groups = 1ull << (k - F);
    const f128 *in = src ? src : data;
    const uint64_t mask = src ? src_pos_mask : ~0ull;
    // 8 waves (a lane's VGPR budget holds two groups of 16 elements); two
    // groups per wave when 16 groups share the block, so each table serves
    // twice the products
    if (k - l - F >= 4)
      hipLaunchKernelGGL((ntt_pass_shared_kernel<4, 8, 2>), dim3((unsigned)(groups / 16)), dim3(512), 0, st, data,
                         tw, k, l, in, mask);
    else
      hipLaunchKernelGGL((ntt_pass_shared_kernel<4, 8, 1>), dim3((unsigned)(groups / 8)), dim3(512), 0, st, data,
                         tw, k, l, in, mask);
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
    return FLOCK_OK;
  }
  ScopedPhase ph(ctx, "ntt_tail_kernel", st);
  uint64_t groups = 1ull << (k - F);
  dim3 grid((unsigned)groups), block(64);
  switch (F) {
    case 1: hipLaunchKernelGGL(ntt_tail_kernel<1>, grid, block, 0, st, data, tw, k, l); break;
    case 2: hipLaunchKernelGGL(ntt_tail_kernel<2>, grid, block, 0, st, data, tw, k, l); break;
    case 3: hipLaunchKernelGGL(ntt_tail_kernel<3>, grid, block, 0, st, data, tw, k, l); break;
    case 4: hipLaunchKernelGGL(ntt_tail_kernel<4>, grid, block, 0, st, data, tw, k, l); break;
    default: ctx->last_error = "ntt_pass: unsupported F"; return FLOCK_E_ARG;
  }
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int ntt_forward_interleaved_device(Ctx *ctx, hipStream_t st, void *d_data, int k, int num_lanes, int start_layer,
                                   int max_fuse, const void *d_msg, const void *d_twiddles) {
  if (num_lanes != 64) {
    ctx->last_error = "ntt: only 64 interleaved lanes are implemented";
    return FLOCK_E_SHAPE;
  }
  const F128 *tw = d_twiddles ? (const F128 *)d_twiddles : ctx->device_twiddles(k);
  if (!tw) return FLOCK_E_DEVICE;
  int l = start_layer;
  bool first = true;
  while (l < k) {
    int F = std::min(max_fuse, k - l);
    int rc = (first && d_msg)
                 ? ntt_pass(ctx, st, (f128 *)d_data, (const f128 *)tw, k, l, F, (const f128 *)d_msg,
                            (1ull << (k - start_layer)) - 1)
                 : ntt_pass(ctx, st, (f128 *)d_data, (const f128 *)tw, k, l, F);
    if (rc) return rc;
    first = false;
    l += F;
  }
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

NttVariant ntt_pass_variant(int k, int l, int F, uint64_t *wgs) {
  const uint64_t groups = 1ull << (k - F);
  if (ntt_pass_is_shared(k, l, F)) {
    const bool g2 = k - l - F >= 4;
    if (wgs) *wgs = groups / (g2 ? 16 : 8);
    return g2 ? NTT_SHARED_G2 : NTT_SHARED_G1;
  }
  if (wgs) *wgs = groups;
  return (NttVariant)(NTT_TAIL1 + (F - 1));
}

int ntt_batch_device(Ctx *ctx, hipStream_t st, NttVariant v, const NttJob *d_jobs, uint32_t n_jobs, uint64_t wgs) {
  const dim3 grid((unsigned)wgs);
  switch (v) {
    case NTT_SHARED_G2: {
      ScopedPhase ph(ctx, "ntt_pass_shared_kernel", st);
      hipLaunchKernelGGL((ntt_pass_shared_batch_kernel<4, 8, 2>), grid, dim3(512), 0, st, d_jobs, n_jobs);
      break;
    }
    case NTT_SHARED_G1: {
      ScopedPhase ph(ctx, "ntt_pass_shared_kernel", st);
      hipLaunchKernelGGL((ntt_pass_shared_batch_kernel<4, 8, 1>), grid, dim3(512), 0, st, d_jobs, n_jobs);
      break;
    }
    case NTT_TAIL1: hipLaunchKernelGGL(ntt_tail_batch_kernel<1>, grid, dim3(64), 0, st, d_jobs, n_jobs); break;
    case NTT_TAIL2: hipLaunchKernelGGL(ntt_tail_batch_kernel<2>, grid, dim3(64), 0, st, d_jobs, n_jobs); break;
    case NTT_TAIL3: hipLaunchKernelGGL(ntt_tail_batch_kernel<3>, grid, dim3(64), 0, st, d_jobs, n_jobs); break;
    case NTT_TAIL4: hipLaunchKernelGGL(ntt_tail_batch_kernel<4>, grid, dim3(64), 0, st, d_jobs, n_jobs); break;
    default: ctx->last_error = "ntt batch: bad variant"; return FLOCK_E_ARG;
  }
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
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
  const uint64_t msg_len = 1ull << log_msg;
  const uint64_t cw_len = msg_len << p->log_inv_rate;
  ScopedPhase ph(ctx, "commit", st);
  // The first layers equal replication; a shared first pass reads the
  // message directly, otherwise replicate explicitly.
  const int rate = (int)p->log_inv_rate;
  const bool fused = rate < k && ntt_pass_is_shared(k, rate, std::min(4, k - rate));
  if (!fused) {
    hipLaunchKernelGGL(replicate_kernel, dim3(2048), dim3(256), 0, st, (const uint4 *)d_z, (uint4 *)d_codeword,
                       msg_len, cw_len);
    FLOCK_HIP_CHECK(ctx, hipGetLastError());
  }
  int rc = ntt_forward_interleaved_device(ctx, st, d_codeword, k, 1 << p->log_batch_size, rate, 4,
                                          fused ? d_z : nullptr, d_twiddles);
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
