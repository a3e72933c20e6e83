// BLAKE3 R1CS witness generation on gfx950: one lane per compression, the
// packed (z, a, b) bits streamed out in slot order.
//
// Slot layout per 16,384-bit block (crates/flock-prover/src/r1cs_hashes/
// blake3.rs:178-189): cv | out_lo | const | m | t_lo | t_hi | blen | flags |
// 56 G records of 250 bits | out_hi | zero padding.  Row witnesses follow the
// row semantics of build_matrices (blake3.rs:1116-1125): an ADD x + y stores
// carry_aux = (x^cin)&(y^cin) in z, x^cin in a, y^cin in b (31 bits, cin =
// sum^x^y); input and lin-id slots store (v, v, all-ones); the constant slot
// (1, 1, 1).  c = z (C = I), so no c buffer is written.
#include <hip/hip_runtime.h>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "ctx.h"
#include "prove_kernels.h"

namespace flock {

namespace {

// Three word-aligned bit streams advanced in lockstep.  Finished words go to
// a 16-word LDS ring per stream (one column per lane); every 16 words the
// wavefront writes the ring out so that 16 lanes cover one compression's
// 128 contiguous bytes (four compressions per store instruction).
constexpr int RING = 16;
constexpr int RING_STRIDE = 65;  // words per ring row: 64 lanes + 1 against bank conflicts

struct Writer {
  uint64_t *z, *a, *b;       // global streams, compression 0 of this wavefront
  uint64_t *sz, *sa, *sb;    // LDS rings
  uint32_t lane;
  uint64_t n_valid;          // compressions of this wavefront that exist
  uint64_t pz = 0, pa = 0, pb = 0;
  uint32_t used = 0, word = 0;

  __device__ __forceinline__ void flush(uint32_t first_word) {
    __syncthreads();
    const uint32_t w = lane & 15u;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const uint32_t c = 4u * q + (lane >> 4);
      if (c < n_valid) {
        const uint64_t g = (uint64_t)c * 256 + first_word + w;
        z[g] = sz[w * RING_STRIDE + c];
        a[g] = sa[w * RING_STRIDE + c];
        b[g] = sb[w * RING_STRIDE + c];
      }
    }
    __syncthreads();
  }
  __device__ __forceinline__ void emit(uint64_t vz, uint64_t va, uint64_t vb) {
    const uint32_t slot = word % RING;
    sz[slot * RING_STRIDE + lane] = vz;
    sa[slot * RING_STRIDE + lane] = va;
    sb[slot * RING_STRIDE + lane] = vb;
    if (slot == RING - 1) flush(word - (RING - 1));
    ++word;
  }
  __device__ __forceinline__ void push(uint32_t vz, uint32_t va, uint32_t vb, uint32_t w) {
    const uint64_t mask = (w == 32) ? 0xFFFFFFFFull : ((1ull << w) - 1ull);
    uint64_t xz = (uint64_t)vz & mask, xa = (uint64_t)va & mask, xb = (uint64_t)vb & mask;
    pz |= xz << used;
    pa |= xa << used;
    pb |= xb << used;
    used += w;
    if (used >= 64) {
      emit(pz, pa, pb);
      used -= 64;
      const uint32_t sh = w - used;  // bits of this value already written
      pz = used ? xz >> sh : 0;
      pa = used ? xa >> sh : 0;
      pb = used ? xb >> sh : 0;
    }
  }
  __device__ __forceinline__ void lin(uint32_t v) { push(v, v, 0xFFFFFFFFu, 32); }
  __device__ __forceinline__ uint32_t add(uint32_t x, uint32_t y) {
    const uint32_t s = x + y;
    const uint32_t cin = s ^ x ^ y;
    const uint32_t l = x ^ cin, r = y ^ cin;
    push(l & r, l, r, 31);
    return s;
  }
};

__device__ __forceinline__ uint32_t rotr(uint32_t x, uint32_t n) { return __builtin_amdgcn_alignbit(x, x, n); }

// Message word of operand mx (0) / my (1) of G g in round r: PERM^r schedule.
constexpr uint8_t MSG_IDX[7][8][2] = {
    {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {8, 9}, {10, 11}, {12, 13}, {14, 15}},
    {{2, 6}, {3, 10}, {7, 0}, {4, 13}, {1, 11}, {12, 5}, {9, 14}, {15, 8}},
    {{3, 4}, {10, 12}, {13, 2}, {7, 14}, {6, 5}, {9, 0}, {11, 15}, {8, 1}},
    {{10, 7}, {12, 9}, {14, 3}, {13, 15}, {4, 0}, {11, 2}, {5, 8}, {1, 6}},
    {{12, 13}, {9, 11}, {15, 10}, {14, 8}, {7, 2}, {5, 3}, {0, 1}, {6, 4}},
    {{9, 14}, {11, 5}, {8, 12}, {15, 1}, {13, 3}, {0, 10}, {2, 6}, {4, 7}},
    {{11, 15}, {5, 0}, {1, 9}, {8, 6}, {14, 10}, {2, 12}, {3, 4}, {7, 13}},
};
constexpr uint8_t G_LANES_D[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                        {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

}  // namespace

// The 64 compressions from wg * 64 of one instance, one per lane.
__device__ __forceinline__ void witness_body(const flock_compression *__restrict__ blocks, uint64_t n_real,
                                             uint64_t n_total, uint64_t *__restrict__ z, uint64_t *__restrict__ a,
                                             uint64_t *__restrict__ b, uint64_t wg,
                                             uint64_t (&ring)[3][RING * RING_STRIDE]) {
  const uint32_t lane = threadIdx.x;
  const uint64_t blk0 = wg * 64;
  const uint64_t blk = blk0 + lane;
  uint32_t cv[8], m[16], tlo = 0, thi = 0, blen = 0, flags = 0;
  if (blk < n_real) {
    const flock_compression &c = blocks[blk];
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = c.cv[i];
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = c.m[i];
    tlo = (uint32_t)c.counter;
    thi = (uint32_t)(c.counter >> 32);
    blen = c.block_len;
    flags = c.flags;
  } else {
    // Padding slots hold the compression of the all-zero input (blake3.rs:2954).
#pragma unroll
    for (int i = 0; i < 8; ++i) cv[i] = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) m[i] = 0;
  }
  // out_lo (slot words 4..7) first, so every word is produced in order.
  uint32_t out[16];
  b3::compress_full(cv, m, ((uint64_t)thi << 32) | tlo, blen, flags, out);
  Writer w;
  w.z = z + blk0 * 256;
  w.a = a + blk0 * 256;
  w.b = b + blk0 * 256;
  w.sz = ring[0];
  w.sa = ring[1];
  w.sb = ring[2];
  w.lane = lane;
  w.n_valid = n_total - blk0 < 64 ? n_total - blk0 : 64;
#pragma unroll
  for (int i = 0; i < 8; ++i) w.lin(cv[i]);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint64_t lo = (uint64_t)out[2 * q] | ((uint64_t)out[2 * q + 1] << 32);
    w.emit(lo, lo, ~0ull);
  }
  w.push(1, 1, 1, 1);
#pragma unroll
  for (int i = 0; i < 16; ++i) w.lin(m[i]);
  w.lin(tlo);
  w.lin(thi);
  w.lin(blen);
  w.lin(flags);
  uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    b3::IV[0], b3::IV[1], b3::IV[2], b3::IV[3], tlo, thi, blen, flags};
#pragma unroll
  for (int r = 0; r < 7; ++r) {
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const int la = G_LANES_D[g][0], lb = G_LANES_D[g][1], lc = G_LANES_D[g][2], ld = G_LANES_D[g][3];
      const uint32_t mx = m[MSG_IDX[r][g][0]], my = m[MSG_IDX[r][g][1]];
      uint32_t A = s[la], B = s[lb], C = s[lc], D = s[ld];
      const uint32_t t0 = w.add(A, B);
      const uint32_t a1 = w.add(t0, mx);
      const uint32_t d1 = rotr(D ^ a1, 16);
      const uint32_t c1 = w.add(C, d1);
      const uint32_t b1 = rotr(B ^ c1, 12);
      const uint32_t t1 = w.add(a1, b1);
      const uint32_t a2 = w.add(t1, my);
      const uint32_t d2 = rotr(d1 ^ a2, 8);
      const uint32_t c2 = w.add(c1, d2);
      const uint32_t bn = rotr(b1 ^ c2, 7);
      w.lin(bn);
      w.lin(d2);
      s[la] = a2;
      s[lb] = bn;
      s[lc] = c2;
      s[ld] = d2;
    }
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) w.lin(s[i + 8] ^ cv[i]);  // out_hi
  // the partial word, then zero padding up to 256 words
  if (w.used) {
    w.emit(w.pz, w.pa, w.pb);
    w.used = 0;
  }
#pragma unroll
  for (int k = 0; k < 256; ++k)
    if (k >= (int)w.word) w.emit(0, 0, 0);
}

__global__ __launch_bounds__(64) void witness_kernel(const flock_compression *__restrict__ blocks, uint64_t n_real,
                                                     uint64_t n_total, uint64_t *__restrict__ z,
                                                     uint64_t *__restrict__ a, uint64_t *__restrict__ b) {
  __shared__ uint64_t ring[3][RING * RING_STRIDE];
  witness_body(blocks, n_real, n_total, z, a, b, blockIdx.x, ring);
}

// Many instances in one launch: workgroup w belongs to the last job whose
// first workgroup is at most w (jobs sorted by wg_begin).
__global__ __launch_bounds__(64) void witness_batch_kernel(const WitnessJob *__restrict__ jobs, uint32_t n_jobs) {
  __shared__ uint64_t ring[3][RING * RING_STRIDE];
  uint32_t j = 0;
  for (uint32_t i = 1; i < n_jobs; ++i)
    if (jobs[i].wg_begin <= blockIdx.x) j = i;
  const WitnessJob J = jobs[j];
  witness_body(J.blocks, J.n_real, J.n_total, J.z, J.a, J.b, blockIdx.x - J.wg_begin, ring);
}

int witness_batch_device(Ctx *ctx, hipStream_t st, const WitnessJob *d_jobs, uint32_t n_jobs, uint64_t total_wgs) {
  ScopedPhase ph(ctx, "witness", st);
  hipLaunchKernelGGL(witness_batch_kernel, dim3((unsigned)total_wgs), dim3(64), 0, st, d_jobs, n_jobs);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

int witness_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_real, int n_log, void *d_z,
                   void *d_a, void *d_b) {
  ScopedPhase ph(ctx, "witness", st);
  const uint64_t n_total = 1ull << n_log;
  hipLaunchKernelGGL(witness_kernel, dim3((unsigned)((n_total + 63) / 64)), dim3(64), 0, st, d_blocks, n_real,
                     n_total, (uint64_t *)d_z, (uint64_t *)d_a, (uint64_t *)d_b);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock

namespace flock {

// generate_compressions (benchmark-tools/common/src/lib.rs:20-33): one
// SplitMix64 stream from seed ^ rotl(log2, 29), 25 low-32-bit outputs per
// compression (8 cv, 16 message, counter).  SplitMix64's state advances by a
// constant, so output k is mix(s0 + k * gamma) and every compression is
// independent.
__device__ __forceinline__ uint32_t splitmix_at(uint64_t s0, uint64_t k) {
  uint64_t z = s0 + k * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)(z ^ (z >> 31));
}

__global__ void generate_kernel(uint64_t s0, uint64_t n, flock_compression *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flock_compression c;
  const uint64_t k0 = 25 * i + 1;
#pragma unroll
  for (int w = 0; w < 8; ++w) c.cv[w] = splitmix_at(s0, k0 + w);
#pragma unroll
  for (int w = 0; w < 16; ++w) c.m[w] = splitmix_at(s0, k0 + 8 + w);
  c.counter = splitmix_at(s0, k0 + 24);
  c.block_len = 64;
  c.flags = 11;
  out[i] = c;
}

int generate_compressions_device(Ctx *ctx, hipStream_t st, uint32_t log2, uint64_t seed, flock_compression *d_out) {
  const uint64_t n = 1ull << log2;
  const uint64_t l = log2;
  const uint64_t s0 = seed ^ ((l << 29) | (l >> 35));
  hipLaunchKernelGGL(generate_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s0, n, d_out);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  return FLOCK_OK;
}

}  // namespace flock
