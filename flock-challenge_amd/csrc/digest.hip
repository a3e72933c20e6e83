// The statement digest's spine on the GPU (stmt_digest.h): the framing of
// A_0, B_0 and C_0 written into HBM as u64 words from the matrices' rows,
// one lane per 1 KiB BLAKE3 chunk for the chunk chaining values, then one
// launch per tree level for the parents, keeping each level's node 1.
#include <hip/hip_runtime.h>
#include <string.h>

#include <vector>

#include "../../include/flock_gpu.h"
#include "blake3_dev.h"
#include "ctx.h"
#include "r1cs.h"
#include "stmt_digest.h"

namespace flock {
namespace r1cs {

namespace {

// Words of one matrix's framing from word `base`: rows, cols, then per row
// its length and its column indices.
__global__ __launch_bounds__(256) void frame_rows_kernel(const uint32_t *__restrict__ ptr,
                                                         const uint32_t *__restrict__ cols, uint32_t rows,
                                                         uint64_t base, uint64_t *__restrict__ w) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r == 0) {
    w[base] = rows;
    w[base + 1] = rows;
  }
  if (r >= rows) return;
  const uint32_t b = ptr[r], e = ptr[r + 1];
  uint64_t *o = w + base + 2 + b + r;
  o[0] = e - b;
  for (uint32_t k = b; k < e; ++k) o[1 + k - b] = cols[k];
}

// C_0 = I: rows, cols, then (1, i) per row.
__global__ __launch_bounds__(256) void frame_identity_kernel(uint32_t rows, uint64_t base, uint64_t *__restrict__ w) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    w[base] = rows;
    w[base + 1] = rows;
  }
  if (i >= rows) return;
  w[base + 2 + 2 * (uint64_t)i] = 1;
  w[base + 3 + 2 * (uint64_t)i] = i;
}

// Chaining value of chunk j >= 1 of header || body, i.e. body bytes
// [1024 j - 43, ..).  The body sits in HBM as dwords; 1024 j - 43 is one byte
// past a dword boundary, so each message word is the funnel shift of two
// neighbouring dwords.  Bytes past the body's end read as zero (the buffer
// is zero-padded), so the last block comes out zero-padded as BLAKE3 wants.
__global__ __launch_bounds__(256) void chunk_cv_kernel(const uint32_t *__restrict__ body, uint64_t total_bytes,
                                                       uint64_t n_chunks, uint4 *__restrict__ cvs) {
  const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1;
  if (j >= n_chunks) return;
  const uint64_t off = 1024 * j - DIGEST_HEADER_BYTES;  // body byte offset, = 1 (mod 4)
  const uint32_t *d = body + (off >> 2);
  const uint64_t left = total_bytes - 1024 * j;
  const uint32_t len = left < 1024 ? (uint32_t)left : 1024u;
  const uint32_t blocks = (len + 63) / 64;
  uint32_t cv[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) cv[i] = b3::IV[i];
  for (uint32_t b = 0; b < blocks; ++b) {
    uint32_t m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = __builtin_amdgcn_alignbit(d[16 * b + k + 1], d[16 * b + k], 8);
    const uint32_t bl = b + 1 == blocks ? len - 64 * b : 64u;
    const uint32_t flags = (b == 0 ? b3::CHUNK_START : 0u) | (b + 1 == blocks ? b3::CHUNK_END : 0u);
    b3::compress_cv(cv, m, j, bl, flags);
  }
  cvs[2 * j] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
  cvs[2 * j + 1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
}

// One tree level: out[i] = parent(in[2i], in[2i+1]) for i >= 1 (node 0 is
// on the m-dependent left edge and is never used), the odd last node carried.
__global__ __launch_bounds__(256) void parent_level_kernel(const uint4 *__restrict__ in, uint64_t n,
                                                           uint4 *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t pairs = n / 2;
  if (i < pairs && i > 0) {
    uint32_t m[16], cv[8];
    const uint4 *p = in + 4 * i;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      m[4 * q] = p[q].x;
      m[4 * q + 1] = p[q].y;
      m[4 * q + 2] = p[q].z;
      m[4 * q + 3] = p[q].w;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) cv[k] = b3::IV[k];
    b3::compress_cv(cv, m, 0, 64, b3::PARENT);
    out[2 * i] = make_uint4(cv[0], cv[1], cv[2], cv[3]);
    out[2 * i + 1] = make_uint4(cv[4], cv[5], cv[6], cv[7]);
  } else if (i == pairs && (n & 1)) {
    out[2 * i] = in[2 * (n - 1)];
    out[2 * i + 1] = in[2 * (n - 1) + 1];
  }
}

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

}  // namespace

bool spine_device(Ctx *ctx, Spine &out) {
  if (hipSetDevice(ctx->device) != hipSuccess) return false;
  const Matrices &mx = matrices();
  const uint64_t na = mx.a.cols.size(), nb = mx.b.cols.size();
  const uint64_t words = digest_body_word_count();
  const uint64_t total = DIGEST_HEADER_BYTES + 8 * words;
  const uint64_t n = (total + 1023) / 1024;
  const size_t body_bytes = 8 * words + 4096;  // zero padding past the end
  DevBuf body, csr, lev0, lev1;
  const size_t csr_words = 2 * (K + 1) + na + nb;
  if (hipMalloc(&body.p, body_bytes) != hipSuccess || hipMalloc(&csr.p, 4 * csr_words) != hipSuccess ||
      hipMalloc(&lev0.p, 32 * n) != hipSuccess || hipMalloc(&lev1.p, 32 * (n / 2 + 1)) != hipSuccess)
    return false;
  hipStream_t st = ctx->stream;
  uint32_t *dp = (uint32_t *)csr.p;
  uint32_t *a_ptr = dp, *a_cols = a_ptr + K + 1, *b_ptr = a_cols + na, *b_cols = b_ptr + K + 1;
  bool ok = hipMemsetAsync(body.p, 0, body_bytes, st) == hipSuccess &&
            hipMemcpyAsync(a_ptr, mx.a.ptr.data(), 4 * (K + 1), hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(a_cols, mx.a.cols.data(), 4 * na, hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(b_ptr, mx.b.ptr.data(), 4 * (K + 1), hipMemcpyHostToDevice, st) == hipSuccess &&
            hipMemcpyAsync(b_cols, mx.b.cols.data(), 4 * nb, hipMemcpyHostToDevice, st) == hipSuccess;
  if (!ok) return false;
  uint64_t *w = (uint64_t *)body.p;
  const unsigned rg = (K + 255) / 256;
  hipLaunchKernelGGL(frame_rows_kernel, dim3(rg), dim3(256), 0, st, a_ptr, a_cols, (uint32_t)K, 0ull, w);
  hipLaunchKernelGGL(frame_rows_kernel, dim3(rg), dim3(256), 0, st, b_ptr, b_cols, (uint32_t)K, 2 + K + na, w);
  hipLaunchKernelGGL(frame_identity_kernel, dim3(rg), dim3(256), 0, st, (uint32_t)K, 4 + 2 * (uint64_t)K + na + nb,
                     w);
  hipLaunchKernelGGL(chunk_cv_kernel, dim3((unsigned)((n - 1 + 255) / 256)), dim3(256), 0, st,
                     (const uint32_t *)body.p, total, n, (uint4 *)lev0.p);
  if (hipGetLastError() != hipSuccess) return false;
  Spine sp;
  sp.n_chunks = n;
  std::vector<std::array<uint32_t, 8>> sib;
  void *in = lev0.p, *outp = lev1.p;
  for (uint64_t len = n; len >= 2; len = (len + 1) / 2) {
    std::array<uint32_t, 8> s;
    if (hipMemcpyAsync(s.data(), (const uint8_t *)in + 32, 32, hipMemcpyDeviceToHost, st) != hipSuccess ||
        ctx->wait_stream(st))
      return false;
    sib.push_back(s);
    if (len == 2) break;
    const uint64_t outs = len / 2 + (len & 1);
    hipLaunchKernelGGL(parent_level_kernel, dim3((unsigned)((outs + 255) / 256)), dim3(256), 0, st,
                       (const uint4 *)in, len, (uint4 *)outp);
    if (hipGetLastError() != hipSuccess) return false;
    std::swap(in, outp);
  }
  if (ctx->wait_stream(st)) return false;
  sp.sib = sib;
  out = sp;
  return true;
}

}  // namespace r1cs
}  // namespace flock
