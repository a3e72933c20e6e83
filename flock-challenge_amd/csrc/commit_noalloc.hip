// The PCS commit that allocates nothing, for callers that cannot call
// hipMalloc (code under HIP graph capture, serving loops, harnesses that
// forbid it).  It runs the same kernels as commit_device; the only device
// memory the regular path allocates itself, the twiddle table, lives here in
// a fixed file-scope array, one copy per GPU, held until the process ends.
//
// The table is rebuilt by every call: a few kilobytes of subspace rows go
// into the array's tail and one kernel spans the 2^k - 1 twiddles from them,
// so no call depends on what an earlier call left there.  Calls that overlap
// on one GPU share the array and race: callers serialize them.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/flock_gpu.h"
#include "commit.h"
#include "commit_noalloc.h"
#include "ctx.h"
#include "gf128_dev.h"

namespace flock {

namespace {

constexpr size_t TW_ENTRIES = (size_t)1 << COMMIT_NOALLOC_MAX_K;  // >= 2^k - 1 twiddles
// rows of layers 1 .. k-1: layer l has l of them, at l (l - 1) / 2
constexpr size_t ROW_ENTRIES = (size_t)COMMIT_NOALLOC_MAX_K * (COMMIT_NOALLOC_MAX_K - 1) / 2;

// 2^20 + 190 F128 = 16 MiB + 3 KiB per GPU.
__device__ __align__(256) f128 scratch[TW_ENTRIES + ROW_ENTRIES];

std::mutex g_mu;
std::map<int, f128 *> g_scratch;  // device -> address of that GPU's copy

// Twiddle idx = 2^l - 1 + b of the breadth-first table: the sum of the rows
// of layer l selected by the set bits of b (additive_ntt_f128.rs:106-128
// spans the table the same way, one bit at a time).
__global__ __launch_bounds__(256) void twiddle_span_kernel(const f128 *__restrict__ rows, int k,
                                                           f128 *__restrict__ tw) {
  const uint64_t n = (1ull << k) - 1ull;
  for (uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (uint64_t)gridDim.x * blockDim.x) {
    const int l = 63 - __builtin_clzll(idx + 1);
    const uint64_t b = idx + 1 - (1ull << l);
    const f128 *row = rows + (size_t)l * (l - 1) / 2;
    f128 acc = f128_zero();
    for (int i = 0; i < l; ++i)
      if ((b >> i) & 1) f128_xor_into(acc, row[i]);
    tw[idx] = acc;
  }
}

}  // namespace

int commit_noalloc_init(Ctx *ctx) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch.count(ctx->device)) return FLOCK_OK;
  FLOCK_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  // The array's name is a host-side stand-in: the device address comes from
  // the runtime, per device.
  f128 *p = nullptr;
  FLOCK_HIP_CHECK(ctx, hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
  g_scratch[ctx->device] = p;
  return FLOCK_OK;
}

int commit_device_noalloc(Ctx *ctx, const flock_pcs_params *p, const void *d_z, void *d_codeword, void *d_tree,
                          uint8_t root[32], hipStream_t st) {
  if (const char *why = commit_params_unsupported(p)) {
    ctx->last_error = std::string("commit: ") + why;
    return FLOCK_E_SHAPE;
  }
  const int k = (int)p->m - 7 - (int)p->log_batch_size + (int)p->log_inv_rate;
  if (k > COMMIT_NOALLOC_MAX_K) {
    ctx->last_error = "commit (no allocation): k_code " + std::to_string(k) + " exceeds the fixed twiddle array (" +
                      std::to_string(COMMIT_NOALLOC_MAX_K) + ")";
    return FLOCK_E_SHAPE;
  }
  if (int rc = commit_noalloc_init(ctx)) return rc;
  f128 *base;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = g_scratch[ctx->device];
  }
  f128 *tw = base, *rows = base + TW_ENTRIES;
  // rows of layer l (tw(l, b) = sum over set bits i of b of row_l[1 + i]),
  // packed at l (l - 1) / 2; kept alive until the commit's final wait
  const std::vector<std::vector<F128>> ev = ntt_subspace_rows(k);
  std::vector<F128> packed(ROW_ENTRIES);
  for (int l = 1; l < k; ++l)
    for (int i = 0; i < l; ++i) packed[(size_t)l * (l - 1) / 2 + i] = ev[k - l - 1][1 + i];
  const size_t nrows = (size_t)k * (k - 1) / 2;
  if (nrows) FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(rows, packed.data(), nrows * 16, hipMemcpyHostToDevice, st));
  const uint64_t n = (1ull << k) - 1ull;
  const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(twiddle_span_kernel, dim3(grid), dim3(256), 0, st, (const f128 *)rows, k, tw);
  FLOCK_HIP_CHECK(ctx, hipGetLastError());
  // the same kernels as commit_device; it waits for the stream before it
  // returns, so `packed` outlives its copy
  return commit_device(ctx, p, d_z, d_codeword, d_tree, root, st, nullptr, tw);
}

}  // namespace flock
