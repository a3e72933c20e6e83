// extern "C" boundary (include/flock_gpu.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>

#include "../../include/flock_gpu.h"
#include "commit.h"
#include "ctx.h"

struct flock_ctx {
  flock::Ctx c;
};

using flock::Ctx;

extern "C" int flock_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

extern "C" int flock_device_alloc(flock_ctx *ctx, size_t bytes, void **d_ptr) {
  if (!ctx || !d_ptr) return FLOCK_E_ARG;
  *d_ptr = nullptr;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  if (hipMalloc(d_ptr, bytes ? bytes : 1) != hipSuccess) {
    ctx->c.last_error = "hipMalloc failed";
    *d_ptr = nullptr;
    return FLOCK_E_OOM;
  }
  return FLOCK_OK;
}

extern "C" int flock_device_free(flock_ctx *ctx, void *d_ptr) {
  if (!ctx) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipStreamSynchronize(ctx->c.stream));
  FLOCK_HIP_CHECK(&ctx->c, hipFree(d_ptr));
  return FLOCK_OK;
}

extern "C" int flock_memcpy_h2d(flock_ctx *ctx, void *d_dst, const void *src, size_t bytes) {
  if (!ctx || (!d_dst && bytes) || (!src && bytes)) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
  FLOCK_HIP_CHECK(&ctx->c, hipStreamSynchronize(ctx->c.stream));
  return FLOCK_OK;
}

extern "C" int flock_memcpy_d2h(flock_ctx *ctx, void *dst, const void *d_src, size_t bytes) {
  if (!ctx || (!d_src && bytes) || (!dst && bytes)) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
  FLOCK_HIP_CHECK(&ctx->c, hipStreamSynchronize(ctx->c.stream));
  return FLOCK_OK;
}

extern "C" int flock_synchronize(flock_ctx *ctx) {
  if (!ctx) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipStreamSynchronize(ctx->c.stream));
  return FLOCK_OK;
}

extern "C" int flock_ctx_create(int device, flock_ctx **out) {
  if (!out) return FLOCK_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return FLOCK_E_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return FLOCK_E_DEVICE;
  flock_ctx *c = new (std::nothrow) flock_ctx();
  if (!c) return FLOCK_E_OOM;
  c->c.device = device;
  if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FLOCK_E_DEVICE;
  }
  *out = c;
  return FLOCK_OK;
}

extern "C" void flock_ctx_destroy(flock_ctx *ctx) { delete ctx; }

extern "C" const char *flock_last_error(const flock_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

static hipStream_t pick_stream(flock_ctx *ctx, void *stream) {
  return stream ? (hipStream_t)stream : ctx->c.stream;
}

extern "C" int flock_commit_device(flock_ctx *ctx, const flock_pcs_params *params, const void *d_z, void *d_codeword,
                                   void *d_tree, uint8_t root[32], void *stream) {
  if (!ctx || !params || !d_z || !d_codeword || !d_tree || !root) return FLOCK_E_ARG;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  return flock::commit_device(&ctx->c, params, d_z, d_codeword, d_tree, root, pick_stream(ctx, stream));
}

extern "C" int flock_commit(flock_ctx *ctx, const flock_pcs_params *params, const void *z_packed, size_t z_bytes,
                            uint8_t root[32]) {
  if (!ctx || !params || !z_packed || !root) return FLOCK_E_ARG;
  Ctx *c = &ctx->c;
  if (hipSetDevice(c->device) != hipSuccess) return FLOCK_E_DEVICE;
  if (params->m < 7 + params->log_batch_size || params->m > 40) {
    c->last_error = "commit: bad m";
    return FLOCK_E_SHAPE;
  }
  const size_t msg_bytes = (size_t)16 << (params->m - 7);
  if (z_bytes != msg_bytes) {
    c->last_error = "commit: z_packed must be 2^(m-7) F128";
    return FLOCK_E_SHAPE;
  }
  const int k = (int)params->m - 7 - (int)params->log_batch_size + (int)params->log_inv_rate;
  const size_t cw_bytes = msg_bytes << params->log_inv_rate;
  const size_t tree_bytes = (size_t)32 * ((2ull << k) - 1);
  void *dz = c->get_scratch("commit.z", msg_bytes);
  void *dcw = c->get_scratch("commit.codeword", cw_bytes);
  void *dtree = c->get_scratch("commit.tree", tree_bytes);
  if (!dz || !dcw || !dtree) return FLOCK_E_OOM;
  FLOCK_HIP_CHECK(c, hipMemcpyAsync(dz, z_packed, msg_bytes, hipMemcpyHostToDevice, c->stream));
  return flock::commit_device(c, params, dz, dcw, dtree, root, c->stream);
}

extern "C" int flock_timing_enable(flock_ctx *ctx, int on) {
  if (!ctx) return FLOCK_E_ARG;
  ctx->c.resolve_phases();
  ctx->c.phase_totals.clear();
  ctx->c.timing = on != 0;
  return FLOCK_OK;
}

extern "C" int flock_timing_report(flock_ctx *ctx, char *buf, size_t cap) {
  if (!ctx || !buf || cap == 0) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  ctx->c.resolve_phases();
  std::string out;
  char line[256];
  for (auto &kv : ctx->c.phase_totals) {
    snprintf(line, sizeof line, "%s %ld %.6f\n", kv.first.c_str(), kv.second.second, kv.second.first);
    out += line;
  }
  if (out.size() + 1 > cap) return FLOCK_E_CAP;
  memcpy(buf, out.c_str(), out.size() + 1);
  return FLOCK_OK;
}
