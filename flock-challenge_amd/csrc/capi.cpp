// extern "C" boundary (include/flock_gpu.h).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <new>

#include "../../include/flock_gpu.h"
#include "batch.h"
#include "commit.h"
#include "commit_noalloc.h"
#include "ctx.h"
#include "ligerito_configs.h"
#include "prover.h"
#include "r1cs.h"
#include "stmt_digest.h"
#include "host_blake3.h"
#include "verifier.h"

struct flock_ctx {
  flock::Ctx c;
};

using flock::Ctx;

// A context whose wait for the device timed out may still have work running
// on its stream: every later call on it fails.
static int usable(const flock_ctx *ctx) {
  if (!ctx) return FLOCK_E_ARG;
  return ctx->c.timed_out ? FLOCK_E_DEVICE : FLOCK_OK;
}

// A failure that came with (or after) a timed-out wait is a device error,
// whatever the failing step reported (a scratch allocation that had to wait
// first reports out-of-memory).
static int device_rc(const Ctx *c, int rc) { return rc && c->timed_out ? FLOCK_E_DEVICE : rc; }

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
  if (int u = usable(ctx)) return u;
  if (!ctx) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  if (int rc = ctx->c.wait_stream(ctx->c.stream)) return rc;
  FLOCK_HIP_CHECK(&ctx->c, hipFree(d_ptr));
  return FLOCK_OK;
}

extern "C" int flock_memcpy_h2d(flock_ctx *ctx, void *d_dst, const void *src, size_t bytes) {
  if (int u = usable(ctx)) return u;
  if (!ctx || (!d_dst && bytes) || (!src && bytes)) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
  if (int rc = ctx->c.wait_stream(ctx->c.stream)) return rc;
  return FLOCK_OK;
}

extern "C" int flock_memcpy_d2h(flock_ctx *ctx, void *dst, const void *d_src, size_t bytes) {
  if (int u = usable(ctx)) return u;
  if (!ctx || (!d_src && bytes) || (!dst && bytes)) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
  if (int rc = ctx->c.wait_stream(ctx->c.stream)) return rc;
  return FLOCK_OK;
}

extern "C" int flock_memcpy_d2d(flock_ctx *ctx, void *d_dst, const void *d_src, size_t bytes) {
  if (int u = usable(ctx)) return u;
  if (!ctx || (!d_src && bytes) || (!d_dst && bytes)) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  FLOCK_HIP_CHECK(&ctx->c, hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
  if (int rc = ctx->c.wait_stream(ctx->c.stream)) return rc;
  return FLOCK_OK;
}

extern "C" int flock_synchronize(flock_ctx *ctx) {
  if (int u = usable(ctx)) return u;
  if (!ctx) return FLOCK_E_ARG;
  hipSetDevice(ctx->c.device);
  if (int rc = ctx->c.wait_stream(ctx->c.stream)) return rc;
  return FLOCK_OK;
}

extern "C" int flock_ctx_create(int device, flock_ctx **out) {
  if (!out) return FLOCK_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return FLOCK_E_DEVICE;
  if (hipSetDevice(device) != hipSuccess) return FLOCK_E_DEVICE;
  // The prover waits on the stream after every round message (the host holds
  // the Fiat-Shamir transcript); spinning instead of sleeping shortens each
  // wait.  Only possible before the device is first used; otherwise ignored.
  (void)hipSetDeviceFlags(hipDeviceScheduleSpin);
  (void)hipGetLastError();
  flock_ctx *c = new (std::nothrow) flock_ctx();
  if (!c) return FLOCK_E_OOM;
  c->c.device = device;
  if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FLOCK_E_DEVICE;
  }
  // this GPU's copy of the no-allocation commit's fixed array
  if (flock::commit_noalloc_init(&c->c)) {
    delete c;
    return FLOCK_E_DEVICE;
  }
  *out = c;
  return FLOCK_OK;
}

extern "C" void flock_ctx_destroy(flock_ctx *ctx) { delete ctx; }

extern "C" int flock_ctx_set_option(flock_ctx *ctx, const char *name, int64_t value) {
  if (!ctx || !name) return FLOCK_E_ARG;
  if (!strcmp(name, "device_transcript")) {
    ctx->c.device_transcript = value != 0;
    return FLOCK_OK;
  }
  if (!strcmp(name, "zc_graph")) {
    ctx->c.zc_graph = value != 0;
    return FLOCK_OK;
  }
  ctx->c.last_error = std::string("unknown option ") + name;
  return FLOCK_E_ARG;
}

extern "C" int flock_ctx_set_wait_timeout(flock_ctx *ctx, double seconds) {
  if (!ctx || !(seconds == seconds)) return FLOCK_E_ARG;
  ctx->c.wait_timeout_s = seconds;
  return FLOCK_OK;
}

extern "C" const char *flock_last_error(const flock_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

static hipStream_t pick_stream(flock_ctx *ctx, void *stream) {
  return stream ? (hipStream_t)stream : ctx->c.stream;
}

extern "C" int flock_pcs_params_sizes(const flock_pcs_params *params, uint64_t *msg_f128, uint64_t *codeword_f128,
                                      uint64_t *tree_nodes) {
  if (!params) return FLOCK_E_ARG;
  if (flock::commit_params_unsupported(params)) return FLOCK_E_SHAPE;
  const int k = (int)params->m - 7 - (int)params->log_batch_size + (int)params->log_inv_rate;
  if (msg_f128) *msg_f128 = 1ull << (params->m - 7);
  if (codeword_f128) *codeword_f128 = 1ull << (k + (int)params->log_batch_size);
  if (tree_nodes) *tree_nodes = (2ull << k) - 1;
  return FLOCK_OK;
}

extern "C" int flock_commit_device(flock_ctx *ctx, const flock_pcs_params *params, const void *d_z, void *d_codeword,
                                   void *d_tree, uint8_t root[32], void *stream) {
  if (int u = usable(ctx)) return u;
  if (!ctx || !params || !d_z || !d_codeword || !d_tree || !root) return FLOCK_E_ARG;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  return device_rc(&ctx->c, flock::commit_device(&ctx->c, params, d_z, d_codeword, d_tree, root, pick_stream(ctx, stream)));
}

extern "C" int flock_commit_device_noalloc(flock_ctx *ctx, const flock_pcs_params *params, const void *d_z,
                                           void *d_codeword, void *d_tree, uint8_t root[32], void *stream) {
  if (int u = usable(ctx)) return u;
  if (!params || !d_z || !d_codeword || !d_tree || !root) return FLOCK_E_ARG;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  return device_rc(&ctx->c, flock::commit_device_noalloc(&ctx->c, params, d_z, d_codeword, d_tree, root,
                                                         pick_stream(ctx, stream)));
}

// ------------------------------------------------------------------ NTT plans
extern "C" int flock_ntt_autotune(flock_ctx *ctx, const flock_pcs_params *params, int reps, char *json, size_t cap) {
  if (int u = usable(ctx)) return u;
  if (!params || !json || cap == 0) return FLOCK_E_ARG;
  Ctx *c = &ctx->c;
  if (const char *why = flock::commit_params_unsupported(params)) {
    c->last_error = std::string("ntt autotune: ") + why;
    return FLOCK_E_SHAPE;
  }
  if (hipSetDevice(c->device) != hipSuccess) return FLOCK_E_DEVICE;
  const int k = (int)params->m - 7 - (int)params->log_batch_size + (int)params->log_inv_rate;
  std::string out;
  const int rc = flock::ntt_autotune(c, k, (int)params->log_inv_rate, reps, out);
  if (rc) return device_rc(c, rc);
  if (out.size() + 1 > cap) {
    c->last_error = "ntt autotune: json buffer too small (" + std::to_string(out.size() + 1) + " bytes needed)";
    return FLOCK_E_CAP;
  }
  memcpy(json, out.c_str(), out.size() + 1);
  return FLOCK_OK;
}

// The plan table is process-wide: these need no GPU, and ctx (may be NULL)
// only receives the error text.
extern "C" int flock_ntt_plan_set(flock_ctx *ctx, uint32_t k, uint32_t start, const int32_t *passes,
                                  uint32_t n_passes) {
  if ((n_passes && !passes) || n_passes > 64) return FLOCK_E_ARG;
  std::string err;
  std::vector<flock::NttPass> p;
  for (uint32_t i = 0; i < n_passes; ++i) {
    const int32_t kn = passes[3 * i + 2];
    if (kn < 0 || kn >= flock::NTT_KERNELS) {
      if (ctx) ctx->c.last_error = "ntt plan: unknown kernel " + std::to_string(kn);
      return FLOCK_E_ARG;
    }
    p.push_back({passes[3 * i], passes[3 * i + 1], (flock::NttKernel)kn});
  }
  const int rc = flock::ntt_plan_install((int)k, (int)start, p, err);
  if (rc && ctx) ctx->c.last_error = err;
  return rc;
}

extern "C" int flock_ntt_plan_get(flock_ctx *ctx, uint32_t k, uint32_t start, int32_t *passes, uint32_t cap,
                                  uint32_t *n_passes) {
  (void)ctx;
  if (!n_passes || (cap && !passes) || start >= k || k > 40) return FLOCK_E_ARG;
  const std::vector<flock::NttPass> p = flock::ntt_plan((int)k, (int)start);
  *n_passes = (uint32_t)p.size();
  if (p.size() > cap) return FLOCK_E_CAP;
  for (size_t i = 0; i < p.size(); ++i) {
    passes[3 * i] = p[i].l;
    passes[3 * i + 1] = p[i].F;
    passes[3 * i + 2] = (int32_t)p[i].kernel;
  }
  return FLOCK_OK;
}

extern "C" int flock_ntt_plans_load(flock_ctx *ctx, const char *path) {
  std::string err;
  const int rc = flock::ntt_plans_load(path, err);
  if (rc && ctx) ctx->c.last_error = err;
  return rc;
}

extern "C" int flock_commit(flock_ctx *ctx, const flock_pcs_params *params, const void *z_packed, size_t z_bytes,
                            uint8_t root[32]) {
  if (int u = usable(ctx)) return u;
  if (!ctx || !params || !z_packed || !root) return FLOCK_E_ARG;
  Ctx *c = &ctx->c;
  if (hipSetDevice(c->device) != hipSuccess) return FLOCK_E_DEVICE;
  if (const char *why = flock::commit_params_unsupported(params)) {
    c->last_error = std::string("commit: ") + why;
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
  if (!dz || !dcw || !dtree) return device_rc(c, FLOCK_E_OOM);
  FLOCK_HIP_CHECK(c, hipMemcpyAsync(dz, z_packed, msg_bytes, hipMemcpyHostToDevice, c->stream));
  return device_rc(c, flock::commit_device(c, params, dz, dcw, dtree, root, c->stream));
}

// ------------------------------------------------------------------ prover
struct flock_setup {
  flock_ctx *ctx;
  uint64_t n_blocks;
  uint32_t m;
  uint8_t digest[32];
};

extern "C" int flock_setup_create(flock_ctx *ctx, uint64_t n_blocks, uint32_t profile, uint32_t merkle_hash,
                                  flock_setup **out) {
  if (!ctx || !out) return FLOCK_E_ARG;
  *out = nullptr;
  if (profile != 0 || merkle_hash != 1) {
    ctx->c.last_error = "setup: only the Fast profile with the Blake3 Merkle hash is implemented";
    return FLOCK_E_ARG;
  }
  if (n_blocks == 0 || n_blocks > (1ull << 21)) {
    ctx->c.last_error = "setup: n_blocks must be in 1..2^21";
    return FLOCK_E_SHAPE;
  }
  uint32_t n_log = 0;
  while ((1ull << n_log) < (n_blocks < 8 ? 8 : n_blocks)) ++n_log;
  const uint32_t m = flock::r1cs::K_LOG + n_log;
  bool found = false;
  for (int i = 0; i < FLOCK_LIG_N_CONFIGS; ++i)
    if (FLOCK_LIG_CONFIGS[i].m == (int)m && FLOCK_LIG_CONFIGS[i].profile == 0) found = true;
  if (!found) {
    ctx->c.last_error = "setup: no Ligerito configuration for m = " + std::to_string(m) + " (supported: 22..35)";
    return FLOCK_E_SHAPE;
  }
  flock_setup *s = new (std::nothrow) flock_setup();
  if (!s) return FLOCK_E_OOM;
  s->ctx = ctx;
  s->n_blocks = n_blocks;
  s->m = m;
  flock::r1cs::matrices();
  // the digest's m-independent spine: persisted, else hashed on this GPU
  flock::r1cs::spine(&ctx->c);
  flock::r1cs::statement_digest((int)m, s->digest);
  *out = s;
  return FLOCK_OK;
}

extern "C" void flock_setup_destroy(flock_setup *setup) { delete setup; }

extern "C" int flock_statement_digest(uint32_t m, uint8_t digest[32]) {
  if (!digest || m < 14 || m > 40) return FLOCK_E_ARG;
  flock::r1cs::statement_digest((int)m, digest);
  return FLOCK_OK;
}

extern "C" int flock_statement_setup(flock_ctx *ctx, char *from, size_t from_cap) {
  if (int u = usable(ctx)) return u;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  std::string src;
  flock::r1cs::spine(&ctx->c, &src);
  if (from && from_cap) {
    const size_t k = src.size() < from_cap - 1 ? src.size() : from_cap - 1;
    memcpy(from, src.data(), k);
    from[k] = 0;
  }
  return FLOCK_OK;
}

extern "C" int flock_setup_info(const flock_setup *setup, uint32_t *m, uint8_t digest[32]) {
  if (!setup) return FLOCK_E_ARG;
  if (m) *m = setup->m;
  if (digest) memcpy(digest, setup->digest, 32);
  return FLOCK_OK;
}

static int finish_proof(Ctx *c, std::vector<uint8_t> &bytes, uint8_t *out, size_t cap, size_t *out_len) {
  c->last_proof.swap(bytes);
  if (out_len) *out_len = c->last_proof.size();
  if (!out || cap < c->last_proof.size()) {
    c->last_error = "proof needs " + std::to_string(c->last_proof.size()) + " bytes";
    return FLOCK_E_CAP;
  }
  memcpy(out, c->last_proof.data(), c->last_proof.size());
  return FLOCK_OK;
}

static int check_prove_args(flock_setup *setup, uint64_t n_blocks, const uint8_t *domain, size_t domain_len) {
  if (!setup || !setup->ctx || (!domain && domain_len)) return FLOCK_E_ARG;
  if (int u = usable(setup->ctx)) return u;
  if (n_blocks != setup->n_blocks) {
    setup->ctx->c.last_error = "prove: n_blocks " + std::to_string(n_blocks) + " != setup's " +
                               std::to_string(setup->n_blocks);
    return FLOCK_E_SHAPE;
  }
  if (hipSetDevice(setup->ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  return FLOCK_OK;
}

extern "C" int flock_prove_blake3_device(flock_setup *setup, const flock_compression *d_blocks, uint64_t n_blocks,
                                         const uint8_t *domain, size_t domain_len, uint8_t *out, size_t cap,
                                         size_t *out_len, void *stream) {
  int rc = check_prove_args(setup, n_blocks, domain, domain_len);
  if (rc) return rc;
  if (!d_blocks) return FLOCK_E_ARG;
  Ctx *c = &setup->ctx->c;
  std::vector<uint8_t> bytes;
  rc = flock::prove_blake3_device(c, pick_stream(setup->ctx, stream), d_blocks, n_blocks, domain, domain_len, bytes);
  if (rc) return device_rc(c, rc);
  return finish_proof(c, bytes, out, cap, out_len);
}

extern "C" int flock_prove_blake3(flock_setup *setup, const flock_compression *blocks, uint64_t n_blocks,
                                  const uint8_t *domain, size_t domain_len, uint8_t *out, size_t cap,
                                  size_t *out_len) {
  int rc = check_prove_args(setup, n_blocks, domain, domain_len);
  if (rc) return rc;
  if (!blocks) return FLOCK_E_ARG;
  Ctx *c = &setup->ctx->c;
  void *d = c->get_scratch("prove.blocks", n_blocks * sizeof(flock_compression));
  if (!d) return FLOCK_E_OOM;
  FLOCK_HIP_CHECK(c, hipMemcpyAsync(d, blocks, n_blocks * sizeof(flock_compression), hipMemcpyHostToDevice, c->stream));
  std::vector<uint8_t> bytes;
  rc = flock::prove_blake3_device(c, c->stream, (const flock_compression *)d, n_blocks, domain, domain_len, bytes);
  if (rc) return device_rc(c, rc);
  return finish_proof(c, bytes, out, cap, out_len);
}

extern "C" int flock_generate_compressions_device(flock_ctx *ctx, uint32_t log2, uint64_t seed,
                                                  flock_compression *d_out) {
  if (int u = usable(ctx)) return u;
  if (!ctx || !d_out || log2 > 30) return FLOCK_E_ARG;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  return flock::generate_compressions_device(&ctx->c, ctx->c.stream, log2, seed, d_out);
}

extern "C" int flock_prove_blake3_seeded(flock_setup *setup, uint32_t log2, uint64_t seed, const uint8_t *domain,
                                         size_t domain_len, uint8_t *out, size_t cap, size_t *out_len) {
  if (log2 > 30) return FLOCK_E_ARG;
  int rc = check_prove_args(setup, 1ull << log2, domain, domain_len);
  if (rc) return rc;
  Ctx *c = &setup->ctx->c;
  void *d = c->get_scratch("prove.blocks", (1ull << log2) * sizeof(flock_compression));
  if (!d) return FLOCK_E_OOM;
  if ((rc = flock::generate_compressions_device(c, c->stream, log2, seed, (flock_compression *)d))) return rc;
  std::vector<uint8_t> bytes;
  rc = flock::prove_blake3_device(c, c->stream, (const flock_compression *)d, 1ull << log2, domain, domain_len, bytes);
  if (rc) return device_rc(c, rc);
  return finish_proof(c, bytes, out, cap, out_len);
}

extern "C" int flock_verify_blake3(const uint8_t *proof, size_t len, uint64_t n_blocks, const uint8_t *domain,
                                   size_t domain_len, const uint8_t expected_root[32], char *err, size_t err_cap) {
  if ((!proof && len) || (!domain && domain_len) || n_blocks == 0) return FLOCK_E_ARG;
  std::string why;
  int rc = flock::verify_blake3(proof, len, n_blocks, domain, domain_len, expected_root, why);
  if (err && err_cap) {
    const size_t k = why.size() < err_cap - 1 ? why.size() : err_cap - 1;
    memcpy(err, why.data(), k);
    err[k] = 0;
  }
  return rc;
}

extern "C" int flock_commitment_root(flock_setup *setup, const flock_compression *blocks, uint64_t n_blocks,
                                     uint8_t root[32]) {
  int rc = check_prove_args(setup, n_blocks, nullptr, 0);
  if (rc) return rc;
  if (!blocks || !root) return FLOCK_E_ARG;
  Ctx *c = &setup->ctx->c;
  void *d = c->get_scratch("prove.blocks", n_blocks * sizeof(flock_compression));
  if (!d) return FLOCK_E_OOM;
  FLOCK_HIP_CHECK(c, hipMemcpyAsync(d, blocks, n_blocks * sizeof(flock_compression), hipMemcpyHostToDevice, c->stream));
  return device_rc(c, flock::commitment_root_device(c, c->stream, (const flock_compression *)d, n_blocks, root));
}

extern "C" int flock_commitment_root_seeded(flock_setup *setup, uint32_t log2, uint64_t seed, uint8_t root[32]) {
  if (log2 > 30 || !root) return FLOCK_E_ARG;
  int rc = check_prove_args(setup, 1ull << log2, nullptr, 0);
  if (rc) return rc;
  Ctx *c = &setup->ctx->c;
  void *d = c->get_scratch("prove.blocks", (1ull << log2) * sizeof(flock_compression));
  if (!d) return FLOCK_E_OOM;
  if ((rc = flock::generate_compressions_device(c, c->stream, log2, seed, (flock_compression *)d))) return rc;
  return device_rc(c, flock::commitment_root_device(c, c->stream, (const flock_compression *)d, 1ull << log2, root));
}

extern "C" void flock_blake3(const void *data, size_t len, uint8_t out[32]) {
  if (!out) return;
  flock::hb3::hash(data, len, out);
}

extern "C" int flock_last_proof(flock_ctx *ctx, uint8_t *out, size_t cap, size_t *out_len) {
  if (!ctx) return FLOCK_E_ARG;
  if (out_len) *out_len = ctx->c.last_proof.size();
  if (!out || cap < ctx->c.last_proof.size()) return FLOCK_E_CAP;
  memcpy(out, ctx->c.last_proof.data(), ctx->c.last_proof.size());
  return FLOCK_OK;
}

extern "C" int flock_timing_enable(flock_ctx *ctx, int on) {
  if (!ctx) return FLOCK_E_ARG;
  ctx->c.resolve_phases();
  ctx->c.phase_totals.clear();
  ctx->c.timing = on != 0;
  return FLOCK_OK;
}

extern "C" int flock_timing_select(flock_ctx *ctx, const char *names) {
  if (!ctx) return FLOCK_E_ARG;
  ctx->c.timing_only.clear();
  if (names) {
    std::string all(names), cur;
    for (char ch : all + ",") {
      if (ch == ',') {
        if (!cur.empty()) ctx->c.timing_only.insert(cur);
        cur.clear();
      } else {
        cur += ch;
      }
    }
  }
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

extern "C" int flock_prove_batch(flock_ctx *ctx, const flock_batch_item *items, uint32_t n_items,
                                 const uint8_t *domain, size_t domain_len, uint32_t n_workers, int use_graph,
                                 uint8_t *out, size_t cap, uint64_t *offsets, double *done_s, double *front_s,
                                 int *replayed) {
  if (int u = usable(ctx)) return u;
  if ((!items && n_items) || !offsets || (!domain && domain_len)) return FLOCK_E_ARG;
  if (hipSetDevice(ctx->c.device) != hipSuccess) return FLOCK_E_DEVICE;
  std::vector<flock::BatchItem> v(n_items);
  for (uint32_t i = 0; i < n_items; ++i) {
    v[i].blocks = items[i].d_blocks;
    v[i].n_blocks = items[i].n_blocks;
    v[i].log2 = items[i].log2;
    v[i].seed = items[i].seed;
    if (!v[i].blocks && v[i].log2 > 21) return FLOCK_E_ARG;
  }
  std::vector<std::vector<uint8_t>> proofs;
  std::vector<double> done;
  flock::BatchStats stats;
  const int rc = flock::prove_batch(&ctx->c, v, domain, domain_len, (int)n_workers, use_graph != 0, proofs, done,
                                    &stats);
  if (rc) return device_rc(&ctx->c, rc);
  uint64_t at = 0;
  for (uint32_t i = 0; i < n_items; ++i) {
    offsets[i] = at;
    at += proofs[i].size();
  }
  offsets[n_items] = at;
  if (done_s)
    for (uint32_t i = 0; i < n_items; ++i) done_s[i] = done[i];
  if (front_s) *front_s = stats.front_s;
  if (replayed) *replayed = stats.graph_replayed ? 1 : 0;
  if (!out || cap < at) {
    ctx->c.last_error = "batch: the proofs need " + std::to_string(at) + " bytes";
    return FLOCK_E_CAP;
  }
  for (uint32_t i = 0; i < n_items; ++i)
    if (!proofs[i].empty()) memcpy(out + offsets[i], proofs[i].data(), proofs[i].size());
  return FLOCK_OK;
}
