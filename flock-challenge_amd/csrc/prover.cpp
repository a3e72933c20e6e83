// The BLAKE3 proof on one MI355X: Blake3Setup::prove_fast
// (crates/flock-prover/src/r1cs_hashes/blake3.rs:3405) composed as
// prove_ligerito (crates/flock-prover/src/prover.rs:194-295).  All bulk work
// runs in HIP kernels on the context stream; the host keeps the Fiat-Shamir
// transcript and the O(1)-size algebra between rounds.
#include "prover.h"

#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "commit.h"
#include "fs_dev.h"
#include "host_algebra.h"
#include "ligerito_configs.h"
#include "prove_kernels.h"
#include "r1cs.h"
#include "transcript.h"
#include "zc_tables.h"

namespace flock {

static_assert(Transcript::STATE_BYTES == FS_STATE_BYTES, "host and device transcript states differ");

namespace {

const F128 ONE = F128::one();

// S -> Lambda extension matrix: M[j][s] = value at 64 + j of the degree-<64
// polynomial through (s, [s == t]) over GF(2^8) (the inverse-S / forward-
// Lambda LCH NTT pair of zerocheck/univariate_skip.rs:106-120 computes the
// same map).  Filled with the round-1 tables.
uint8_t EXT_M[64][64];

const ZcTables &zc_tables() {
  static ZcTables T;
  static std::once_flag once;
  std::call_once(once, [] {
    memset(&T, 0, sizeof T);
    auto &M = EXT_M;
    for (int j = 0; j < 64; ++j)
      for (int s = 0; s < 64; ++s) {
        uint8_t num = 1, den = 1;
        for (int t = 0; t < 64; ++t) {
          if (t == s) continue;
          num = f8_mul(num, (uint8_t)((64 + j) ^ t));
          den = f8_mul(den, (uint8_t)(s ^ t));
        }
        M[j][s] = f8_mul(num, f8_inv(den));
      }
    for (int v = 0; v < 256; ++v)
      for (int p = 0; p < 8; ++p)
        for (int lam = 0; lam < 64; ++lam) {
          uint8_t acc = 0;
          for (int t = 0; t < 8; ++t)
            if ((v >> t) & 1) acc ^= M[lam][8 * p + t];
          const int h = lam >> 5, bit = lam & 31;
          for (int t = 0; t < 8; ++t)
            if ((acc >> t) & 1) T.extb[((t >> 2) * 4096 + v * 16 + 2 * p + h) * 4 + (t & 3)] |= 1u << bit;
        }
    F128 beta[4];
    for (int i = 0; i < 4; ++i) beta[i] = medium(i);
    for (int xm = 0; xm < 16; ++xm) {
      F128 e4 = ONE;
      for (int i = 0; i < 4; ++i) e4 = e4 * (((xm >> i) & 1) ? beta[i] : ONE + beta[i]);
      for (int t = 0; t < 8; ++t) {
        F128 v = e4 * phi8((uint8_t)(1 << t));
        memcpy(T.pt[xm][t], &v, 16);
      }
    }
  });
  return T;
}

// The transposes of A_0 and B_0 as the lincheck kernel reads them, in one
// upload: per matrix, ptr (K + 1 words), split (K words: the first entry of
// each column whose row is in the upper half, rows >= K / 2) and the rows as
// 16-bit indices, padded by 8 entries for the kernel's 16-byte loads.
struct LcImage {
  std::vector<uint8_t> bytes;
  size_t off[2][3];  // [matrix][ptr, split, rows]
  bool ok = true;
  LcCsc at(const uint8_t *base, int mi) const {
    return {(const uint32_t *)(base + off[mi][0]), (const uint32_t *)(base + off[mi][1]),
            (const uint16_t *)(base + off[mi][2])};
  }
};

const LcImage &lc_image() {
  static LcImage img;
  static std::once_flag once;
  std::call_once(once, [] {
    const r1cs::Matrices &mx = r1cs::matrices();
    static_assert(r1cs::K <= 1 << 16, "16-bit row indices");
    size_t at = 0;
    auto reserve = [&](size_t bytes) {
      const size_t o = at;
      at = (at + bytes + 255) & ~(size_t)255;
      return o;
    };
    const r1cs::Csr *t[2] = {&mx.a_t, &mx.b_t};
    for (int mi = 0; mi < 2; ++mi) {
      img.off[mi][0] = reserve(4ull * (r1cs::K + 1));
      img.off[mi][1] = reserve(4ull * r1cs::K);
      img.off[mi][2] = reserve(2ull * (t[mi]->cols.size() + 8));
    }
    img.bytes.assign(at, 0);
    for (int mi = 0; mi < 2; ++mi) {
      const r1cs::Csr &c = *t[mi];
      uint32_t *ptr = (uint32_t *)(img.bytes.data() + img.off[mi][0]);
      uint32_t *split = (uint32_t *)(img.bytes.data() + img.off[mi][1]);
      uint16_t *rows = (uint16_t *)(img.bytes.data() + img.off[mi][2]);
      for (int col = 0; col <= r1cs::K; ++col) ptr[col] = c.ptr[col];
      for (size_t k = 0; k < c.cols.size(); ++k) rows[k] = (uint16_t)c.cols[k];
      for (int col = 0; col < r1cs::K; ++col) {
        uint32_t k = c.ptr[col];
        while (k < c.ptr[col + 1] && c.cols[k] < (uint32_t)r1cs::K / 2) ++k;
        for (uint32_t q = k; q < c.ptr[col + 1]; ++q)
          if (c.cols[q] < (uint32_t)r1cs::K / 2) img.ok = false;  // rows not ascending: the split is wrong
        split[col] = k;
      }
    }
  });
  return img;
}

// Node indices of merkle_multi_proof (merkle.rs:879-925) in emission order.
std::vector<uint64_t> multiproof_nodes(size_t num_leaves, const std::vector<size_t> &pos) {
  std::vector<uint64_t> nodes;
  if (pos.empty() || num_leaves == 1) return nodes;
  std::vector<size_t> act = pos;  // already sorted + distinct
  size_t ls = 0, ll = num_leaves;
  while (ll > 1) {
    std::vector<size_t> nxt;
    for (size_t i = 0; i < act.size();) {
      size_t p = act[i];
      if (i + 1 < act.size() && act[i + 1] == (p ^ 1)) {
        i += 2;
      } else {
        nodes.push_back(ls + (p ^ 1));
        i += 1;
      }
      nxt.push_back(p >> 1);
    }
    act.swap(nxt);
    ls += ll;
    ll >>= 1;
  }
  return nodes;
}

struct Bytes {
  std::vector<uint8_t> v;
  void put(const void *p, size_t n) { v.insert(v.end(), (const uint8_t *)p, (const uint8_t *)p + n); }
  void u64(uint64_t x) { put(&x, 8); }
  void u32(uint32_t x) { put(&x, 4); }
  void f(F128 x) { put(&x, 16); }
  void fvec(const F128 *p, size_t n) {
    u64(n);
    put(p, 16 * n);
  }
};

struct Opening {
  std::vector<F128> rows;  // n x lanes
  size_t lanes = 0, n = 0;
  std::vector<uint8_t> hashes;  // 32 x k
  // rows then hashes in the context's pinned opening area, copied there
  // without waiting; valid once the stream has drained past those copies
  const uint8_t *pending = nullptr;
};

void finish_opening(Opening &o) {
  if (!o.pending) return;
  memcpy(o.rows.data(), o.pending, o.rows.size() * 16);
  memcpy(o.hashes.data(), o.pending + o.rows.size() * 16, o.hashes.size());
  o.pending = nullptr;
}

void put_opening(Bytes &b, const Opening &o) {
  b.u64(o.n);
  for (size_t i = 0; i < o.n; ++i) b.fvec(&o.rows[i * o.lanes], o.lanes);
  b.u64(o.hashes.size() / 32);
  b.put(o.hashes.data(), o.hashes.size());
}

F128 fv(const uint4 &u) { return F128((uint64_t)u.x | ((uint64_t)u.y << 32), (uint64_t)u.z | ((uint64_t)u.w << 32)); }

// A Ligero commitment on the device (ligero_commit, pcs/ligerito.rs:3340-3560).
struct Witness {
  void *cw = nullptr, *tree = nullptr;
  size_t block_len = 0, lanes = 0;
  int dim = 0, rate = 0;
  uint8_t root[32];
};

}  // namespace

// ------------------------------------------------------------------ the prover
struct Prover {
  Ctx *ctx;
  hipStream_t st;
  int m, L;
  const flock_lig_config *cfg = nullptr;
  std::vector<const char *> scratch_names;

  bool pow_armed = false;  // the device's best nonce holds all-ones
  // openings copied back without a wait since the stream last drained
  int openings_in_flight = 0;

  void *dev(const char *name, size_t bytes) { return ctx->get_scratch(name, bytes); }

  // Wait until everything queued on the stream so far has finished: a
  // signal kernel publishes a new sequence number in the mapped readback
  // area and the host polls it.
  int wait_signal() {
    const uint32_t seq = ctx->next_readback_seq();
    if (int rc = signal_device(ctx, st, ctx->readback_flag_dev(), seq)) return rc;
    return wait_seq(seq);
  }
  // Poll until seq is published (with the stream's error state checked now
  // and then, so a failed kernel ends the wait, and a wall-clock deadline,
  // so a kernel that never finishes does too).
  int wait_seq(uint32_t seq) {
    if (int rc = ctx->wait_flag_seq((const uint32_t *)(ctx->readback + Ctx::READBACK_FLAG), seq, st)) return rc;
    openings_in_flight = 0;  // everything queued before seq has finished
    return FLOCK_OK;
  }
  int sync() {
    if (int rc = ctx->wait_stream(st)) return rc;
    openings_in_flight = 0;
    return FLOCK_OK;
  }

  // Device-to-host read; waits for the stream (the host needs the bytes).
  // Small reads land in pinned memory first.
  int d2h(void *h, const void *d, size_t n) {
    HostPhase hp(ctx, "wait_d2h");
    if (n <= ctx->landing_bytes) {
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(ctx->landing, d, n, hipMemcpyDeviceToHost, st));
      if (int rc = wait_signal()) return rc;
      memcpy(h, ctx->landing, n);
      return FLOCK_OK;
    }
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, st));
    return sync();
  }
  // Two small device ranges (together at most the landing area) with one
  // wait for the stream.
  int d2h2(void *h0, const void *d0, size_t n0, void *h1, const void *d1, size_t n1) {
    if (n0 + n1 > ctx->landing_bytes) {
      int rc = d2h(h0, d0, n0);
      return rc ? rc : d2h(h1, d1, n1);
    }
    HostPhase hp(ctx, "wait_d2h");
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(ctx->landing, d0, n0, hipMemcpyDeviceToHost, st));
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(ctx->landing + n0, d1, n1, hipMemcpyDeviceToHost, st));
    if (int rc = wait_signal()) return rc;
    memcpy(h0, ctx->landing, n0);
    memcpy(h1, ctx->landing + n0, n1);
    return FLOCK_OK;
  }
  // A small result a kernel wrote into the mapped readback area and then
  // published as seq: wait for it, then read it.
  int fetch(void *h, size_t n, uint32_t seq) {
    HostPhase hp(ctx, "wait_d2h");
    if (int rc = wait_seq(seq)) return rc;
    memcpy(h, ctx->readback, n);
    return FLOCK_OK;
  }
  // Host-to-device upload without waiting: the bytes are staged in the
  // pinned ring, which is only recycled after the stream has drained.
  int h2d(void *d, const void *h, size_t n) {
    uint8_t *p = ctx->pinned_alloc(n);
    if (!p) {
      if (int rc = sync()) return rc;
      ctx->pinned_reset();
      p = ctx->pinned_alloc(n);
    }
    if (!p) {  // larger than the ring: plain synchronous copy
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, st));
      return sync();
    }
    memcpy(p, h, n);
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(d, p, n, hipMemcpyHostToDevice, st));
    return FLOCK_OK;
  }

  // eq table over n variables into device buffer `out` (LSB-first):
  // eq(r, x) = lo(x_0..) * mid(..) * hi(..) with host tables of at most 2^8,
  // 2^8 and 2^(n-16) entries, expanded on the device in two steps.
  int eq_device(const F128 *r, int n, void *out) {
    const int a = std::min(n, 8), b = std::min(n - a, 8), c = n - a - b;
    // the factor tables lo | mid | hi go up in one copy (into a buffer of at
    // least 1024 entries, so it is not regrown between calls)
    const size_t na = (size_t)1 << a, nb = (size_t)1 << b, nc = c ? (size_t)1 << c : 0;
    std::vector<F128> tabs(na + nb + nc);
    {
      HostPhase hp(ctx, "eq_tables");
      build_eq(r, a, tabs.data());
      build_eq(r + a, b, tabs.data() + na);
      if (c) build_eq(r + a + b, c, tabs.data() + na + nb);
    }
    uint8_t *d = (uint8_t *)dev("eq.tabs", std::max<size_t>(tabs.size(), 1024) * 16);
    if (!d) return FLOCK_E_OOM;
    if (int rc = h2d(d, tabs.data(), tabs.size() * 16)) return rc;
    const uint8_t *dlo = d, *dmid = d + na * 16, *dhi = dmid + nb * 16;
    if (!c) return eq_expand_device(ctx, st, dlo, a, dmid, b, out);
    void *dt = dev("eq.lomid", 16ull << (a + b));
    if (!dt) return FLOCK_E_OOM;
    if (int rc = eq_expand_device(ctx, st, dlo, a, dmid, b, dt)) return rc;
    return eq_expand_device(ctx, st, dt, a + b, dhi, c, out);
  }

  // Fold grinds of up to this many bits run on the host: the expected 2^bits
  // single-block compressions (a few microseconds) cost less than one device
  // scan and its readback.
  static constexpr uint32_t HOST_GRIND_BITS = 6;

  int grind(Transcript &t, uint32_t bits, uint64_t *nonce_out) {
    uint64_t nonce = 0;
    if (bits && bits <= HOST_GRIND_BITS) {
      // the smallest satisfying nonce, as the device scan finds it
      HostPhase hp(ctx, "pow_host");
      uint32_t m[16] = {0};
      t.state_digest((uint8_t *)m);
      for (;; ++nonce) {
        m[8] = (uint32_t)nonce;
        m[9] = (uint32_t)(nonce >> 32);
        uint32_t o[16];
        hb3::compress(hb3::IV, m, 0, 64, hb3::CHUNK_START | hb3::CHUNK_END | hb3::ROOT, o);
        if ((__builtin_bswap32(o[0]) >> (32 - bits)) == 0) break;
      }
    } else if (bits) {
      ScopedPhase ph(ctx, "pow", st);
      // the best nonce (first 8 bytes; the digest goes as kernel arguments)
      uint8_t *work = (uint8_t *)dev("pow.work", FLOCK_POW_WORK_BYTES);
      if (!work) return FLOCK_E_OOM;
      if (!pow_armed) {  // every search's take re-arms it afterwards
        unsigned long long none = ~0ull;
        if (int rc = h2d(work, &none, 8)) return rc;
        pow_armed = true;
      }
      uint8_t dg[32];
      t.state_digest(dg);
      for (uint64_t start = 0;;) {
        const uint32_t seq = ctx->next_readback_seq();
        uint64_t span = 0;
        int rc = pow_chain_device(ctx, st, work, dg, bits, start, &span);
        if (!rc) rc = pow_take_device(ctx, st, work, ctx->readback_dev, seq);
        unsigned long long got = 0;
        if (!rc) rc = fetch(&got, 8, seq);
        if (rc) return rc;
        if (got != ~0ull) {
          nonce = got;
          break;
        }
        start += span;
      }
    }
    t.observe_bytes(&nonce, 8);
    *nonce_out = nonce;
    return FLOCK_OK;
  }

  int commit_level(const void *f, int log_msg_cols, int k, int rate, const char *tag, Witness &w) {
    ScopedPhase ph(ctx, "ligero_commit", st);
    w.lanes = (size_t)1 << k;
    w.dim = log_msg_cols + rate;
    w.rate = rate;
    w.block_len = (size_t)1 << w.dim;
    std::string cwn = std::string("lig.cw.") + tag, trn = std::string("lig.tree.") + tag;
    w.cw = dev(cwn.c_str(), w.block_len * w.lanes * 16);
    w.tree = dev(trn.c_str(), (2 * w.block_len - 1) * 32);
    if (!w.cw || !w.tree) return FLOCK_E_OOM;
    const uint64_t msg = (uint64_t)1 << (log_msg_cols + k);
    int rc = replicate_device(ctx, st, f, msg, w.cw, w.block_len * w.lanes);
    if (rc) return rc;
    const F128 *tw = ctx->device_twiddles(w.dim);
    if (!tw) return FLOCK_E_DEVICE;
    if ((rc = ntt_narrow_device(ctx, st, w.cw, tw, w.dim, rate, k, 0))) return rc;
    if ((rc = merkle_tree_device(ctx, st, w.cw, (uint32_t)(16 * w.lanes), w.block_len, w.tree))) return rc;
    return d2h(w.root, (const uint8_t *)w.tree + 32 * (2 * w.block_len - 2), 32);
  }

  // With defer, the rows and hashes go to the pinned opening area without a
  // wait (finish_opening reads them after the stream has drained), unless
  // the area is full.
  int open(const Witness &w, const std::vector<size_t> &q, Opening &o, bool defer) {
    ScopedPhase ph(ctx, "open_rows", st);
    o.n = q.size();
    o.lanes = w.lanes;
    o.rows.assign(o.n * o.lanes, F128());
    std::vector<uint64_t> idx(q.begin(), q.end());
    std::vector<uint64_t> nodes = multiproof_nodes(w.block_len, q);
    size_t ni = idx.size() + nodes.size();
    uint64_t *didx = (uint64_t *)dev("open.idx", ni * 8 + 8);
    void *drows = dev("open.rows", o.rows.size() * 16 + 16), *dnodes = dev("open.nodes", nodes.size() * 32 + 32);
    if (!didx || !drows || !dnodes) return FLOCK_E_OOM;
    std::vector<uint64_t> all(idx);
    all.insert(all.end(), nodes.begin(), nodes.end());
    int rc = h2d(didx, all.data(), all.size() * 8);
    if (!rc) rc = gather_device(ctx, st, w.cw, didx, (uint32_t)idx.size(), (uint32_t)(16 * w.lanes), drows);
    if (!rc) rc = gather_device(ctx, st, w.tree, didx + idx.size(), (uint32_t)nodes.size(), 32, dnodes);
    o.hashes.resize(nodes.size() * 32);
    if (rc) return rc;
    const size_t rb = o.rows.size() * 16, hb = o.hashes.size();
    if (defer && ctx->open_area && ctx->open_area_off + rb + hb <= ctx->open_area_bytes) {
      uint8_t *dst = ctx->open_area + ctx->open_area_off;
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(dst, drows, rb, hipMemcpyDeviceToHost, st));
      if (hb) FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(dst + rb, dnodes, hb, hipMemcpyDeviceToHost, st));
      ctx->open_area_off += (rb + hb + 255) & ~(size_t)255;
      o.pending = dst;
      ++openings_in_flight;
      return FLOCK_OK;
    }
    // rows and sibling hashes come back with one wait
    return nodes.empty() ? d2h(o.rows.data(), drows, o.rows.size() * 16)
                         : d2h2(o.rows.data(), drows, o.rows.size() * 16, o.hashes.data(), dnodes, o.hashes.size());
  }

  // induce_sumcheck_poly (pcs/ligerito.rs:1989-2160) as a transposed NTT:
  // basis = R^T F^T s where s[q_i] = eq(alpha, i) and F R is the RS encoding
  // of the Ligero level (replicate, layers rate..dim-1).
  int induce(const Witness &w, const std::vector<size_t> &q, const std::vector<F128> &alpha, int n_cols, void *basis) {
    ScopedPhase ph(ctx, "induce_basis", st);
    std::vector<F128> aw = eq_table(alpha.data(), (int)alpha.size());
    aw.resize(q.size());
    // positions (padded to an even count) then values, uploaded in one copy
    const size_t n_idx = (q.size() + 1) & ~(size_t)1;
    std::vector<uint64_t> up(n_idx + 2 * q.size());
    std::copy(q.begin(), q.end(), up.begin());
    memcpy(up.data() + n_idx, aw.data(), q.size() * 16);
    void *sp = dev("induce.sparse", w.block_len * 16);
    uint64_t *dup = (uint64_t *)dev("induce.up", up.size() * 8 + 16);
    if (!sp || !dup) return FLOCK_E_OOM;
    FLOCK_HIP_CHECK(ctx, hipMemsetAsync(sp, 0, w.block_len * 16, st));
    int rc = h2d(dup, up.data(), up.size() * 8);
    if (!rc) rc = scatter_device(ctx, st, sp, dup, dup + n_idx, (uint32_t)q.size());
    const F128 *tw = ctx->device_twiddles(w.dim);
    if (!tw) return FLOCK_E_DEVICE;
    if (!rc) rc = ntt_narrow_device(ctx, st, sp, tw, w.dim, w.rate, 0, 1);
    if (!rc) rc = fold_replicas_device(ctx, st, sp, (uint64_t)1 << n_cols, 1u << w.rate, basis);
    return rc;
  }

  int msg(const void *f, const void *b, uint64_t len, int with_eval, F128 out[3]) {
    const uint32_t seq = ctx->next_readback_seq();
    int rc = lig_msg_device(ctx, st, f, b, len / 2, with_eval, ctx->readback_dev, seq);
    uint4 h[3];
    if (!rc) rc = fetch(h, 48, seq);
    for (int i = 0; i < 3; ++i) out[i] = fv(h[i]);
    return rc;
  }
};

int commitment_root_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_blocks,
                           uint8_t root[32]) {
  int n_log = 0;
  while ((1ull << n_log) < (n_blocks < 8 ? 8 : n_blocks)) ++n_log;
  const int m = r1cs::K_LOG + n_log, L = m - 7;
  const size_t words = (size_t)1 << (m - 6);
  // the prover's buffers: a context that proves and verifies reuses them
  void *z = ctx->get_scratch("w.z", words * 8), *a = ctx->get_scratch("w.a", words * 8),
       *b = ctx->get_scratch("w.b", words * 8);
  if (!z || !a || !b) return FLOCK_E_OOM;
  int rc = witness_device(ctx, st, d_blocks, n_blocks, n_log, z, a, b);
  if (rc) return rc;
  flock_pcs_params pp{(uint32_t)m, 1, 6, 0, 1};
  const size_t npos0 = (size_t)1 << (L - 6 + 1);
  void *cw0 = ctx->get_scratch("c.codeword", npos0 * 64 * 16), *tree0 = ctx->get_scratch("c.tree", (2 * npos0 - 1) * 32);
  if (!cw0 || !tree0) return FLOCK_E_OOM;
  if ((rc = commit_device(ctx, &pp, z, cw0, tree0, root, st))) return rc;
  return ctx->wait_stream(st);
}

namespace {

const flock_lig_config *lig_config(Ctx *ctx, int m) {
  for (int i = 0; i < FLOCK_LIG_N_CONFIGS; ++i)
    if (FLOCK_LIG_CONFIGS[i].m == m && FLOCK_LIG_CONFIGS[i].profile == 0) return &FLOCK_LIG_CONFIGS[i];
  ctx->last_error = "prove: no Ligerito configuration for m = " + std::to_string(m) + " (supported: 22..35)";
  return nullptr;
}

// Host staging areas of the context (allocated once) and the per-proof
// resets; the stream is idle between proves, so the upload ring can be
// recycled.
int prepare_ctx(Ctx *ctx, hipStream_t st) {
  if (int e = ctx->wait_stream(st)) return e;
  ctx->pinned_reset();
  if (!ctx->landing) {
    void *p = nullptr;
    if (hipHostMalloc(&p, 1 << 20, hipHostMallocDefault) == hipSuccess) {
      ctx->landing = (uint8_t *)p;
      ctx->landing_bytes = 1 << 20;
    }
  }
  if (!ctx->open_area) {
    void *p = nullptr;
    if (hipHostMalloc(&p, 4 << 20, hipHostMallocDefault) == hipSuccess) {
      ctx->open_area = (uint8_t *)p;
      ctx->open_area_bytes = 4 << 20;
    }
  }
  ctx->open_area_off = 0;
  if (!ctx->readback) {
    void *p = nullptr, *dp = nullptr;
    FLOCK_HIP_CHECK(ctx, hipHostMalloc(&p, Ctx::READBACK_BYTES, hipHostMallocMapped | hipHostMallocCoherent));
    if (hipHostGetDevicePointer(&dp, p, 0) != hipSuccess) {
      hipHostFree(p);
      ctx->last_error = "prove: cannot map the readback area into the device";
      return FLOCK_E_DEVICE;
    }
    ctx->readback = (uint8_t *)p;
    ctx->readback_dev = dp;
  }
  if (!ctx->landing) {
    ctx->last_error = "prove: landing area unavailable";
    return FLOCK_E_DEVICE;
  }
  return FLOCK_OK;
}

// Everything after the commitment, over its buffers: zerocheck round 1
// onwards.  The root is read from root_src once root_ready (may be null:
// already there) has completed, so round 1's row products, queued first,
// overlap the wait.  ph_all (may be null) is ended before the final wait.
int prove_back(Ctx *ctx, hipStream_t st, int m, const flock_lig_config *cfg, const ProofFront &fb,
               hipEvent_t root_ready, const uint8_t *root_src, const uint8_t *domain, size_t domain_len,
               std::vector<uint8_t> &out, ScopedPhase *ph_all);

}  // namespace

int prove_blake3_device(Ctx *ctx, hipStream_t st, const flock_compression *d_blocks, uint64_t n_blocks,
                        const uint8_t *domain, size_t domain_len, std::vector<uint8_t> &out) {
  if (n_blocks == 0) {
    ctx->last_error = "prove: need at least one compression";
    return FLOCK_E_SHAPE;
  }
  int n_log = 0;
  while ((1ull << n_log) < std::max<uint64_t>(n_blocks, 8)) ++n_log;
  const int m = r1cs::K_LOG + n_log;
  const flock_lig_config *cfg = lig_config(ctx, m);
  if (!cfg) return FLOCK_E_SHAPE;
  if (int e = prepare_ctx(ctx, st)) return e;
  ScopedPhase ph_all(ctx, "prove", st);
  HostPhase hp_all(ctx, "prove_wall");
  int rc = 0;
  const int L = m - 7;
  const size_t words = (size_t)1 << (m - 6);  // uint64 words per bit-vector

  // ---- witness + commit
  ProofFront fb;
  fb.z = ctx->get_scratch("w.z", words * 8);
  fb.a = ctx->get_scratch("w.a", words * 8);
  fb.b = ctx->get_scratch("w.b", words * 8);
  if (!fb.z || !fb.a || !fb.b) return FLOCK_E_OOM;
  if ((rc = witness_device(ctx, st, d_blocks, n_blocks, n_log, fb.z, fb.a, fb.b))) return rc;
  flock_pcs_params pp{(uint32_t)m, 1, 6, 0, 1};
  const size_t npos0 = (size_t)1 << (L - 6 + 1);
  fb.codeword = ctx->get_scratch("c.codeword", npos0 * 64 * 16);
  fb.tree = ctx->get_scratch("c.tree", (2 * npos0 - 1) * 32);
  if (!fb.codeword || !fb.tree) return FLOCK_E_OOM;
  // The root comes back through the landing area under an event.
  hipEvent_t root_ready = ctx->take_event();
  if (!root_ready) {
    ctx->last_error = "prove: no event for the root";
    return FLOCK_E_DEVICE;
  }
  if ((rc = commit_device(ctx, &pp, fb.z, fb.codeword, fb.tree, ctx->landing, st, root_ready))) return rc;
  return prove_back(ctx, st, m, cfg, fb, root_ready, ctx->landing, domain, domain_len, out, &ph_all);
}

int prove_after_commit_device(Ctx *ctx, hipStream_t st, int m, const ProofFront &fb, const uint8_t root[32],
                              const uint8_t *domain, size_t domain_len, std::vector<uint8_t> &out) {
  const flock_lig_config *cfg = lig_config(ctx, m);
  if (!cfg) return FLOCK_E_SHAPE;
  if (int e = prepare_ctx(ctx, st)) return e;
  ScopedPhase ph_all(ctx, "prove", st);
  return prove_back(ctx, st, m, cfg, fb, nullptr, root, domain, domain_len, out, &ph_all);
}

namespace {

int prove_back(Ctx *ctx, hipStream_t st, int m, const flock_lig_config *cfg, const ProofFront &fb,
               hipEvent_t root_ready, const uint8_t *root_src, const uint8_t *domain, size_t domain_len,
               std::vector<uint8_t> &out, ScopedPhase *ph_all) {
  Prover P{ctx, st, m, m - 7, cfg, {}};
  int rc = 0;
  const int L = m - 7, n = m - 6;
  const size_t words = (size_t)1 << (m - 6);  // uint64 words per bit-vector
  void *z = fb.z, *a = fb.a, *b = fb.b, *cw0 = fb.codeword, *tree0 = fb.tree;
  flock_pcs_params pp{(uint32_t)m, 1, 6, 0, 1};
  const int k0 = L - 6 + 1;
  const size_t npos0 = (size_t)1 << k0;
  // zerocheck round 1's products depend on the rows only: queued before the
  // wait for the root, so the device works while the host builds the
  // transcript and derives the challenges
  void *dzt = P.dev("zc.tables", sizeof(ZcTables));
  if (!dzt) return FLOCK_E_OOM;
  if (!ctx->zc_tables_uploaded) {
    if ((rc = P.h2d(dzt, &zc_tables(), sizeof(ZcTables)))) return rc;
    ctx->zc_tables_uploaded = true;
  }
  if ((rc = zc_y_device(ctx, st, a, b, m, dzt))) return rc;
  uint8_t root[32];
  if (root_ready) {
    HostPhase hp(ctx, "wait_d2h");
    if (int e = ctx->wait_event(root_ready)) return e;
    ctx->event_pool.push_back(root_ready);
  }
  memcpy(root, root_src, 32);

  Transcript t(domain, domain_len);
  {
    uint8_t dig[32];
    r1cs::statement_digest(m, dig);
    t.label("flock-r1cs-v0");
    t.observe_bytes(dig, 32);
    t.observe_bytes(root, 32);
  }

  Bytes B;
  B.put("FLOCK", 5);
  uint8_t hv[2] = {4, 2};
  B.put(hv, 2);
  B.put(root, 32);
  B.u64(pp.m);
  B.u64(pp.log_inv_rate);
  B.u64(pp.log_batch_size);
  B.u32(pp.profile);
  B.u32(pp.merkle_hash);

  // ---- zerocheck (zerocheck.rs:482-1364)
  t.label("flock-zerocheck-v0");
  std::vector<F128> r(m);
  {
    std::vector<F128> rs = t.sample_vec(6), ro = t.sample_vec(m - 13);
    for (int i = 0; i < 6; ++i) r[i] = rs[i];
    for (int i = 0; i < 3; ++i) r[6 + i] = phi8(SMALL_CHAL[i]);
    for (int i = 0; i < 4; ++i) r[9 + i] = medium(i);
    for (int i = 0; i < m - 13; ++i) r[13 + i] = ro[i];
  }
  // The multilinear rounds' eq tables in split form, E_i = Elo_i (x) Ehi_i
  // (round i's table is E_i = eq(r[7 + i ..]) over n - 1 - i variables):
  // round 0 splits at ZS_LOG + 1 and rounds 1 .. n_split at ZS_LOG, so both
  // start from Ehi = eq(r[7 + ZS_LOG + 1 ..]), which each later split round
  // folds for the next.  They depend on r only and are prepared while the
  // device works on round 1.  Rounds with half <= 2^ZC_WHOLE_LOG use the
  // whole table instead: a split round's workgroup walks 2^ZS_LOG pairs per
  // x_hi in a serial chain, so with few x_hi blocks it is latency-bound
  // (about 25 us however small), while the whole-table kernel spreads the
  // pairs over all CUs and materializing E costs at most 4 MiB.
  constexpr int ZC_WHOLE_LOG = 18;
  const int n_split = std::max(1, std::min(n - 2 - ZS_LOG, n - 2 - ZC_WHOLE_LOG));
  if (n - 2 - ZS_LOG < 1) {  // m >= 22 for every Ligerito configuration
    ctx->last_error = "zerocheck: m too small for the split eq tables";
    return FLOCK_E_SHAPE;
  }
  void *dElo = nullptr, *Ehi = nullptr, *Ehi2 = nullptr;
  const size_t elo0_len = (size_t)2 << ZS_LOG;
  F128 r1ab[64], r1c[64];
  {
    ScopedPhase ph(ctx, "zc_round1_all", st);
    void *dehi = P.dev("zc.ehi", ((size_t)1 << (m - 13)) * 16);
    void *dout = P.dev("zc.r1q", 64 * 128 * 16);
    void *dhic = P.dev("zc.hic", ((size_t)1 << (m - 11)) * 16), *dq = P.dev("zc.cq", 2048 * 16);
    if (!dehi || !dout || !dhic || !dq) return FLOCK_E_OOM;
    if ((rc = P.eq_device(&r[13], m - 13, dehi))) return rc;
    if ((rc = zc_round1_device(ctx, st, m, dehi, dout))) return rc;
    // P^C(lambda_j) = sum_s phi8(M[j][s]) C_s with C_s = sum_x eq(x) bit_s(c[x]);
    // elements are row pairs (x = 2e + h), e = il + 16 ih.
    if ((rc = P.eq_device(&r[11], m - 11, dhic))) return rc;
    if ((rc = bitfold_device(ctx, st, z, 4, 1ull << (m - 11), dhic, nullptr, dq))) return rc;
    {
      dElo = P.dev("zc.elo", (elo0_len + ((size_t)n_split << ZS_LOG)) * 16);
      Ehi = P.dev("zc.ehi.a", 16ull << (n - 2 - ZS_LOG));
      Ehi2 = P.dev("zc.ehi.b", 16ull << (n - 2 - ZS_LOG));
      if (!dElo || !Ehi || !Ehi2) return FLOCK_E_OOM;
      std::vector<F128> elo(elo0_len + ((size_t)n_split << ZS_LOG));
      {
        HostPhase hp(ctx, "eq_tables");
        const std::vector<F128> e0 = eq_table(&r[7], ZS_LOG + 1);
        std::copy(e0.begin(), e0.end(), elo.begin());
        for (int j = 1; j <= n_split; ++j) {
          const std::vector<F128> e = eq_table(&r[7 + j], ZS_LOG);
          std::copy(e.begin(), e.end(), elo.begin() + elo0_len + ((size_t)(j - 1) << ZS_LOG));
        }
      }
      if ((rc = P.h2d(dElo, elo.data(), elo.size() * 16))) return rc;
      if ((rc = P.eq_device(&r[7 + ZS_LOG + 1], n - 2 - ZS_LOG, Ehi))) return rc;
    }
    // both round-1 results come back with one wait
    std::vector<F128> qab(64 * 128), q(2048);
    if ((rc = P.d2h2(qab.data(), dout, qab.size() * 16, q.data(), dq, q.size() * 16))) return rc;
    {
      // P^AB(lambda_j) = sum_{x_mid, t} E4(x_mid) phi8(2^t) bitfold of bit t of
      // Y_j: element 4 x_mid + 2 h + t / 4, bit 32 (t % 4) + b for j = 32 h + b
      const ZcTables &T = zc_tables();
      for (int j = 0; j < 64; ++j) {
        const int h = j >> 5, bit = j & 31;
        F256 acc;  // unreduced (reduction is linear)
        for (int xm = 0; xm < 16; ++xm)
          for (int t = 0; t < 8; ++t) {
            F128 w;
            memcpy(&w, T.pt[xm][t], 16);
            acc ^= mul_unreduced(w, qab[(4 * xm + 2 * h + (t >> 2)) * 128 + 32 * (t & 3) + bit]);
          }
        r1ab[j] = reduce(acc);
      }
    }
    HostPhase hp(ctx, "round1_combine");
    std::vector<F128> lo = eq_table(&r[7], 4);
    F128 cs[64];
    const F128 w0 = ONE + r[6], w1 = r[6];
    for (int sb = 0; sb < 64; ++sb) {
      F256 s0, s1;
      for (int il = 0; il < 16; ++il) {
        s0 ^= mul_unreduced(lo[il], q[il * 128 + sb]);
        s1 ^= mul_unreduced(lo[il], q[il * 128 + 64 + sb]);
      }
      cs[sb] = w0 * reduce(s0) + w1 * reduce(s1);
    }
    for (int j = 0; j < 64; ++j) {
      F256 acc;
      for (int sb = 0; sb < 64; ++sb)
        if (EXT_M[j][sb]) acc ^= mul_unreduced(phi8(EXT_M[j][sb]), cs[sb]);
      r1c[j] = reduce(acc);
    }
  }
  t.observe_slice(r1ab, 64);
  t.observe_slice(r1c, 64);
  const F128 zz = t.sample();
  F128 wl[64], ws[64];
  {
    HostPhase hp(ctx, "lagrange_z");
    lagrange(zz, 64, 64, 64, 64, wl);
    lagrange(zz, 0, 64, 0, 64, ws);
  }
  const F128 final_c = dot(wl, r1c, 64);
  std::vector<F128> rhos;
  std::vector<std::pair<F128, F128>> zc_rounds;
  F128 final_a, final_b;
  {
    const size_t rows = words;
    // After round 1, A[x] = L(a[x]) with L(row) = sum_s bit_s(row) L_s(z); round
    // 0 and its fold evaluate L on the rows, so only the half-size folded
    // tables are materialized.
    void *A = P.dev("zc.A", rows * 8), *Bv = P.dev("zc.B", rows * 8);
    void *A2 = P.dev("zc.A2", rows * 4), *B2 = P.dev("zc.B2", rows * 4);
    // whole eq tables for the rounds after n_split (from 2^(n - 2 - n_split)
    // entries down)
    const size_t e_len = (size_t)1 << (n - 2 - n_split);
    void *E = P.dev("zc.E", e_len * 16), *E2 = P.dev("zc.E2", e_len * 8);
    void *dlw8 = P.dev("zc.lw8", 2048 * 16), *dlwr8 = P.dev("zc.lwr8", 2048 * 16), *dlw = P.dev("zc.lw", 256 * 16);
    if (!A || !Bv || !A2 || !B2 || !E || !E2 || !dlw8 || !dlwr8 || !dlw) return FLOCK_E_OOM;
    // byte tables: out[v * 8 + p] = sum_{t < 8, bit t of v} w[8 p + t], each
    // entry one XOR from the entry without v's lowest set bit
    auto byte_tables = [](const F128 *w, F128 *out) {
      for (int p = 0; p < 8; ++p) {
        out[p] = F128();
        for (int v = 1; v < 256; ++v) out[v * 8 + p] = out[(v & (v - 1)) * 8 + p] + w[8 * p + __builtin_ctz(v)];
      }
    };
    {
      std::vector<F128> lw8(2048);  // of L_s(z)
      byte_tables(ws, lw8.data());
      if ((rc = P.h2d(dlw8, lw8.data(), lw8.size() * 16))) return rc;
    }
    {
      // nibble tables: lw[v * 16 + p] = sum_{t < 4, bit t of v} L_{4p + t}(z)
      F128 lw[256];
      for (int p = 0; p < 16; ++p) {
        lw[p] = F128();
        for (int v = 1; v < 16; ++v) lw[v * 16 + p] = lw[(v & (v - 1)) * 16 + p] + ws[4 * p + __builtin_ctz(v)];
      }
      if ((rc = P.h2d(dlw, lw, sizeof lw))) return rc;
    }
    // the first whole-table round (n_split + 1) starts from eq itself; it
    // depends on the point only, so it is queued now
    if (n_split + 1 < n && (rc = P.eq_device(&r[8 + n_split], n - 2 - n_split, E))) return rc;
    // With the device transcript, the rounds' messages go to device memory,
    // the kernel that reduces each absorbs it and squeezes the round's
    // challenge into device memory, and the folds read it from there: the
    // rounds run with no host round trip, and the host reads everything back
    // once.
    const bool dfs = ctx->device_transcript;
    uint8_t *d_fs = nullptr, *d_msgs = nullptr, *d_rhos = nullptr, *d_ws = nullptr;
    if (dfs) {
      d_fs = (uint8_t *)P.dev("fs.state", FS_STATE_BYTES);
      d_msgs = (uint8_t *)P.dev("zc.msgs", 32 * (size_t)(n + 1));
      d_rhos = (uint8_t *)P.dev("zc.rhos", 16 * (size_t)n);
      d_ws = (uint8_t *)P.dev("zc.ws", 64 * 16);
      if (!d_fs || !d_msgs || !d_rhos || !d_ws) return FLOCK_E_OOM;
      uint8_t st_bytes[FS_STATE_BYTES];
      t.export_state(st_bytes);
      if ((rc = P.h2d(d_fs, st_bytes, FS_STATE_BYTES)) || (rc = P.h2d(d_ws, ws, sizeof ws))) return rc;
    }
    // The reductions' partials: sized once for the largest grid (4096
    // workgroups x 2 values), so no launcher regrows it mid-sequence and a
    // captured graph keeps pointing at it.
    void *d_partial = P.dev("zc.round.partial", (size_t)4096 * 2 * 16);
    if (!d_partial) return FLOCK_E_OOM;
    // With the device transcript the rounds are a fixed sequence of launches
    // for this m and these buffers: captured once in a HIP graph and
    // replayed by later proofs.
    std::vector<const void *> graph_key;
    if (dfs && ctx->zc_graph)
      graph_key = {(const void *)(uintptr_t)m, A, Bv, A2, B2, E, E2, dElo, Ehi, Ehi2, d_fs, d_msgs, d_rhos,
                   d_ws, dlw8, dlwr8, dlw, a, b, d_partial};
    hipGraphExec_t *graph = nullptr;
    bool capturing = false;
    if (!graph_key.empty()) {
      auto it = ctx->zc_graphs.find(graph_key);
      if (it != ctx->zc_graphs.end()) {
        graph = &it->second;
      } else {
        FLOCK_HIP_CHECK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));  // other threads keep using their streams
        capturing = true;
      }
    }
    const bool replay = graph != nullptr;  // the loop below does not run
    const bool timing_was = ctx->timing;
    if (capturing) ctx->timing = false;  // phase events are not part of the graph
    size_t len = rows;
    // A whole-table round folds the previous round's tables itself (one
    // launch fewer per round): then the fold by prev_rho is pending when the
    // round starts.
    bool fold_pending = false;
    F128 prev_rho;
    const uint8_t *prev_rho_dev = nullptr;
    for (int i = 0; i < n && !graph; ++i) {
      const size_t half = len / 2;
      const bool split = i <= n_split;
      const uint32_t seq = dfs ? 0 : ctx->next_readback_seq();
      void *out2 = dfs ? (void *)(d_msgs + 32 * (size_t)i) : ctx->readback_dev;
      // device transcript: the round's reduction kernel absorbs the message
      // and samples the challenge into rho_dev
      const uint8_t *rho_dev = dfs ? d_rhos + 16 * (size_t)i : nullptr;
      FsStep fs;
      if (dfs) fs = FsStep{d_fs, (void *)rho_dev};
      if (fold_pending) {
        rc = zc_fold_round_device(ctx, st, A, Bv, prev_rho, E, half, A2, B2, i + 1 < n ? E2 : nullptr, out2, seq,
                                  prev_rho_dev, fs);
        std::swap(A, A2);
        std::swap(Bv, B2);
        fold_pending = false;
      } else if (i == 0) {
        rc = zc_round0_split_device(ctx, st, a, b, dlw, dElo, Ehi, half, out2, seq, fs);
      } else {  // 1 <= i <= n_split: every later round is fused with its fold
        rc = zc_round_split_device(ctx, st, A, Bv,
                                   (const uint8_t *)dElo + (elo0_len + ((size_t)(i - 1) << ZS_LOG)) * 16, Ehi, half,
                                   out2, i < n_split ? Ehi2 : nullptr, seq, fs);
        std::swap(Ehi, Ehi2);
      }
      if (rc) break;
      if (!split) std::swap(E, E2);
      F128 rho;
      if (!dfs) {
        uint4 h[2];
        if ((rc = P.fetch(h, 32, seq))) return rc;
        F128 g1 = fv(h[0]), gi = fv(h[1]);
        zc_rounds.push_back({g1, gi});
        t.observe(g1);
        t.observe(gi);
        rho = t.sample();
        rhos.push_back(rho);
      }
      if (i == 0) {
        if (dfs) {
          if ((rc = rho_bytetab_device(ctx, st, d_ws, rho_dev, dlwr8))) break;
        } else {
          F128 wr[64];  // rho L_s(z)
          for (int q = 0; q < 64; ++q) wr[q] = rho * ws[q];
          std::vector<F128> lwr8(2048);
          byte_tables(wr, lwr8.data());
          if ((rc = P.h2d(dlwr8, lwr8.data(), lwr8.size() * 16))) return rc;
        }
        if ((rc = zc_fold0_device(ctx, st, a, b, dlw8, dlwr8, half, A, Bv))) break;
      } else if (i + 1 < n && i + 1 > n_split) {
        prev_rho = rho;  // folded by round i + 1's launch
        prev_rho_dev = rho_dev;
        fold_pending = true;
      } else {
        if ((rc = fold_pair_device(ctx, st, A, Bv, half, rho, A2, B2, rho_dev))) break;
        std::swap(A, A2);
        std::swap(Bv, B2);
      }
      len = half;
    }
    if (capturing) {
      ctx->timing = timing_was;
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      FLOCK_HIP_CHECK(ctx, e);
      hipGraphExec_t ex = nullptr;
      const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      FLOCK_HIP_CHECK(ctx, ie);
      graph = &ctx->zc_graphs.emplace(graph_key, ex).first->second;
    }
    if (rc) return rc;
    if (graph) {
      FLOCK_HIP_CHECK(ctx, hipGraphLaunch(*graph, st));
      // where the final tables are when the loop did not run: every round
      // after round 0 folds once into the other buffer pair (n - 1 swaps)
      if (replay && ((n - 1) & 1)) {
        std::swap(A, A2);
        std::swap(Bv, B2);
      }
    }
    if (dfs) {
      // one readback: messages, challenges, the final evaluations and the
      // transcript state, which the host takes over
      const size_t nm = 32 * (size_t)n, nr = 16 * (size_t)n;
      uint8_t *land = ctx->landing;
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land, d_msgs, nm, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm, d_rhos, nr, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm + nr, A, 16, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm + nr + 16, Bv, 16, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm + nr + 32, d_fs, FS_STATE_BYTES, hipMemcpyDeviceToHost, st));
      {
        HostPhase hp(ctx, "wait_d2h");
        if ((rc = P.wait_signal())) return rc;
      }
      for (int i = 0; i < n; ++i) {
        F128 g1, gi, rh;
        memcpy(&g1, land + 32 * i, 16);
        memcpy(&gi, land + 32 * i + 16, 16);
        memcpy(&rh, land + nm + 16 * i, 16);
        zc_rounds.push_back({g1, gi});
        rhos.push_back(rh);
      }
      memcpy(&final_a, land + nm + nr, 16);
      memcpy(&final_b, land + nm + nr + 16, 16);
      t.import_state(land + nm + nr + 32);
    } else {
      uint4 h[2];
      if ((rc = P.d2h2(&h[0], A, 16, &h[1], Bv, 16))) return rc;
      final_a = fv(h[0]);
      final_b = fv(h[1]);
    }
    t.observe(final_a);
    t.observe(final_b);
  }
  B.fvec(r1ab, 64);
  B.fvec(r1c, 64);
  B.u64(zc_rounds.size());
  for (auto &pr : zc_rounds) {
    B.f(pr.first);
    B.f(pr.second);
  }
  B.f(final_a);
  B.f(final_b);
  B.f(final_c);

  // The zerocheck point's part of the ring-switch claims (x_full of claim 1)
  // and its eq tables' buffers: known now, expanded during the lincheck.
  // eq(x_full[1..]) over L variables is kept as two factors, over the low
  // rs_k variables and the rest (rs_basis_device).
  const std::vector<F128> xf1(r.begin() + 6, r.end());
  const size_t N = (size_t)1 << L;
  const int rs_k = std::max(std::min(13, L - 8), std::min(L - 1, 10));
  void *rs_h1 = P.dev("rs.hi1", (N / 16) * 16), *rs_lo1 = P.dev("rs.lo1", 16ull << rs_k);
  void *rs_hk1 = P.dev("rs.hk1", 16ull << (L - rs_k));
  if (!rs_h1 || !rs_lo1 || !rs_hk1) return FLOCK_E_OOM;

  // ---- lincheck (lincheck.rs:1480-1664)
  F128 lc_rounds[8][2], zpart[64], r_inner_skip, w_claim, r_inner_rest[8];
  {
    ScopedPhase ph(ctx, "lincheck", st);
    t.label("flock-lincheck-v0");
    const F128 alpha = t.sample(), beta = t.sample();
    const LcImage &img = lc_image();
    if (!img.ok) {
      ctx->last_error = "lincheck: transposed matrix rows are not ascending";
      return FLOCK_E_DEVICE;
    }
    uint8_t *dimg = (uint8_t *)P.dev("lc.csc", img.bytes.size());
    void *deq = P.dev("lc.eq", 16 << 14), *dcomb = P.dev("lc.comb", 16 << 14), *dzv = P.dev("lc.zv", 16 << 14);
    void *deqo = P.dev("lc.eqo", 16ull << (m - 14));
    if (!dimg || !deq || !dcomb || !dzv || !deqo) return FLOCK_E_OOM;
    if (!ctx->csc_uploaded) {
      if ((rc = P.h2d(dimg, img.bytes.data(), img.bytes.size()))) return rc;
      ctx->csc_uploaded = true;
    }
    const LcCsc lca = img.at(dimg, 0), lcb = img.at(dimg, 1);
    // quirky eq table (lincheck.rs:1127-1142): index i_skip + 64 i_rest,
    // the outer product lam(i_skip) eq(rhos[0..8], i_rest) expanded on the device
    {
      F128 lr[64 + 256];  // lam | eq(rhos[0..8]), uploaded in one copy
      {
        HostPhase hp(ctx, "lincheck_eq");
        lagrange(zz, 0, 64, 0, 64, lr);
        build_eq(rhos.data(), 8, lr + 64);
      }
      uint8_t *dlr = (uint8_t *)P.dev("lc.lamer", sizeof lr);
      if (!dlr) return FLOCK_E_OOM;
      if ((rc = P.h2d(dlr, lr, sizeof lr)) || (rc = eq_expand_device(ctx, st, dlr, 6, dlr + 64 * 16, 8, deq)))
        return rc;
    }
    if ((rc = lc_comb_device(ctx, st, lca, lcb, deq, alpha, beta, r1cs::Z_CONST, dcomb))) return rc;
    if ((rc = P.eq_device(&rhos[8], m - 14, deqo))) return rc;
    // z_vec[128 c + r] = sum_blk eq_outer[blk] bit_r(16-byte chunk c of block blk)
    if ((rc = bitfold_device(ctx, st, z, 7, 1ull << (m - 14), deqo, nullptr, dzv))) return rc;
    // comb and z_vec come back through the landing area; the wait covers only
    // their copies, so the ring-switch eq tables of the zerocheck point
    // (which do not depend on the lincheck rounds) run on the device while
    // the host does the rounds.
    std::vector<F128> comb(1 << 14), zv(1 << 14);
    const size_t lc_bytes = comb.size() * 16;
    hipEvent_t lc_ready = ctx->take_event();
    if (2 * lc_bytes > ctx->landing_bytes || !lc_ready) {
      ctx->last_error = "lincheck: landing area or event unavailable";
      return FLOCK_E_DEVICE;
    }
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(ctx->landing, dcomb, lc_bytes, hipMemcpyDeviceToHost, st));
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(ctx->landing + lc_bytes, dzv, lc_bytes, hipMemcpyDeviceToHost, st));
    FLOCK_HIP_CHECK(ctx, hipEventRecord(lc_ready, st));
    if ((rc = P.eq_device(&xf1[5], L - 4, rs_h1)) || (rc = P.eq_device(&xf1[1], rs_k, rs_lo1)) ||
        (rc = P.eq_device(&xf1[1 + rs_k], L - rs_k, rs_hk1)))
      return rc;
    {
      HostPhase hp(ctx, "wait_d2h");
      if (int e = ctx->wait_event(lc_ready)) return e;
      ctx->event_pool.push_back(lc_ready);
    }
    memcpy(comb.data(), ctx->landing, lc_bytes);
    memcpy(zv.data(), ctx->landing + lc_bytes, lc_bytes);
    HostPhase hp(ctx, "lincheck_rounds");
    // Round k's message (binding the top bit) over the arrays as folded by
    // rounds 0..k-1.  The fold of round k - 1 is fused into round k's pass:
    // position i of the pass computes the folded values at i and i + half
    // from the unfolded ones at i, i + half, i + len, i + half + len and
    // writes them in place (no position is read by another chunk).  Message
    // products accumulate unreduced (reduction is linear) per chunk.
    HostPool &pool = ctx->host_pool();
    size_t len = 1 << 14;  // length after the folds applied so far
    F128 rr[8];
    for (int rnd = 0; rnd < 8; ++rnd) {
      const size_t half = len / 2;
      const bool fold = rnd > 0;
      const F128 rp = fold ? rr[rnd - 1] : F128();
      // chunks of at least 512 positions: smaller rounds run on this thread
      const size_t grain = std::max<size_t>(512, (half + 63) / 64);
      const size_t nch = (half + grain - 1) / grain;
      std::vector<F256> p1(nch), pi(nch);
      pool.parallel_for(half, grain, [&](size_t c, size_t b, size_t e) {
        F256 a1, ai;
        for (size_t i = b; i < e; ++i) {
          F128 c0 = comb[i], c1 = comb[i + half], z0 = zv[i], z1 = zv[i + half];
          if (fold) {
            c0 += rp * (comb[i + len] + c0);
            c1 += rp * (comb[i + half + len] + c1);
            z0 += rp * (zv[i + len] + z0);
            z1 += rp * (zv[i + half + len] + z1);
            comb[i] = c0;
            comb[i + half] = c1;
            zv[i] = z0;
            zv[i + half] = z1;
          }
          a1 ^= mul_unreduced(c1, z1);
          ai ^= mul_unreduced(c0 + c1, z0 + z1);
        }
        p1[c] = a1;
        pi[c] = ai;
      });
      F256 s1, si;
      for (size_t c = 0; c < nch; ++c) {
        s1 ^= p1[c];
        si ^= pi[c];
      }
      const F128 e1 = reduce(s1), einf = reduce(si);
      lc_rounds[rnd][0] = e1;
      lc_rounds[rnd][1] = einf;
      t.observe(e1);
      t.observe(einf);
      rr[rnd] = t.sample();
      len = half;
    }
    // the last round's fold (only z_vec is needed further)
    for (size_t i = 0; i < len; ++i) zv[i] += rr[7] * (zv[i + len] + zv[i]);
    for (int i = 0; i < 64; ++i) zpart[i] = zv[i];
    t.observe_slice(zpart, 64);
    r_inner_skip = t.sample();
    F128 lam[64];
    lagrange(r_inner_skip, 0, 64, 0, 64, lam);
    w_claim = dot(lam, zpart, 64);
    for (int i = 0; i < 8; ++i) r_inner_rest[i] = rr[7 - i];
  }
  B.u64(8);
  for (int i = 0; i < 8; ++i) {
    B.f(lc_rounds[i][0]);
    B.f(lc_rounds[i][1]);
  }
  B.fvec(zpart, 64);

  // ---- claims (r1cs.rs:286-330): x_full = x_inner_rest ++ x_outer
  std::vector<F128> xf[2];
  xf[0].assign(r_inner_rest, r_inner_rest + 8);
  xf[0].insert(xf[0].end(), rhos.begin() + 8, rhos.end());
  xf[1] = xf1;

  // ---- PCS open: ring switch (ring_switch.rs:2708-2780, 3040-3117)
  t.label("flock-pcs-open-batch-v0");
  void *rs_lo0 = P.dev("rs.lo0", 16ull << rs_k), *rs_hk0 = P.dev("rs.hk0", 16ull << (L - rs_k));
  void *cb = P.dev("lig.cb", N * 16), *dtab = P.dev("rs.tab", 2 * 32 * 16 * 16);
  if (!rs_lo0 || !rs_hk0 || !cb || !dtab) return FLOCK_E_OOM;
  F128 shv[2][128], sclaim[2], eqr[2][128];
  {
    // s_hat_v[k][r] = sum_i bit_r(z[i]) eq_k(i), eq_k(i) = lo_k(i mod 16) hi_k(i / 16)
    void *h0 = P.dev("rs.hi0", (N / 16) * 16), *h1 = rs_h1;
    void *dq = P.dev("rs.q", 2 * 2048 * 16);
    if (!h0 || !h1 || !dq) return FLOCK_E_OOM;
    // (the tables of claim 1 were expanded during the lincheck)
    if ((rc = P.eq_device(&xf[0][5], L - 4, h0))) return rc;
    if ((rc = bitfold_device(ctx, st, z, 4, N / 16, h0, h1, dq))) return rc;
    if ((rc = P.eq_device(&xf[0][1], rs_k, rs_lo0)) || (rc = P.eq_device(&xf[0][1 + rs_k], L - rs_k, rs_hk0)))
      return rc;
    std::vector<F128> q(2 * 2048);
    if ((rc = P.d2h(q.data(), dq, q.size() * 16))) return rc;
    for (int k = 0; k < 2; ++k) {
      std::vector<F128> lo = eq_table(&xf[k][1], 4);
      for (int rb = 0; rb < 128; ++rb) {
        F256 acc;
        for (int il = 0; il < 16; ++il) acc ^= mul_unreduced(lo[il], q[k * 2048 + il * 128 + rb]);
        shv[k][rb] = reduce(acc);
      }
    }
  }
  for (int k = 0; k < 2; ++k) {
    t.label("flock-ring-switch-v0");
    t.observe_slice(shv[k], 128);
    std::vector<F128> rd = t.sample_vec(7);
    build_eq(rd.data(), 7, eqr[k]);
    // s_hat_u = bit transpose of s_hat_v (ring_switch.rs:1654-1718)
    F128 u[128];
    for (int bb = 0; bb < 128; ++bb) {
      uint64_t lo = 0, hi = 0;
      for (int i = 0; i < 128; ++i) {
        uint64_t bit = bb < 64 ? (shv[k][i].lo >> bb) & 1 : (shv[k][i].hi >> (bb - 64)) & 1;
        if (i < 64) lo |= bit << i;
        else hi |= bit << (i - 64);
      }
      u[bb] = F128(lo, hi);
    }
    sclaim[k] = dot(u, eqr[k], 128);
  }
  F128 gam[2] = {t.sample(), t.sample()};
  {
    // nibble tables of gamma_k eq(r''_k)
    std::vector<F128> tab(2 * 32 * 16);
    for (int k = 0; k < 2; ++k) {
      F128 wv[128];
      for (int bb = 0; bb < 128; ++bb) wv[bb] = gam[k] * eqr[k][bb];
      for (int p = 0; p < 32; ++p)
        for (int v = 0; v < 16; ++v) {
          F128 s;
          for (int q = 0; q < 4; ++q)
            if ((v >> q) & 1) s += wv[4 * p + q];
          tab[k * 512 + p * 16 + v] = s;
        }
    }
    if ((rc = P.h2d(dtab, tab.data(), tab.size() * 16))) return rc;
    if ((rc = rs_basis_device(ctx, st, dtab, rs_lo0, rs_lo1, rs_k, rs_hk0, rs_hk1, L - rs_k, cb))) return rc;
  }
  const F128 target = gam[0] * sclaim[0] + gam[1] * sclaim[1];
  B.u64(2);
  B.fvec(shv[0], 128);
  B.fvec(shv[1], 128);

  // ---- Ligerito (pcs/ligerito.rs:6222-6980)
  std::vector<F128> msgs;
  std::vector<uint64_t> nonces, fold_nonces;
  std::vector<F128> oods, yr;
  std::vector<std::array<uint8_t, 32>> roots;
  std::vector<Opening> rec_open;
  Opening open0, open_last;
  {
    ScopedPhase ph(ctx, "ligerito", st);
    t.label("flock-ligerito-basis-v0");
    t.observe(target);
    t.observe_bytes(root, 32);
    void *f = z, *fb = P.dev("lig.f.a", N * 8), *fc = P.dev("lig.f.b", N * 8);
    void *cbb = P.dev("lig.cb.b", N * 8), *cbc = P.dev("lig.cb.c", N * 8);
    void *basis = P.dev("lig.basis", N * 8);
    if (!fb || !fc || !cbb || !cbc || !basis) return FLOCK_E_OOM;
    void *cur_cb = cb;
    size_t len = N;
    F128 o3[3];
    auto push = [&](F128 u0, F128 u2) {
      msgs.push_back(u0);
      msgs.push_back(u2);
      t.observe(u0);
      t.observe(u2);
    };
    auto fold = [&](F128 rv) -> int {
      void *nf = (f == fb) ? fc : fb, *ncb = (cur_cb == cbb) ? cbc : cbb;
      // fold and message in one launch, up to 2^16 pairs (above, the separate
      // fold's full byte tables of rv cost less than the fused kernel's combs)
      if (len >= 4 && len / 4 <= (1u << 16)) {
        const uint32_t seq = ctx->next_readback_seq();
        int e = lig_fold_msg_device(ctx, st, f, cur_cb, rv, len / 4, nf, ncb, ctx->readback_dev, seq);
        uint4 h[3];
        if (!e) e = P.fetch(h, 48, seq);
        if (e) return e;
        f = nf;
        cur_cb = ncb;
        len /= 2;
        push(fv(h[0]), fv(h[1]));
        return FLOCK_OK;
      }
      int e = fold_pair_device(ctx, st, f, cur_cb, len / 2, rv, nf, ncb);
      if (e) return e;
      f = nf;
      cur_cb = ncb;
      len /= 2;
      if ((e = P.msg(f, cur_cb, len, 0, o3))) return e;
      push(o3[0], o3[1]);
      return FLOCK_OK;
    };
    // The same fold and message with the challenge in device memory, the
    // message written to out3 (device) and observed by the device transcript.
    const bool dfs_lig = ctx->device_transcript;
    constexpr int LIG_MAX_FOLDS = 16;
    uint8_t *d_lfs = nullptr, *d_lwork = nullptr, *d_lmsg = nullptr, *d_lr = nullptr, *d_lnonce = nullptr;
    if (dfs_lig) {
      d_lfs = (uint8_t *)P.dev("fs.state", FS_STATE_BYTES);
      d_lwork = (uint8_t *)P.dev("lig.fs.work", FLOCK_POW_WORK_BYTES + 16);  // + the status word
      d_lmsg = (uint8_t *)P.dev("lig.fs.msgs", 48 * LIG_MAX_FOLDS);
      d_lr = (uint8_t *)P.dev("lig.fs.r", 16 * LIG_MAX_FOLDS);
      d_lnonce = (uint8_t *)P.dev("lig.fs.nonce", 8 * LIG_MAX_FOLDS);
      if (!d_lfs || !d_lwork || !d_lmsg || !d_lr || !d_lnonce) return FLOCK_E_OOM;
    }
    auto fold_dev = [&](const uint8_t *d_r, uint8_t *out3, void *d_digest) -> int {
      void *nf = (f == fb) ? fc : fb, *ncb = (cur_cb == cbb) ? cbc : cbb;
      const FsStep fs{d_lfs, nullptr, d_digest};
      int e;
      if (len >= 4 && len / 4 <= (1u << 16)) {
        e = lig_fold_msg_device(ctx, st, f, cur_cb, F128(), len / 4, nf, ncb, out3, 0, d_r, fs);
        f = nf;
        cur_cb = ncb;
        len /= 2;
        return e;
      }
      if ((e = fold_pair_device(ctx, st, f, cur_cb, len / 2, F128(), nf, ncb, d_r))) return e;
      f = nf;
      cur_cb = ncb;
      len /= 2;
      return lig_msg_device(ctx, st, f, cur_cb, len / 2, 0, out3, 0, fs);
    };
    // A level's `count` fold rounds: round j's proof of work (fold_bits - j
    // bits, when positive), its challenge, the fold and the next message
    // (pcs/ligerito.rs:6300-6340).  With the device transcript the rounds
    // run on the device back to back: the state goes up once, and the
    // messages, nonces and the state come back once.
    auto fold_rounds = [&](int fold_bits, int count) -> int {
      if (!dfs_lig) {
        for (int j = 0; j < count; ++j) {
          const int bits = fold_bits - j;
          if (bits > 0) {
            uint64_t nc;
            if (int e = P.grind(t, (uint32_t)bits, &nc)) return e;
            fold_nonces.push_back(nc);
          }
          if (int e = fold(t.sample())) return e;
        }
        return FLOCK_OK;
      }
      if (count > LIG_MAX_FOLDS) {
        ctx->last_error = "ligerito: too many fold rounds in one level";
        return FLOCK_E_SHAPE;
      }
      uint8_t init[FS_STATE_BYTES];
      t.export_state(init);
      if (int e = P.h2d(d_lfs, init, FS_STATE_BYTES)) return e;
      // the scan's best nonce (none yet), the first round's digest and the
      // status word (0)
      uint8_t work[FLOCK_POW_WORK_BYTES + 16];
      memset(work, 0, sizeof work);
      memset(work, 0xFF, 8);
      t.state_digest(work + 32);
      if (int e = P.h2d(d_lwork, work, sizeof work)) return e;
      uint8_t *d_status = d_lwork + FLOCK_POW_WORK_BYTES;
      auto grid_pow = [&](int j) { return fold_bits - j > (int)FLOCK_POW_WAVE_BITS; };
      for (int j = 0; j < count; ++j) {
        const int bits = fold_bits - j;
        if (int e = lig_challenge_device(ctx, st, d_lfs, bits > 0 ? (uint32_t)bits : 0u, d_lwork, true,
                                         d_lnonce + 8 * j, d_lr + 16 * j, d_status))
          return e;
        // the message's transcript step leaves the next proof of work's base
        void *next_dg = j + 1 < count && grid_pow(j + 1) ? lig_pow_digest(d_lwork) : nullptr;
        if (int e = fold_dev(d_lr + 16 * j, d_lmsg + 48 * j, next_dg)) return e;
      }
      const size_t nm = 48 * (size_t)count, nn = 8 * (size_t)count;
      if (nm + nn + 16 + FS_STATE_BYTES > ctx->landing_bytes) {
        ctx->last_error = "ligerito: landing area too small";
        return FLOCK_E_DEVICE;
      }
      uint8_t *land = ctx->landing;
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land, d_lmsg, nm, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm, d_lnonce, nn, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm + nn, d_status, 16, hipMemcpyDeviceToHost, st));
      FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(land + nm + nn + 16, d_lfs, FS_STATE_BYTES, hipMemcpyDeviceToHost, st));
      {
        HostPhase hp(ctx, "wait_d2h");
        if (int e = P.wait_signal()) return e;
      }
      uint32_t status;
      memcpy(&status, land + nm + nn, 4);
      if (status) {
        ctx->last_error = "ligerito: the proof of work found no nonce in its search";
        return FLOCK_E_DEVICE;
      }
      for (int j = 0; j < count; ++j) {
        if (fold_bits - j > 0) {
          uint64_t nc;
          memcpy(&nc, land + nm + 8 * j, 8);
          fold_nonces.push_back(nc);
        }
        F128 u0, u2;
        memcpy(&u0, land + 48 * j, 16);
        memcpy(&u2, land + 48 * j + 16, 16);
        msgs.push_back(u0);  // observed on the device
        msgs.push_back(u2);
      }
      t.import_state(land + nm + nn + 16);
      return FLOCK_OK;
    };
    auto glue = [&](const void *bn, F128 be) { return axpy_device(ctx, st, cur_cb, bn, len, be); };
    if ((rc = P.msg(f, cur_cb, len, 0, o3))) return rc;
    push(o3[0], o3[1]);
    const int ik = cfg->initial_k;
    if ((rc = fold_rounds(cfg->levels[0].fold_grinding_bits, ik))) return rc;
    const int n1 = L - ik;
    Witness prev;
    prev.cw = cw0;
    prev.tree = tree0;
    prev.lanes = 64;
    prev.block_len = npos0;
    prev.dim = k0;
    prev.rate = 1;
    Witness nxt;
    if ((rc = P.commit_level(f, n1 - cfg->levels[1].k_recursive, cfg->levels[1].k_recursive,
                             cfg->levels[1].log_inv_rate, "1", nxt)))
      return rc;
    std::array<uint8_t, 32> rt;
    memcpy(rt.data(), nxt.root, 32);
    roots.push_back(rt);
    t.observe_bytes(nxt.root, 32);
    auto ood = [&](int count, int nvars) -> int {
      void *deq = P.dev("lig.eqz", (16ull << nvars));
      if (!deq) return FLOCK_E_OOM;
      for (int o = 0; o < count; ++o) {
        std::vector<F128> zo = t.sample_vec(nvars);
        int e = P.eq_device(zo.data(), nvars, deq);
        if (!e) e = P.msg(f, deq, len, 1, o3);
        if (e) return e;
        t.observe(o3[2]);
        oods.push_back(o3[2]);
        push(o3[0], o3[1]);
        F128 be = t.sample();
        if ((e = glue(deq, be))) return e;
      }
      return FLOCK_OK;
    };
    if ((rc = ood(cfg->levels[1].ood_samples, n1))) return rc;
    uint64_t nc;
    if ((rc = P.grind(t, (uint32_t)cfg->levels[0].grinding_bits, &nc))) return rc;
    nonces.push_back(nc);
    std::vector<size_t> q = sample_queries(t, prev.block_len, (size_t)cfg->levels[0].queries);
    std::vector<F128> al = t.sample_vec((size_t)ceil_log2(q.size()));
    if ((rc = P.open(prev, q, open0, true))) return rc;
    if ((rc = P.induce(prev, q, al, n1, basis))) return rc;
    if ((rc = P.msg(f, basis, len, 0, o3))) return rc;
    push(o3[0], o3[1]);
    F128 be = t.sample();
    if ((rc = glue(basis, be))) return rc;
    prev = nxt;
    const int R = cfg->n_levels - 1;
    for (int i = 0; i < R; ++i) {
      const int ki = cfg->levels[i + 1].k_recursive;
      if ((rc = fold_rounds(cfg->levels[i + 1].fold_grinding_bits, ki))) return rc;
      if (i == R - 1) {
        yr.resize(len);
        if ((rc = P.d2h(yr.data(), f, len * 16))) return rc;
        for (auto &v : yr) t.observe(v);
        if ((rc = P.grind(t, (uint32_t)cfg->levels[i + 1].grinding_bits, &nc))) return rc;
        nonces.push_back(nc);
        q = sample_queries(t, prev.block_len, (size_t)cfg->levels[i + 1].queries);
        if ((rc = P.open(prev, q, open_last, false))) return rc;
        break;
      }
      int nn = 0;
      while (((size_t)1 << nn) < len) ++nn;
      const int kn = cfg->levels[i + 2].k_recursive;
      Witness w2;
      const char *tag = (i % 2 == 0) ? "2" : "1";
      if ((rc = P.commit_level(f, nn - kn, kn, cfg->levels[i + 2].log_inv_rate, tag, w2))) return rc;
      memcpy(rt.data(), w2.root, 32);
      roots.push_back(rt);
      t.observe_bytes(w2.root, 32);
      if ((rc = ood(cfg->levels[i + 2].ood_samples, nn))) return rc;
      if ((rc = P.grind(t, (uint32_t)cfg->levels[i + 1].grinding_bits, &nc))) return rc;
      nonces.push_back(nc);
      q = sample_queries(t, prev.block_len, (size_t)cfg->levels[i + 1].queries);
      al = t.sample_vec((size_t)ceil_log2(q.size()));
      Opening op;
      if ((rc = P.open(prev, q, op, true))) return rc;
      rec_open.push_back(op);
      if ((rc = P.induce(prev, q, al, nn, basis))) return rc;
      if ((rc = P.msg(f, basis, len, 0, o3))) return rc;
      push(o3[0], o3[1]);
      be = t.sample();
      if ((rc = glue(basis, be))) return rc;
      prev = w2;
    }
  }
  // the deferred openings' bytes are in host memory once the stream has
  // drained past their copies (the last opening's wait normally did that)
  if (P.openings_in_flight && (rc = P.wait_signal())) return rc;
  finish_opening(open0);
  for (auto &o : rec_open) finish_opening(o);
  finish_opening(open_last);
  // LigeritoProof (pcs/ligerito.rs:1504-1545)
  {
    uint8_t root0[32];
    memcpy(root0, root, 32);
    B.put(root0, 32);
    put_opening(B, open0);
    B.u64(roots.size());
    for (auto &rt : roots) B.put(rt.data(), 32);
    B.u64(rec_open.size());
    for (auto &o : rec_open) put_opening(B, o);
    B.fvec(yr.data(), yr.size());
    put_opening(B, open_last);
    B.u64(msgs.size() / 2);
    B.put(msgs.data(), msgs.size() * 16);
    B.u64(nonces.size());
    B.put(nonces.data(), nonces.size() * 8);
    B.fvec(oods.data(), oods.size());
    B.u64(fold_nonces.size());
    B.put(fold_nonces.data(), fold_nonces.size() * 8);
  }
  if (ph_all) ph_all->end();
  if ((rc = ctx->wait_stream(st))) return rc;
  out.swap(B.v);
  return FLOCK_OK;
}

}  // namespace

}  // namespace flock
