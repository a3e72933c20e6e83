// The proof verifier: Blake3Setup::verify (crates/flock-prover/src/
// r1cs_hashes/blake3.rs:3698) preceded by the harness's checks of the bundle
// (benchmark-tools/harness/src/main.rs:319-366).  It parses the bundle
// bytes, replays the Fiat-Shamir transcript (Appendix-B order: statement,
// zerocheck, lincheck, ring switch, Ligerito) and checks every sumcheck
// identity, proof of work, Merkle opening and the final residual.  The
// commitment root is compared with a root recomputed from the compressions
// (witness + commit on the GPU) by the caller, when it has them.
//
// Host work only: O(1)-size algebra per round, 218 + 106 + ... opened rows,
// and the combined ring-switch basis evaluated succinctly in the tensor
// algebra GF(2^128) (x) GF(2^128) (a product of L factors) rather than as a
// dense 2^L table.
#include "verifier.h"

#include <string.h>

#include <array>
#include <string>
#include <vector>

#include "host_algebra.h"
#include "host_blake3.h"
#include "host_field.h"
#include "ligerito_configs.h"
#include "r1cs.h"
#include "transcript.h"

namespace flock {

namespace {

const F128 ONE = F128::one();

// ------------------------------------------------------------------ bundle
// R1csProofBundleLigerito (proof_io.rs:161-208): "FLOCK" 04 02 + bincode 1.x
// fixint, fields in struct declaration order (proof.rs:18-22,
// zerocheck.rs:217-231, lincheck.rs:436-447, pcs.rs:51, ring_switch.rs:2503,
// pcs/ligerito.rs:1504-1545).
struct Opening {
  size_t n = 0, lanes = 0;
  std::vector<F128> rows;  // n x lanes
  std::vector<uint8_t> hashes;  // 32 x k
};

struct Bundle {
  uint8_t root[32];
  uint64_t m = 0, log_inv_rate = 0, log_batch_size = 0;
  uint32_t profile = 0, merkle_hash = 0;
  std::vector<F128> r1ab, r1c;
  std::vector<std::array<F128, 2>> zc_rounds;
  F128 final_a, final_b, final_c;
  std::vector<std::array<F128, 2>> lc_rounds;
  std::vector<F128> z_partial;
  std::vector<std::vector<F128>> s_hat_v;
  uint8_t lig_root[32];
  Opening initial;
  std::vector<std::array<uint8_t, 32>> roots;
  std::vector<Opening> rec;
  std::vector<F128> yr;
  Opening final_open;
  std::vector<std::array<F128, 2>> msgs;
  std::vector<uint64_t> nonces;
  std::vector<F128> ood;
  std::vector<uint64_t> fold_nonces;
};

struct Reader {
  const uint8_t *p;
  size_t n, o = 0;
  bool ok = true;
  bool take(void *dst, size_t k) {
    if (!ok || k > n - o) return ok = false;
    memcpy(dst, p + o, k);
    o += k;
    return true;
  }
  uint64_t u64() {
    uint64_t v = 0;
    take(&v, 8);
    return v;
  }
  uint32_t u32() {
    uint32_t v = 0;
    take(&v, 4);
    return v;
  }
  F128 f() {
    F128 v;
    take(&v, 16);
    return v;
  }
  // a length prefix for elements of `size` bytes: at most what is left
  uint64_t len(size_t size) {
    uint64_t k = u64();
    if (ok && (k > (n - o) / size)) ok = false;
    return ok ? k : 0;
  }
  void fvec(std::vector<F128> &v) {
    v.resize(len(16));
    if (!v.empty()) take(v.data(), 16 * v.size());
  }
  void pairs(std::vector<std::array<F128, 2>> &v) {
    v.resize(len(32));
    if (!v.empty()) take(v.data(), 32 * v.size());
  }
  void u64s(std::vector<uint64_t> &v) {
    v.resize(len(8));
    if (!v.empty()) take(v.data(), 8 * v.size());
  }
  void opening(Opening &op) {
    op.n = len(8);
    std::vector<F128> row;
    for (size_t i = 0; i < op.n && ok; ++i) {
      fvec(row);
      if (i == 0) op.lanes = row.size();
      if (row.size() != op.lanes) ok = false;
      op.rows.insert(op.rows.end(), row.begin(), row.end());
    }
    op.hashes.resize(32 * len(32));
    if (!op.hashes.empty()) take(op.hashes.data(), op.hashes.size());
  }
};

bool parse(const uint8_t *b, size_t n, Bundle &P, std::string &err) {
  Reader R{b, n};
  uint8_t head[7];
  if (!R.take(head, 7) || memcmp(head, "FLOCK", 5) || head[5] != 4 || head[6] != 2) {
    err = "bundle: bad header";
    return false;
  }
  R.take(P.root, 32);
  P.m = R.u64();
  P.log_inv_rate = R.u64();
  P.log_batch_size = R.u64();
  P.profile = R.u32();
  P.merkle_hash = R.u32();
  R.fvec(P.r1ab);
  R.fvec(P.r1c);
  R.pairs(P.zc_rounds);
  P.final_a = R.f();
  P.final_b = R.f();
  P.final_c = R.f();
  R.pairs(P.lc_rounds);
  R.fvec(P.z_partial);
  P.s_hat_v.resize(R.len(8));
  for (auto &v : P.s_hat_v) R.fvec(v);
  R.take(P.lig_root, 32);
  R.opening(P.initial);
  P.roots.resize(R.len(32));
  for (auto &rt : P.roots) R.take(rt.data(), 32);
  P.rec.resize(R.len(16));
  for (auto &op : P.rec) R.opening(op);
  R.fvec(P.yr);
  R.opening(P.final_open);
  R.pairs(P.msgs);
  R.u64s(P.nonces);
  R.fvec(P.ood);
  R.u64s(P.fold_nonces);
  if (!R.ok) {
    err = "bundle: truncated or malformed";
    return false;
  }
  if (R.o != n) {
    err = "bundle: trailing bytes";
    return false;
  }
  return true;
}

// ------------------------------------------------------------------ checks
// verify_pow (challenger.rs:706-730): the leading `bits` bits of
// BLAKE3(state digest || nonce_le || 0^24) are zero; the nonce is absorbed.
bool verify_pow(Transcript &t, uint64_t nonce, uint32_t bits) {
  bool ok;
  if (bits == 0) {
    ok = nonce == 0;
  } else {
    uint8_t d[32];
    t.state_digest(d);
    uint32_t m[16] = {0};
    memcpy(m, d, 32);
    m[8] = (uint32_t)nonce;
    m[9] = (uint32_t)(nonce >> 32);
    uint32_t o[16];
    hb3::compress(hb3::IV, m, 0, 64, hb3::CHUNK_START | hb3::CHUNK_END | hb3::ROOT, o);
    const uint64_t be = ((uint64_t)__builtin_bswap32(o[0]) << 32) | __builtin_bswap32(o[1]);
    ok = bits <= 64 && (be >> (64 - bits)) == 0;
  }
  t.observe_bytes(&nonce, 8);
  return ok;
}

// verify_merkle_multi_proof (merkle.rs:930-1005): leaves at sorted distinct
// positions, siblings from the proof in emission order, parents by BLAKE3.
bool merkle_verify_multi(const uint8_t root[32], size_t num_leaves, const std::vector<size_t> &pos,
                         std::vector<std::array<uint8_t, 32>> h, const std::vector<uint8_t> &proof) {
  const size_t n_proof = proof.size() / 32;
  if (num_leaves == 0 || (num_leaves & (num_leaves - 1))) return false;
  if (pos.empty()) return n_proof == 0;
  for (size_t i = 0; i < pos.size(); ++i)
    if (pos[i] >= num_leaves || (i > 0 && pos[i - 1] >= pos[i])) return false;
  if (num_leaves == 1) return n_proof == 0 && memcmp(h[0].data(), root, 32) == 0;
  std::vector<size_t> p = pos;
  size_t used = 0, ll = num_leaves, n = pos.size();
  while (ll > 1) {
    size_t nn = 0;
    for (size_t i = 0; i < n;) {
      const uint8_t *l, *r;
      const size_t q = p[i];
      if (i + 1 < n && p[i + 1] == (q ^ 1)) {
        l = h[i].data();
        r = h[i + 1].data();
        i += 2;
      } else {
        if (used >= n_proof) return false;
        const uint8_t *s = proof.data() + 32 * used++;
        l = (q & 1) ? s : h[i].data();
        r = (q & 1) ? h[i].data() : s;
        i += 1;
      }
      uint32_t m[16], o[16];
      memcpy(m, l, 32);
      memcpy(m + 8, r, 32);
      hb3::compress(hb3::IV, m, 0, 64, hb3::PARENT, o);
      memcpy(h[nn].data(), o, 32);
      p[nn++] = q >> 1;
    }
    n = nn;
    ll >>= 1;
  }
  return used == n_proof && n == 1 && memcmp(h[0].data(), root, 32) == 0;
}

// Opened rows against a Ligero root: leaf = non-root chunk CV of the row's
// bytes (merkle.rs:200-205).
bool verify_opens(const uint8_t root[32], size_t block_len, const std::vector<size_t> &q, const Opening &op,
                  size_t lanes) {
  if (op.n != q.size() || op.lanes != lanes) return false;
  std::vector<std::array<uint8_t, 32>> lh(q.size());
  for (size_t i = 0; i < q.size(); ++i) {
    hb3::Hasher hs;
    hs.update(&op.rows[i * lanes], 16 * lanes);
    hs.finalize_non_root(lh[i].data());
  }
  return merkle_verify_multi(root, block_len, q, lh, op.hashes);
}

// sum_i eq(alpha, i) <row_i, eq(v, .)> (induce_sumcheck_poly's enforced sum,
// pcs/ligerito.rs:1989-2160).
F128 enforced_sum(const Opening &op, const std::vector<F128> &v, const std::vector<F128> &alpha) {
  const std::vector<F128> ev = eq_table(v.data(), (int)v.size()), aw = eq_table(alpha.data(), (int)alpha.size());
  F128 s;
  for (size_t i = 0; i < op.n; ++i) s += aw[i] * dot(&op.rows[i * op.lanes], ev.data(), op.lanes);
  return s;
}

// Subspace-polynomial constants: svk[i] = W_i(beta_i) for the novel basis of
// the additive NTT over span{x^1 ..} (pcs/ligerito.rs:1776-1824).
std::vector<F128> sks_vks(int n) {
  std::vector<F128> out(n + 1), layer(n + 1);
  out[0] = ONE;
  for (int i = 1; i <= n; ++i) layer[i - 1] = F128(1ull << i, 0);
  int cur = n;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < cur; ++j) {
      F128 s = layer[j] * layer[j] + out[i] * layer[j];
      if (j == 0)
        out[i + 1] = s;
      else
        layer[j - 1] = s;
    }
    --cur;
  }
  return out;
}

// Sumcheck message as a quadratic c + b X + a X^2 with c = u0, a = u2,
// b = claim + u2 (the sum over {0,1} equals the running claim).
struct Quad {
  F128 c, b, a;
};
Quad q_from(F128 u0, F128 u2, F128 tr) { return {u0, tr + u2, u2}; }
F128 q_eval(const Quad &q, F128 r) { return q.c + r * q.b + r * r * q.a; }
Quad q_fold(const Quad &p, const Quad &q, F128 al) { return {p.c + al * q.c, p.b + al * q.b, p.a + al * q.a}; }

// One level's induced basis: basis(p) = sum_i eq(alpha, i) prod_k (1 + p_k (1
// + W^_k(q_i))) with W^_k the normalized subspace polynomials at the query
// (ligerito.rs:1989-2160).  The per-query W^ values do not depend on p.
struct Induced {
  int n = 0;
  std::vector<F128> aw;                // eq(alpha, i)
  std::vector<std::vector<F128>> w;    // W^_k(q_i), k < n
  F128 beta;
  int ris_start = 0;
};

Induced make_induced(int n, const std::vector<size_t> &q, const std::vector<F128> &alpha) {
  Induced I;
  I.n = n;
  I.aw = eq_table(alpha.data(), (int)alpha.size());
  const std::vector<F128> svk = sks_vks(n);
  std::vector<F128> isvk(n);
  for (int i = 0; i < n; ++i) isvk[i] = inv(svk[i]);
  I.w.resize(q.size());
  for (size_t i = 0; i < q.size(); ++i) {
    std::vector<F128> &w = I.w[i];
    w.resize(n);
    if (n == 0) continue;
    w[0] = F128((uint64_t)q[i], 0);
    for (int k = 1; k < n; ++k) w[k] = w[k - 1] * w[k - 1] + svk[k - 1] * w[k - 1];
    for (int k = 0; k < n; ++k) w[k] = w[k] * isvk[k];
  }
  return I;
}

F128 induced_eval(const Induced &I, const F128 *p) {
  F128 sum;
  for (size_t i = 0; i < I.w.size(); ++i) {
    F128 prod = ONE;
    for (int k = 0; k < I.n; ++k) prod = prod * (ONE + p[k] * (ONE + I.w[i][k]));
    sum += I.aw[i] * prod;
  }
  return sum;
}

struct Ood {
  std::vector<F128> z;
  F128 beta;
  int ris_start = 0;
};

// ---------------------------------------------------------------- tensor algebra
// The combined ring-switch basis at a point p (the dense table of the
// prover, ring_switch.rs:1720-1760, folded at p):
//   b(p) = sum_k gamma_k sum_b eq(r''_k)[b] c_b,  c_b = sum_i eq(p, i) bit_b(eq(x_k, i)).
// With T = sum_i eq(p, i) (x) eq(x_k, i) = sum_b c_b (x) X^b in GF(2^128) (x)
// GF(2^128), the sum over i factorizes:
//   T = prod_j ((1 + p_j) (x) (1 + x_j) + p_j (x) x_j).
// Multiplying T = sum_b c_b (x) X^b by that factor gives
//   c' = (1 + p_j) c + M_{x_j} c,  (M_x c)_b' = sum_b bit_b'(X^b x) c_b.
struct Tensor {
  F128 c[128];
};

void tensor_apply(Tensor &T, F128 xj, bool p_bit_known, bool p_bit, F128 pj) {
  F128 d[128];
  for (int b = 0; b < 128; ++b) d[b] = F128();
  F128 xb = xj;  // X^b x
  for (int b = 0; b < 128; ++b) {
    const F128 cb = T.c[b];
    uint64_t lo = xb.lo, hi = xb.hi;
    while (lo) {
      const int k = __builtin_ctzll(lo);
      d[k] += cb;
      lo &= lo - 1;
    }
    while (hi) {
      const int k = __builtin_ctzll(hi);
      d[64 + k] += cb;
      hi &= hi - 1;
    }
    xb = mul_x(xb);
  }
  if (p_bit_known) {
    // p_j = 0: c + M c; p_j = 1: M c
    for (int b = 0; b < 128; ++b) T.c[b] = p_bit ? d[b] : T.c[b] + d[b];
  } else {
    const F128 s = ONE + pj;
    for (int b = 0; b < 128; ++b) T.c[b] = s * T.c[b] + d[b];
  }
}

// ------------------------------------------------------------------ protocol
struct ZcClaim {
  F128 z;
  std::vector<F128> r, rhos;
};

struct LcClaim {
  F128 r_inner_skip, w;
  F128 r_inner_rest[8];
};

struct Claim {
  F128 z_skip, value;
  std::vector<F128> x_full;  // L + 1
};

// zerocheck verify (zerocheck.rs:1365-1520)
bool zerocheck_verify(const Bundle &P, int m, Transcript &t, ZcClaim &cl, std::string &err) {
  const int n = m - 6;
  if (P.r1ab.size() != 64 || P.r1c.size() != 64 || (int)P.zc_rounds.size() != n) {
    err = "zerocheck: wrong shape";
    return false;
  }
  t.label("flock-zerocheck-v0");
  cl.r.assign(m, F128());
  {
    std::vector<F128> rs = t.sample_vec(6), ro = t.sample_vec(m - 13);
    for (int i = 0; i < 6; ++i) cl.r[i] = rs[i];
    for (int i = 0; i < 3; ++i) cl.r[6 + i] = phi8(SMALL_CHAL[i]);
    for (int i = 0; i < 4; ++i) cl.r[9 + i] = medium(i);
    for (int i = 0; i < m - 13; ++i) cl.r[13 + i] = ro[i];
  }
  t.observe_slice(P.r1ab.data(), 64);
  t.observe_slice(P.r1c.data(), 64);
  const F128 z = t.sample();
  F128 wl[64], wc[64];
  lagrange(z, 64, 64, 64, 64, wl);
  const F128 pcz = dot(wl, P.r1c.data(), 64);
  if (!(pcz == P.final_c)) {
    err = "zerocheck: final c evaluation";
    return false;
  }
  // P^AB + P^C vanishes on S: interpolate over S u Lambda with zeros on S
  lagrange(z, 64, 64, 0, 128, wc);
  F128 running = pcz;
  for (int i = 0; i < 64; ++i) running += wc[i] * (P.r1ab[i] + P.r1c[i]);
  cl.rhos.resize(n);
  for (int i = 0; i < n; ++i) {
    const F128 req = cl.r[6 + i], g1 = P.zc_rounds[i][0], ginf = P.zc_rounds[i][1];
    const F128 g0 = (running + req * g1) * inv(ONE + req);
    t.observe(g1);
    t.observe(ginf);
    const F128 rho = t.sample();
    cl.rhos[i] = rho;
    const F128 opr = ONE + rho;
    running = g0 * opr + g1 * rho + ginf * (rho * opr);
  }
  if (!(running == P.final_a * P.final_b)) {
    err = "zerocheck: final claim";
    return false;
  }
  t.observe(P.final_a);
  t.observe(P.final_b);
  cl.z = z;
  return true;
}

// lincheck verify (lincheck.rs:1666-1830)
bool lincheck_verify(const Bundle &P, const ZcClaim &zc, Transcript &t, LcClaim &cl, std::string &err) {
  if (P.lc_rounds.size() != 8 || P.z_partial.size() != 64) {
    err = "lincheck: wrong shape";
    return false;
  }
  t.label("flock-lincheck-v0");
  const F128 alpha = t.sample();
  // quirky eq table (lincheck.rs:1127-1142): index i_skip + 64 i_rest
  std::vector<F128> eq(r1cs::K);
  {
    F128 lam[64];
    lagrange(zc.z, 0, 64, 0, 64, lam);
    const std::vector<F128> er = eq_table(zc.rhos.data(), 8);
    for (int ir = 0; ir < 256; ++ir)
      for (int is = 0; is < 64; ++is) eq[is + 64 * ir] = lam[is] * er[ir];
  }
  // comb[c] = alpha (A_0^T eq)[c] + (B_0^T eq)[c] (fold_alpha_batched)
  const r1cs::Matrices &mx = r1cs::matrices();
  std::vector<F128> comb(r1cs::K);
  for (int c = 0; c < r1cs::K; ++c) {
    F128 ta, tb;
    for (uint32_t k = mx.a_t.ptr[c]; k < mx.a_t.ptr[c + 1]; ++k) ta += eq[mx.a_t.cols[k]];
    for (uint32_t k = mx.b_t.ptr[c]; k < mx.b_t.ptr[c + 1]; ++k) tb += eq[mx.b_t.cols[k]];
    comb[c] = alpha * ta + tb;
  }
  F128 target = alpha * P.final_a + P.final_b;
  const F128 beta = t.sample();
  comb[r1cs::Z_CONST] += beta;
  target += beta;
  F128 running = target, rr[8];
  size_t len = r1cs::K;
  for (int rnd = 0; rnd < 8; ++rnd) {
    const F128 e1 = P.lc_rounds[rnd][0], einf = P.lc_rounds[rnd][1];
    t.observe(e1);
    t.observe(einf);
    const F128 r = t.sample();
    const F128 e0 = running + e1, c1 = e0 + e1 + einf;
    running = einf * (r * r) + c1 * r + e0;
    const size_t half = len / 2;
    for (size_t i = 0; i < half; ++i) comb[i] = comb[i] + r * (comb[i + half] + comb[i]);
    len = half;
    rr[rnd] = r;
  }
  t.observe_slice(P.z_partial.data(), 64);
  if (!(dot(comb.data(), P.z_partial.data(), 64) == running)) {
    err = "lincheck: final claim";
    return false;
  }
  cl.r_inner_skip = t.sample();
  F128 lam[64];
  lagrange(cl.r_inner_skip, 0, 64, 0, 64, lam);
  cl.w = dot(lam, P.z_partial.data(), 64);
  for (int i = 0; i < 8; ++i) cl.r_inner_rest[i] = rr[7 - i];
  return true;
}

// PCS open verify: ring switch (ring_switch.rs:3346-3520) for every claim,
// then recursive_verifier_with_basis_succinct (pcs/ligerito.rs:6981-7546).
bool pcs_verify(const Bundle &P, int m, const flock_lig_config &cfg, const Claim cl[2], Transcript &t,
                std::string &err) {
  const int L = m - 7;
  if (P.s_hat_v.size() != 2 || P.s_hat_v[0].size() != 128 || P.s_hat_v[1].size() != 128) {
    err = "ring switch: wrong shape";
    return false;
  }
  t.label("flock-pcs-open-batch-v0");
  F128 eq_rd[2][128], sclaim[2];
  for (int k = 0; k < 2; ++k) {
    const F128 *shv = P.s_hat_v[k].data();
    t.label("flock-ring-switch-v0");
    t.observe_slice(shv, 128);
    // build_claim_weights (ring_switch.rs:138-160)
    F128 lam[64], acc;
    lagrange(cl[k].z_skip, 0, 64, 0, 64, lam);
    const F128 x0 = cl[k].x_full[0];
    for (int i = 0; i < 128; ++i) acc += lam[i & 63] * ((i >> 6) ? x0 : ONE + x0) * shv[i];
    if (!(acc == cl[k].value)) {
      err = "ring switch: claim " + std::to_string(k);
      return false;
    }
    const std::vector<F128> rd = t.sample_vec(7);
    build_eq(rd.data(), 7, eq_rd[k]);
    // s_hat_u = bit transpose of s_hat_v (ring_switch.rs:1654-1718)
    F128 s;
    for (int b = 0; b < 128; ++b) {
      uint64_t lo = 0, hi = 0;
      for (int i = 0; i < 128; ++i) {
        const uint64_t bit = b < 64 ? (shv[i].lo >> b) & 1 : (shv[i].hi >> (b - 64)) & 1;
        if (i < 64)
          lo |= bit << i;
        else
          hi |= bit << (i - 64);
      }
      s += F128(lo, hi) * eq_rd[k][b];
    }
    sclaim[k] = s;
  }
  F128 gam[2];
  for (int k = 0; k < 2; ++k) gam[k] = t.sample();
  F128 tr = gam[0] * sclaim[0] + gam[1] * sclaim[1];

  // ---- Ligerito
  const int ik = cfg.initial_k, R = cfg.n_levels - 1;
  if (memcmp(P.lig_root, P.root, 32)) {
    err = "ligerito: initial root differs from the commitment";
    return false;
  }
  t.label("flock-ligerito-basis-v0");
  t.observe(tr);
  t.observe_bytes(P.lig_root, 32);
  size_t tx = 0, fni = 0, oi = 0, ni = 0;
  auto next_msg = [&](Quad &q, F128 claim) -> bool {
    if (tx >= P.msgs.size()) return false;
    t.observe(P.msgs[tx][0]);
    t.observe(P.msgs[tx][1]);
    q = q_from(P.msgs[tx][0], P.msgs[tx][1], claim);
    ++tx;
    return true;
  };
  auto fail = [&](const char *what) {
    err = std::string("ligerito: ") + what;
    return false;
  };
  Quad run;
  if (!next_msg(run, tr)) return fail("missing sumcheck message");
  std::vector<F128> ris;
  std::vector<Induced> lv;
  std::vector<Ood> od;
  // one sumcheck fold with its fold-grind (ligerito.rs:6369-6385)
  auto fold_round = [&](int bits, std::vector<F128> *level_r) -> bool {
    if (bits > 0) {
      if (fni >= P.fold_nonces.size() || !verify_pow(t, P.fold_nonces[fni++], (uint32_t)bits)) return false;
    }
    const F128 rj = t.sample();
    ris.push_back(rj);
    if (level_r) level_r->push_back(rj);
    tr = q_eval(run, rj);
    return next_msg(run, tr);
  };
  // out-of-domain samples of the committed level (nvars variables)
  auto ood_round = [&](int count, int nvars) -> bool {
    for (int o = 0; o < count; ++o) {
      Ood c;
      c.z = t.sample_vec(nvars);
      if (oi >= P.ood.size()) return false;
      const F128 y = P.ood[oi++];
      t.observe(y);
      Quad iq;
      if (!next_msg(iq, y)) return false;
      c.beta = t.sample();
      run = q_fold(run, iq, c.beta);
      tr += c.beta * y;
      c.ris_start = (int)ris.size();
      od.push_back(std::move(c));
    }
    return true;
  };
  // queries of a committed level: grind, sample, open, induced basis
  auto query_round = [&](uint32_t bits, size_t block_len, size_t queries, const uint8_t *root, const Opening &op,
                         size_t lanes, const std::vector<F128> &fold_r, int log_msg_cols, bool glue) -> bool {
    if (ni >= P.nonces.size() || !verify_pow(t, P.nonces[ni++], bits)) return false;
    const std::vector<size_t> q = sample_queries(t, block_len, queries);
    const std::vector<F128> al = t.sample_vec((size_t)ceil_log2(q.size()));
    if (!verify_opens(root, block_len, q, op, lanes)) return false;
    const F128 es = enforced_sum(op, fold_r, al);
    Quad iq;
    if (glue && !next_msg(iq, es)) return false;
    Induced I = make_induced(log_msg_cols, q, al);
    I.beta = t.sample();
    if (glue) run = q_fold(run, iq, I.beta);
    tr += I.beta * es;
    I.ris_start = (int)ris.size();
    lv.push_back(std::move(I));
    return true;
  };

  std::vector<F128> r0;
  for (int j = 0; j < ik; ++j)
    if (!fold_round(cfg.levels[0].fold_grinding_bits - j, &r0)) return fail("initial folds");
  if (P.roots.empty()) return fail("missing level root");
  t.observe_bytes(P.roots[0].data(), 32);
  const int n1 = L - ik;
  if (!ood_round(cfg.levels[1].ood_samples, n1)) return fail("out-of-domain samples");
  const size_t block_len_0 = (size_t)1 << (L - ik + cfg.levels[0].log_inv_rate);
  if (!query_round((uint32_t)cfg.levels[0].grinding_bits, block_len_0, (size_t)cfg.levels[0].queries, P.lig_root,
                   P.initial, (size_t)1 << ik, r0, n1, true))
    return fail("initial openings");
  const uint8_t *prev_root = P.roots[0].data();
  int prev_k = cfg.levels[1].k_recursive, prev_cols = n1 - prev_k, prev_rate = cfg.levels[1].log_inv_rate;
  int ncur = n1;
  size_t next_root = 1, rec_i = 0;
  for (int i = 0; i < R; ++i) {
    const int ki = cfg.levels[i + 1].k_recursive;
    if (ncur < ki) return fail("level shape");
    std::vector<F128> lr;
    for (int j = 0; j < ki; ++j)
      if (!fold_round(cfg.levels[i + 1].fold_grinding_bits - j, &lr)) return fail("level folds");
    ncur -= ki;
    const size_t prev_block = (size_t)1 << (prev_cols + prev_rate), prev_lanes = (size_t)1 << prev_k;
    if (i == R - 1) {
      if (tx != P.msgs.size() || oi != P.ood.size() || fni != P.fold_nonces.size() || ni + 1 != P.nonces.size() ||
          next_root != P.roots.size() || rec_i != P.rec.size())
        return fail("unused proof items");
      if (P.yr.size() != ((size_t)1 << ncur)) return fail("final vector length");
      for (const F128 &v : P.yr) t.observe(v);
      if (!query_round((uint32_t)cfg.levels[i + 1].grinding_bits, prev_block, (size_t)cfg.levels[i + 1].queries,
                       prev_root, P.final_open, prev_lanes, lr, ncur, false))
        return fail("final openings");
      // residual: sum_y yr[y] (b(ris ++ y) + sum beta_k basis_k(.. y) + ood terms) == tr
      const int yl = ncur, nris = (int)ris.size();
      if (nris + yl != L) return fail("variable count");
      Tensor base[2];
      for (int k = 0; k < 2; ++k) {
        for (int b = 0; b < 128; ++b) base[k].c[b] = F128();
        base[k].c[0] = ONE;
        for (int j = 0; j < nris; ++j) tensor_apply(base[k], cl[k].x_full[1 + j], false, false, ris[j]);
      }
      F128 inner;
      for (size_t y = 0; y < P.yr.size(); ++y) {
        F128 comb;
        for (int k = 0; k < 2; ++k) {
          Tensor T = base[k];
          for (int e = 0; e < yl; ++e) tensor_apply(T, cl[k].x_full[1 + nris + e], true, (y >> e) & 1, F128());
          comb += gam[k] * dot(eq_rd[k], T.c, 128);
        }
        for (const Induced &I : lv) {
          std::vector<F128> p(I.n);
          const int nfold = I.n - yl;
          for (int e = 0; e < nfold; ++e) p[e] = ris[I.ris_start + e];
          for (int e = 0; e < yl; ++e) p[nfold + e] = ((y >> e) & 1) ? ONE : F128();
          comb += I.beta * induced_eval(I, p.data());
        }
        for (const Ood &c : od) {
          const int folded = (int)c.z.size() - yl;
          F128 sc = c.beta;
          for (int e = 0; e < folded; ++e) sc = sc * (ONE + c.z[e] + ris[c.ris_start + e]);
          for (int e = 0; e < yl; ++e) {
            const F128 ze = c.z[folded + e];
            sc = sc * (((y >> e) & 1) ? ze : ONE + ze);
          }
          comb += sc;
        }
        inner += P.yr[y] * comb;
      }
      if (!(inner == tr)) return fail("final residual");
      return true;
    }
    if (next_root >= P.roots.size()) return fail("missing level root");
    const uint8_t *root_next = P.roots[next_root++].data();
    t.observe_bytes(root_next, 32);
    if (!ood_round(cfg.levels[i + 2].ood_samples, ncur)) return fail("out-of-domain samples");
    if (rec_i >= P.rec.size()) return fail("missing level opening");
    if (!query_round((uint32_t)cfg.levels[i + 1].grinding_bits, prev_block, (size_t)cfg.levels[i + 1].queries,
                     prev_root, P.rec[rec_i++], prev_lanes, lr, ncur, true))
      return fail("level openings");
    prev_root = root_next;
    const int kn = cfg.levels[i + 2].k_recursive;
    if (ncur < kn) return fail("level shape");
    prev_k = kn;
    prev_cols = ncur - kn;
    prev_rate = cfg.levels[i + 2].log_inv_rate;
  }
  return fail("no final level");
}

}  // namespace

int verify_blake3(const uint8_t *bytes, size_t len, uint64_t n_blocks, const uint8_t *domain, size_t domain_len,
                  const uint8_t *recomputed_root, std::string &err) {
  Bundle P;
  if (!parse(bytes, len, P, err)) return FLOCK_E_VERIFY;
  int n_log = 0;
  while ((1ull << n_log) < (n_blocks < 8 ? 8 : n_blocks)) ++n_log;
  const int m = r1cs::K_LOG + n_log;
  const flock_lig_config *cfg = nullptr;
  for (int i = 0; i < FLOCK_LIG_N_CONFIGS; ++i)
    if (FLOCK_LIG_CONFIGS[i].m == m && FLOCK_LIG_CONFIGS[i].profile == 0) cfg = &FLOCK_LIG_CONFIGS[i];
  if (!cfg) {
    err = "verify: no Ligerito configuration for m = " + std::to_string(m);
    return FLOCK_E_SHAPE;
  }
  if (P.m != (uint64_t)m || P.log_inv_rate != 1 || P.log_batch_size != 6 || P.profile != 0 || P.merkle_hash != 1) {
    err = "verify: commitment parameters do not match the instance";
    return FLOCK_E_VERIFY;
  }
  if (recomputed_root && memcmp(recomputed_root, P.root, 32)) {
    err = "verify: commitment root differs from the witness of the compressions";
    return FLOCK_E_VERIFY;
  }
  Transcript t(domain, domain_len);
  uint8_t dig[32];
  r1cs::statement_digest(m, dig);
  t.label("flock-r1cs-v0");
  t.observe_bytes(dig, 32);
  t.observe_bytes(P.root, 32);
  ZcClaim zc;
  if (!zerocheck_verify(P, m, t, zc, err)) return FLOCK_E_VERIFY;
  LcClaim lc;
  if (!lincheck_verify(P, zc, t, lc, err)) return FLOCK_E_VERIFY;
  // claims (r1cs.rs:286-330): ab at (r_inner_skip, r_inner_rest ++ x_outer),
  // c at (z, r_rest)
  const int n = m - 6;
  Claim cl[2];
  cl[0].z_skip = lc.r_inner_skip;
  cl[0].value = lc.w;
  cl[0].x_full.assign(lc.r_inner_rest, lc.r_inner_rest + 8);
  cl[0].x_full.insert(cl[0].x_full.end(), zc.rhos.begin() + 8, zc.rhos.end());
  cl[1].z_skip = zc.z;
  cl[1].value = P.final_c;
  cl[1].x_full.assign(zc.r.begin() + 6, zc.r.begin() + 6 + n);
  if (!pcs_verify(P, m, *cfg, cl, t, err)) return FLOCK_E_VERIFY;
  return FLOCK_OK;
}

}  // namespace flock
