// R1CS construction by a symbolic pass over one BLAKE3 compression.  Each
// 32-bit value is carried as 32 GF(2) linear forms over the K witness slots
// (K-bit sets); a carry row reads X[i] ^ cin[i] / Y[i] ^ cin[i] with
// cin[i] = xor_{j<i} carry[j].  XOR of sets equals the reference's sorted,
// square-free row lists (r1cs_hashes/common.rs:768-784), so rows and digest
// come out identical.
#include "r1cs.h"

#include <string.h>

#include <array>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "host_blake3.h"
#include "stmt_digest.h"

namespace flock {
namespace r1cs {

static const int G_LANES[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                  {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

int msg_index(int r, int g, int which) {
  static const int P[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  int perm[16], nxt[16];
  for (int i = 0; i < 16; ++i) perm[i] = i;
  for (int k = 0; k < r; ++k) {
    for (int i = 0; i < 16; ++i) nxt[i] = perm[P[i]];
    memcpy(perm, nxt, sizeof perm);
  }
  return perm[2 * g + which];
}

namespace {
constexpr int KW = K / 64;
using Set = std::vector<uint64_t>;  // KW words
struct Word {
  std::vector<Set> bit;
  Word() : bit(32, Set(KW, 0)) {}
};
inline void toggle(Set &s, int i) { s[i >> 6] ^= 1ull << (i & 63); }
Word slot(int base) {
  Word w;
  for (int i = 0; i < 32; ++i) toggle(w.bit[i], base + i);
  return w;
}
Word constant(uint32_t v) {
  Word w;
  for (int i = 0; i < 32; ++i)
    if ((v >> i) & 1) toggle(w.bit[i], Z_CONST);
  return w;
}
Word wxor(const Word &x, const Word &y) {
  Word o;
  for (int i = 0; i < 32; ++i)
    for (int k = 0; k < KW; ++k) o.bit[i][k] = x.bit[i][k] ^ y.bit[i][k];
  return o;
}
Word wrotr(const Word &x, int n) {
  Word o;
  for (int i = 0; i < 32; ++i) o.bit[i] = x.bit[(i + n) % 32];
  return o;
}

struct Builder {
  std::vector<Set> A, B;
  Builder() : A(K, Set(KW, 0)), B(K, Set(KW, 0)) {}
  Word add(const Word &x, const Word &y, int cb) {
    Word s;
    Set cin(KW, 0);
    for (int i = 0; i < 32; ++i) {
      if (i < 31)
        for (int k = 0; k < KW; ++k) {
          A[cb + i][k] = x.bit[i][k] ^ cin[k];
          B[cb + i][k] = y.bit[i][k] ^ cin[k];
        }
      for (int k = 0; k < KW; ++k) s.bit[i][k] = x.bit[i][k] ^ y.bit[i][k] ^ cin[k];
      if (i < 31) toggle(cin, cb + i);
    }
    return s;
  }
  void lin(int base, const Word &v) {
    for (int i = 0; i < 32; ++i) {
      A[base + i] = v.bit[i];
      std::fill(B[base + i].begin(), B[base + i].end(), 0);
      toggle(B[base + i], Z_CONST);
    }
  }
};

Csr to_csr(const std::vector<Set> &rows) {
  Csr c;
  c.ptr.resize(K + 1);
  for (int r = 0; r < K; ++r) {
    c.ptr[r] = (uint32_t)c.cols.size();
    for (int k = 0; k < KW; ++k) {
      uint64_t v = rows[r][k];
      while (v) {
        c.cols.push_back((uint32_t)(k * 64 + __builtin_ctzll(v)));
        v &= v - 1;
      }
    }
  }
  c.ptr[K] = (uint32_t)c.cols.size();
  return c;
}

Csr transpose(const Csr &m) {
  Csr t;
  std::vector<uint32_t> cnt(K + 1, 0);
  for (uint32_t c : m.cols) cnt[c + 1]++;
  for (int i = 0; i < K; ++i) cnt[i + 1] += cnt[i];
  t.ptr = cnt;
  t.cols.resize(m.cols.size());
  std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
  for (int r = 0; r < K; ++r)
    for (uint32_t k = m.ptr[r]; k < m.ptr[r + 1]; ++k) t.cols[fill[m.cols[k]]++] = (uint32_t)r;
  return t;
}

Matrices build() {
  Builder b;
  toggle(b.A[Z_CONST], Z_CONST);
  toggle(b.B[Z_CONST], Z_CONST);
  const int base[6] = {CV_BASE, M_BASE, T_LO_BASE, T_HI_BASE, BLEN_BASE, FLAGS_BASE};
  const int len[6] = {256, 512, 32, 32, 32, 32};
  for (int r = 0; r < 6; ++r)
    for (int j = 0; j < len[r]; ++j) {
      toggle(b.A[base[r] + j], base[r] + j);
      toggle(b.B[base[r] + j], Z_CONST);
    }
  std::vector<Word> st(16);
  for (int w = 0; w < 8; ++w) st[w] = slot(CV_BASE + 32 * w);
  for (int i = 0; i < 4; ++i) st[8 + i] = constant(hb3::IV[i]);
  st[12] = slot(T_LO_BASE);
  st[13] = slot(T_HI_BASE);
  st[14] = slot(BLEN_BASE);
  st[15] = slot(FLAGS_BASE);
  auto carry = [](int g, int add) { return GS_BASE + G_STRIDE * g + 31 * add; };
  auto linb = [](int g, int which) { return GS_BASE + G_STRIDE * g + 186 + 32 * which; };
  for (int r = 0; r < 7; ++r)
    for (int gi = 0; gi < 8; ++gi) {
      int g = 8 * r + gi;
      const int *L = G_LANES[gi];
      Word a = st[L[0]], bb = st[L[1]], c = st[L[2]], d = st[L[3]];
      Word mx = slot(M_BASE + 32 * msg_index(r, gi, 0)), my = slot(M_BASE + 32 * msg_index(r, gi, 1));
      Word t0 = b.add(a, bb, carry(g, 0));
      Word a1 = b.add(t0, mx, carry(g, 1));
      Word d1 = wrotr(wxor(d, a1), 16);
      Word c1 = b.add(c, d1, carry(g, 2));
      Word b1 = wrotr(wxor(bb, c1), 12);
      Word t1 = b.add(a1, b1, carry(g, 3));
      Word a2 = b.add(t1, my, carry(g, 4));
      Word d2 = wrotr(wxor(d1, a2), 8);
      Word c2 = b.add(c1, d2, carry(g, 5));
      Word bn = wrotr(wxor(b1, c2), 7);
      b.lin(linb(g, 0), bn);
      b.lin(linb(g, 1), d2);
      st[L[0]] = a2;
      st[L[1]] = slot(linb(g, 0));
      st[L[2]] = c2;
      st[L[3]] = slot(linb(g, 1));
    }
  for (int w = 0; w < 8; ++w) {
    b.lin(OUT_LO_BASE + 32 * w, wxor(st[w], st[w + 8]));
    b.lin(OUT_HI_BASE + 32 * w, wxor(st[w + 8], slot(CV_BASE + 32 * w)));
  }
  Matrices m;
  m.a = to_csr(b.A);
  m.b = to_csr(b.B);
  m.a_t = transpose(m.a);
  m.b_t = transpose(m.b);
  return m;
}
}  // namespace

const Matrices &matrices() {
  static Matrices m;
  static std::once_flag once;
  std::call_once(once, [] { m = build(); });
  return m;
}

void statement_digest(int m, uint8_t out[32]) {
  // One digest per m, kept for the process (several sizes may be proved
  // concurrently by a server); each is folded from the spine (stmt_digest.h).
  static std::mutex mu;
  static std::map<int, std::array<uint8_t, 32>> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(m);
    if (it != cache.end()) {
      memcpy(out, it->second.data(), 32);
      return;
    }
  }
  digest_from_spine(spine(), m, out);
  std::array<uint8_t, 32> d;
  memcpy(d.data(), out, 32);
  std::lock_guard<std::mutex> lk(mu);
  cache[m] = d;
}

void precompute_digests(int m_lo, int m_hi) {
  uint8_t d[32];
  for (int m = m_lo; m <= m_hi; ++m) statement_digest(m, d);
}

}  // namespace r1cs
}  // namespace flock
