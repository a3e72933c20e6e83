// Host-side algebra shared by the prover and the verifier.
#include "host_algebra.h"

#include <algorithm>
#include <map>
#include <mutex>

namespace flock {

namespace {
const F128 ONE = F128::one();
}  // namespace

void build_eq(const F128 *r, int n, F128 *t) {
  t[0] = ONE;
  for (int i = 0; i < n; ++i) {
    size_t half = (size_t)1 << i;
    for (size_t x = 0; x < half; ++x) {
      F128 v = t[x];
      F128 vr = v * r[i];
      t[x + half] = vr;
      t[x] = v + vr;
    }
  }
}

std::vector<F128> eq_table(const F128 *r, int n) {
  std::vector<F128> t((size_t)1 << n);
  build_eq(r, n, t.data());
  return t;
}

// Lagrange weights at z over nodes phi8(first .. first + count) within the
// node set phi8(set_first .. set_first + set_count)
// (zerocheck/multilinear.rs:370-389, 509-563): w_i = prod_{j != i} (z + s_j)
// * inv_den_i over the whole set, with the denominators cached per set and
// the numerators from prefix / suffix products.
void lagrange(F128 z, int first, int count, int set_first, int set_count, F128 *w) {
  static std::mutex mu;
  static std::map<int, std::vector<F128>> inv_den;  // keyed by (set_first, set_count)
  std::vector<F128> idn;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = inv_den.find(set_first * 1000 + set_count);
    if (it == inv_den.end()) {
      std::vector<F128> v(set_count);
      for (int i = 0; i < set_count; ++i) {
        F128 si = phi8((uint8_t)(set_first + i)), den = ONE;
        for (int j = 0; j < set_count; ++j)
          if (j != i) den = den * (si + phi8((uint8_t)(set_first + j)));
        v[i] = inv(den);
      }
      it = inv_den.emplace(set_first * 1000 + set_count, v).first;
    }
    idn = it->second;
  }
  std::vector<F128> pre(set_count + 1), suf(set_count + 1);
  pre[0] = ONE;
  for (int j = 0; j < set_count; ++j) pre[j + 1] = pre[j] * (z + phi8((uint8_t)(set_first + j)));
  suf[set_count] = ONE;
  for (int j = set_count - 1; j >= 0; --j) suf[j] = suf[j + 1] * (z + phi8((uint8_t)(set_first + j)));
  for (int i = 0; i < count; ++i) {
    const int li = first + i - set_first;
    w[i] = pre[li] * suf[li + 1] * idn[li];
  }
}

F128 dot(const F128 *a, const F128 *b, size_t n) {
  F128 s;
  for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

uint8_t f8_inv(uint8_t a) {
  uint8_t r = 1, sq = a;
  for (int i = 0; i < 8; ++i) {
    if ((0xFE >> i) & 1) r = f8_mul(r, sq);
    sq = f8_mul(sq, sq);
  }
  return r;
}

F128 medium(int i) {
  F128 g(1ull << (1u << i), 0);
  return g * inv(ONE + g);
}
extern const uint8_t SMALL_CHAL[3] = {0xF7, 0x53, 0xB5};


int ceil_log2(size_t n) {
  int l = 0;
  if (n <= 1) return 0;
  --n;
  while (n) {
    ++l;
    n >>= 1;
  }
  return l;
}

std::vector<size_t> sample_queries(Transcript &t, size_t block_len, size_t count) {
  std::vector<size_t> out;
  while (out.size() < count) {
    F128 v = t.sample();
    size_t q = (size_t)(v.lo % block_len);
    if (std::find(out.begin(), out.end(), q) == out.end()) out.push_back(q);
  }
  std::sort(out.begin(), out.end());
  return out;
}


}  // namespace flock
