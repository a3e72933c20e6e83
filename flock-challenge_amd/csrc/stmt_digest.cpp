// The statement digest through its m-independent spine (stmt_digest.h).
#include "stmt_digest.h"

#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <memory>
#include <mutex>
#include <thread>

#include "host_blake3.h"
#include "r1cs.h"

namespace flock {
namespace r1cs {

namespace {

void put_u64(uint8_t *p, uint64_t v) { memcpy(p, &v, 8); }

// Chaining value of one (non-root) chunk of `len` <= 1024 bytes with chunk
// counter `index`.
void chunk_cv(const uint8_t *p, size_t len, uint64_t index, uint32_t cv[8]) {
  memcpy(cv, hb3::IV, 32);
  const size_t blocks = len ? (len + 63) / 64 : 1;
  for (size_t b = 0; b < blocks; ++b) {
    uint8_t blk[64] = {0};
    const size_t n = std::min<size_t>(64, len - 64 * b);
    memcpy(blk, p + 64 * b, n);
    uint32_t m[16], o[16];
    memcpy(m, blk, 64);
    const uint32_t flags = (b == 0 ? hb3::CHUNK_START : 0) | (b + 1 == blocks ? hb3::CHUNK_END : 0);
    hb3::compress(cv, m, index, (uint32_t)n, flags, o);
    memcpy(cv, o, 32);
  }
}

void parent(const uint32_t l[8], const uint32_t r[8], uint32_t flags, uint32_t out[8]) {
  uint32_t m[16], o[16];
  memcpy(m, l, 32);
  memcpy(m + 8, r, 32);
  hb3::compress(hb3::IV, m, 0, 64, hb3::PARENT | flags, o);
  memcpy(out, o, 32);
}

std::mutex g_mu;
std::unique_ptr<Spine> g_spine;

}  // namespace

uint64_t digest_body_word_count() {
  const Matrices &mx = matrices();
  return (2 + K + mx.a.cols.size()) + (2 + K + mx.b.cols.size()) + (2 + 2 * (uint64_t)K);
}

std::vector<uint64_t> digest_body_words() {
  const Matrices &mx = matrices();
  std::vector<uint64_t> w;
  w.reserve(digest_body_word_count());
  for (const Csr *c : {&mx.a, &mx.b}) {
    w.push_back(K);
    w.push_back(K);
    for (int r = 0; r < K; ++r) {
      w.push_back(c->ptr[r + 1] - c->ptr[r]);
      for (uint32_t k = c->ptr[r]; k < c->ptr[r + 1]; ++k) w.push_back(c->cols[k]);
    }
  }
  w.push_back(K);  // C_0 = I: row i holds column i
  w.push_back(K);
  for (uint64_t i = 0; i < (uint64_t)K; ++i) {
    w.push_back(1);
    w.push_back(i);
  }
  return w;
}

static void digest_header(int m, uint8_t h[DIGEST_HEADER_BYTES]) {
  memcpy(h, "flock-r1cs-stmt-v1", 18);
  put_u64(h + 18, (uint64_t)m);
  put_u64(h + 26, (uint64_t)K_LOG);
  put_u64(h + 34, (uint64_t)K_SKIP);
  h[42] = 0;  // WitnessLayout::RowMajor
}

void digest_chunk0(int m, uint8_t out[1024]) {
  digest_header(m, out);
  // the body's first 981 bytes: the A framing's head, the same for every m
  const Matrices &mx = matrices();
  std::vector<uint64_t> w = {K, K};
  for (int r = 0; r < K && w.size() * 8 < 1024; ++r) {
    w.push_back(mx.a.ptr[r + 1] - mx.a.ptr[r]);
    for (uint32_t k = mx.a.ptr[r]; k < mx.a.ptr[r + 1]; ++k) w.push_back(mx.a.cols[k]);
  }
  memcpy(out + DIGEST_HEADER_BYTES, w.data(), 1024 - DIGEST_HEADER_BYTES);
}

Spine spine_host() {
  const std::vector<uint64_t> w = digest_body_words();
  const uint64_t total = DIGEST_HEADER_BYTES + 8 * w.size();
  Spine sp;
  sp.n_chunks = (total + 1023) / 1024;
  const uint8_t *body = (const uint8_t *)w.data();
  // level 0: chunks 1 .. n-1 (chunk i is body bytes [1024 i - 43, ..))
  std::vector<std::array<uint32_t, 8>> lev(sp.n_chunks);
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::atomic<uint64_t> next{1};
  std::vector<std::thread> ts;
  for (unsigned t = 0; t < nt; ++t)
    ts.emplace_back([&] {
      for (uint64_t i; (i = next.fetch_add(64)) < sp.n_chunks;)
        for (uint64_t j = i; j < std::min<uint64_t>(i + 64, sp.n_chunks); ++j) {
          const uint64_t off = 1024 * j - DIGEST_HEADER_BYTES;
          chunk_cv(body + off, std::min<uint64_t>(1024, total - 1024 * j), j, lev[j].data());
        }
    });
  for (auto &t : ts) t.join();
  // pairwise levels with the odd last node carried up; record node 1
  while (lev.size() >= 2) {
    sp.sib.push_back(lev[1]);
    std::vector<std::array<uint32_t, 8>> up((lev.size() + 1) / 2);
    for (size_t i = 0; i + 1 < lev.size(); i += 2)
      if (i > 0) parent(lev[i].data(), lev[i + 1].data(), 0, up[i / 2].data());
    if (lev.size() & 1) up.back() = lev.back();
    lev.swap(up);
  }
  return sp;
}

void digest_from_spine(const Spine &sp, int m, uint8_t out[32]) {
  uint8_t c0[1024];
  digest_chunk0(m, c0);
  uint32_t cv[8];
  chunk_cv(c0, 1024, 0, cv);
  for (size_t l = 0; l + 1 < sp.sib.size(); ++l) parent(cv, sp.sib[l].data(), 0, cv);
  parent(cv, sp.sib.back().data(), hb3::ROOT, cv);
  memcpy(out, cv, 32);
}

uint64_t matrices_fingerprint() {
  const Matrices &mx = matrices();
  uint64_t h = 0x9E3779B97F4A7C15ull ^ mx.a.cols.size() ^ (mx.b.cols.size() << 32);
  auto mix = [&h](uint64_t v) {
    h ^= v + 0x9E3779B97F4A7C15ull + (h << 6) + (h >> 2);
    h *= 0xBF58476D1CE4E5B9ull;
  };
  for (const Csr *c : {&mx.a, &mx.b}) {
    for (uint32_t p : c->ptr) mix(p);
    // the columns, 4 at a time
    const size_t n = c->cols.size();
    size_t i = 0;
    for (; i + 4 <= n; i += 4)
      mix((uint64_t)c->cols[i] | ((uint64_t)c->cols[i + 1] << 16) ^ ((uint64_t)c->cols[i + 2] << 32) ^
          ((uint64_t)c->cols[i + 3] << 48));
    for (; i < n; ++i) mix(c->cols[i]);
  }
  return h;
}

std::string spine_cache_path() {
  std::string dir;
  if (const char *e = getenv("FLOCK_CACHE_DIR"); e && *e) {
    dir = e;
  } else {
    Dl_info info;
    if (dladdr((void *)&spine_cache_path, &info) && info.dli_fname) {
      std::string lib = info.dli_fname;
      const size_t s = lib.rfind('/');
      dir = (s == std::string::npos ? std::string(".") : lib.substr(0, s)) + "/cache";
    } else {
      dir = "cache";
    }
  }
  char fp[32];
  snprintf(fp, sizeof fp, "%016llx", (unsigned long long)matrices_fingerprint());
  return dir + "/stmt_spine_v1_" + fp + ".txt";
}

static bool load_spine(const std::string &path, Spine &sp) {
  FILE *f = fopen(path.c_str(), "r");
  if (!f) return false;
  unsigned long long n = 0;
  unsigned levels = 0;
  bool ok = fscanf(f, "flock-stmt-spine v1 %llu %u\n", &n, &levels) == 2 && n >= 2 && levels > 0 && levels < 64;
  Spine s;
  s.n_chunks = n;
  for (unsigned l = 0; ok && l < levels; ++l) {
    std::array<uint32_t, 8> cv;
    for (int i = 0; ok && i < 8; ++i) ok = fscanf(f, "%8x", &cv[i]) == 1;
    s.sib.push_back(cv);
  }
  fclose(f);
  if (!ok) return false;
  sp = s;
  return true;
}

static void save_spine(const std::string &path, const Spine &sp) {
  const size_t slash = path.rfind('/');
  if (slash != std::string::npos) mkdir(path.substr(0, slash).c_str(), 0755);
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  FILE *f = fopen(tmp.c_str(), "w");
  if (!f) return;  // the cache is an optimization: a read-only tree just recomputes
  fprintf(f, "flock-stmt-spine v1 %llu %zu\n", (unsigned long long)sp.n_chunks, sp.sib.size());
  for (auto &cv : sp.sib) {
    for (int i = 0; i < 8; ++i) fprintf(f, "%08x", cv[i]);
    fprintf(f, "\n");
  }
  const bool ok = fclose(f) == 0;
  if (ok) rename(tmp.c_str(), path.c_str());
  else unlink(tmp.c_str());
}

// Defined in digest.hip: the spine on the GPU of ctx (false on failure).
bool spine_device(Ctx *ctx, Spine &out);

const Spine &spine(Ctx *ctx, std::string *from) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_spine) {
    if (from) *from = "memory";
    return *g_spine;
  }
  std::unique_ptr<Spine> sp(new Spine());
  const std::string path = spine_cache_path();
  std::string src = "file";
  if (!load_spine(path, *sp)) {
    src = "gpu";
    if (!ctx || !spine_device(ctx, *sp) || !sp->valid()) {
      src = "host";
      *sp = spine_host();
    }
    save_spine(path, *sp);
  }
  if (from) *from = src;
  g_spine = std::move(sp);
  return *g_spine;
}

}  // namespace r1cs
}  // namespace flock
