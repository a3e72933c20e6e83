// Batched proving: many independent proofs of uneven sizes, the fronts
// (witness and commitment) of the small ones fused across instances.
//
// The front is the part of a proof that needs no challenge: the witness of
// every compression, the interleaved NTT of the rate-1/2 codeword and its
// Merkle tree.  For small instances (at most 2^FUSE_MAX_LOG2 compressions,
// whose launches are short enough for launch overhead to matter) it runs
// as one launch per step for all of them at once: one witness launch, one launch per NTT pass stage and
// pass variant (pass p of every instance that has one; instances of
// different sizes have different pass lists), one Merkle-leaf launch, one
// launch per Merkle subtree stage, one root gather.  Each launch reads a job
// table (one entry per instance: buffers, shape, first workgroup), so a
// workgroup finds its instance with a scan of at most n entries.  The launch
// sequence depends only on the batch's shape (the instances' sizes), so it
// is captured once per shape in a HIP graph and replayed: later batches of
// the same shape rewrite the job tables and launch the graph.
//
// Once the fused fronts are done, worker threads, each with its own context
// and stream, take the larger instances as whole proofs (largest first) and
// the rest of the small instances' proofs (zerocheck onwards, which waits on
// Fiat-Shamir challenges), one worker at a time on the small ones while
// large ones remain.  The proofs overlap on the device as in the proof
// server.  Every proof is the same bytes as a solo proof of its input.
#include "batch.h"

#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "commit.h"
#include "prove_kernels.h"
#include "prover.h"
#include "r1cs.h"

namespace flock {

namespace {

constexpr size_t JOB_AREA_BYTES = 4u << 20;  // fixed, so captured graphs keep pointing at it

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

using Inst = FrontJob;

// The launch plan of a batch: job tables (host copies, concatenated into one
// upload) and the launches that read them.
struct Launch {
  enum Kind { WITNESS, NTT, LEAF, SUBTREE, ROOTS } kind;
  int F = 0;  // NTT: the pass's layer count and kernel
  NttKernel kernel = NTT_SHARED_W8G2;
  size_t offset = 0;  // byte offset of its jobs in the job area
  uint32_t n_jobs = 0;
  uint64_t wgs = 0;
};

struct Plan {
  std::vector<uint8_t> jobs;  // job area image
  std::vector<Launch> launches;
  size_t roots_offset = 0;  // where the gathered roots land in the job area
  std::vector<uint64_t> shape;
  template <class J>
  size_t add(const std::vector<J> &v) {
    const size_t at = align256(jobs.size());
    jobs.resize(at + v.size() * sizeof(J));
    memcpy(jobs.data() + at, v.data(), v.size() * sizeof(J));
    return at;
  }
};

}  // namespace

struct FrontGraphs {
  std::map<std::vector<uint64_t>, hipGraphExec_t> graphs;
  ~FrontGraphs() {
    for (auto &kv : graphs) hipGraphExecDestroy(kv.second);
  }
};

namespace {

// The fused fronts on one context: its stream and a fixed job area.
struct FrontRunner {
  Ctx *ctx;
  FrontGraphs *cache;  // null: launch directly, no graph

  // The plan for the instances (whose buffers are already assigned).
  Plan plan(std::vector<Inst> &in) {
    Plan p;
    for (auto &x : in) p.shape.push_back(((uint64_t)x.n_log << 56) | x.n_real);
    std::vector<WitnessJob> wj;
    uint64_t wg = 0;
    for (auto &x : in) {
      const uint64_t n_total = 1ull << x.n_log;
      wj.push_back({x.blocks, x.n_real, n_total, (uint64_t *)x.fb.z, (uint64_t *)x.fb.a, (uint64_t *)x.fb.b, wg});
      wg += (n_total + 63) / 64;
    }
    p.launches.push_back({Launch::WITNESS, 0, NTT_SHARED_W8G2, p.add(wj), (uint32_t)wj.size(), wg});
    // NTT pass stages: pass s of every instance (its plan's), one launch per
    // (F, kernel); the first pass of each reads the message (positions
    // masked to the un-replicated half), as commit_device does
    std::vector<std::vector<std::pair<int, NttJob>>> stages;  // (F * NTT_KERNELS + kernel, job)
    for (auto &x : in) {
      const uint4 *tw = (const uint4 *)ctx->device_twiddles(x.k);
      const std::vector<NttPass> plan = ntt_plan(x.k, 1);
      for (size_t s = 0; s < plan.size(); ++s) {
        const NttPass &q = plan[s];
        NttJob j{(uint4 *)x.fb.codeword, tw, s == 0 ? (const uint4 *)x.fb.z : nullptr,
                 s == 0 ? (1ull << (x.k - 1)) - 1 : ~0ull, x.k, q.l,
                 ntt_pass_wgs(x.k, q.F, q.kernel)};  // wg_begin holds the count for now
        if (stages.size() <= s) stages.emplace_back();
        stages[s].push_back({q.F * NTT_KERNELS + (int)q.kernel, j});
      }
    }
    for (auto &st : stages)
      for (int key = 0; key < 5 * NTT_KERNELS; ++key) {
        std::vector<NttJob> jobs;
        uint64_t w = 0;
        for (auto &pj : st)
          if (pj.first == key) {
            NttJob j = pj.second;
            const uint64_t cnt = j.wg_begin;
            j.wg_begin = w;
            w += cnt;
            jobs.push_back(j);
          }
        if (!jobs.empty())
          p.launches.push_back({Launch::NTT, key / NTT_KERNELS, (NttKernel)(key % NTT_KERNELS), p.add(jobs),
                                (uint32_t)jobs.size(), w});
      }
    // Merkle leaves, then subtree stages of up to 9 levels
    std::vector<MerkleJob> lj;
    wg = 0;
    for (auto &x : in) {
      const uint64_t npos = 1ull << x.k;
      lj.push_back({x.fb.codeword, x.fb.tree, npos, 0, 1024u, 0, wg});
      wg += (npos + 255) / 256;
    }
    p.launches.push_back({Launch::LEAF, 0, NTT_SHARED_W8G2, p.add(lj), (uint32_t)lj.size(), wg});
    std::vector<std::vector<MerkleJob>> sub;
    for (auto &x : in) {
      uint64_t rs = 0, rl = 1ull << x.k;
      for (int q = 0; rl > 1; ++q) {
        int L = 0;
        while (L < 9 && (2ull << L) <= rl) ++L;
        if ((int)sub.size() <= q) sub.emplace_back();
        sub[q].push_back({nullptr, x.fb.tree, rl, rs, 0u, L, rl >> L});
        for (int i = 0; i < L; ++i) {
          rs += rl;
          rl >>= 1;
        }
      }
    }
    for (auto &st : sub) {
      uint64_t w = 0;
      for (auto &j : st) {
        const uint64_t cnt = j.wg_begin;
        j.wg_begin = w;
        w += cnt;
      }
      p.launches.push_back({Launch::SUBTREE, 0, NTT_SHARED_W8G2, p.add(st), (uint32_t)st.size(), w});
    }
    std::vector<MerkleJob> rj;
    for (auto &x : in) rj.push_back({nullptr, x.fb.tree, (2ull << x.k) - 1, 0, 0u, 0, 0});
    p.launches.push_back({Launch::ROOTS, 0, NTT_SHARED_W8G2, p.add(rj), (uint32_t)rj.size(), 0});
    p.roots_offset = align256(p.jobs.size());
    p.jobs.resize(p.roots_offset + 32 * in.size());
    return p;
  }

  int issue(const Plan &p, uint8_t *d_jobs, hipStream_t st) {
    for (const Launch &L : p.launches) {
      uint8_t *j = d_jobs + L.offset;
      int rc = FLOCK_OK;
      switch (L.kind) {
        case Launch::WITNESS: rc = witness_batch_device(ctx, st, (const WitnessJob *)j, L.n_jobs, L.wgs); break;
        case Launch::NTT: rc = ntt_batch_device(ctx, st, L.F, L.kernel, (const NttJob *)j, L.n_jobs, L.wgs); break;
        case Launch::LEAF: rc = merkle_leaf_batch_device(ctx, st, (const MerkleJob *)j, L.n_jobs, L.wgs); break;
        case Launch::SUBTREE: rc = merkle_subtree_batch_device(ctx, st, (const MerkleJob *)j, L.n_jobs, L.wgs); break;
        case Launch::ROOTS:
          rc = gather_roots_device(ctx, st, (const MerkleJob *)j, L.n_jobs, d_jobs + p.roots_offset);
          break;
      }
      if (rc) return rc;
    }
    return FLOCK_OK;
  }

  // The fused fronts of every instance; fills each instance's root.
  int front(std::vector<Inst> &in, bool use_graph, bool *replayed) {
    hipStream_t st = ctx->stream;
    Plan p = plan(in);
    if (p.jobs.size() > JOB_AREA_BYTES) {
      ctx->last_error = "batch: too many instances for the job area";
      return FLOCK_E_SHAPE;
    }
    uint8_t *d_jobs = (uint8_t *)ctx->get_scratch("batch.jobs", JOB_AREA_BYTES);
    if (!d_jobs) return FLOCK_E_OOM;
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(d_jobs, p.jobs.data(), p.roots_offset, hipMemcpyHostToDevice, st));
    *replayed = false;
    if (!use_graph || !cache) {
      if (int rc = issue(p, d_jobs, st)) return rc;
    } else {
      auto it = cache->graphs.find(p.shape);
      if (it == cache->graphs.end()) {
        // capture the launch sequence of this shape once (phase timing off:
        // its events are not part of the graph)
        const bool timing = ctx->timing;
        ctx->timing = false;
        hipGraph_t g = nullptr;
        FLOCK_HIP_CHECK(ctx, hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));  // other threads keep using their streams
        const int rc = issue(p, d_jobs, st);
        const hipError_t e = hipStreamEndCapture(st, &g);
        ctx->timing = timing;
        if (rc) {
          if (g) hipGraphDestroy(g);
          return rc;
        }
        FLOCK_HIP_CHECK(ctx, e);
        hipGraphExec_t ex = nullptr;
        const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        FLOCK_HIP_CHECK(ctx, ie);
        it = cache->graphs.emplace(p.shape, ex).first;
      } else {
        *replayed = true;
      }
      FLOCK_HIP_CHECK(ctx, hipGraphLaunch(it->second, st));
    }
    std::vector<uint8_t> roots(32 * in.size());
    FLOCK_HIP_CHECK(ctx, hipMemcpyAsync(roots.data(), d_jobs + p.roots_offset, roots.size(), hipMemcpyDeviceToHost, st));
    if (int rc = ctx->wait_stream(st)) return rc;
    for (size_t i = 0; i < in.size(); ++i) memcpy(in[i].root, roots.data() + 32 * i, 32);
    return FLOCK_OK;
  }
};

}  // namespace

int run_fronts(Ctx *ctx, std::vector<FrontJob> &jobs, FrontGraphs *graphs, bool *replayed) {
  for (auto &j : jobs) {
    j.m = r1cs::K_LOG + j.n_log;
    j.k = j.m - 12;
    if (j.m < 22 || j.n_real == 0 || j.n_real > (1ull << j.n_log)) {
      ctx->last_error = "fronts: a job needs 256 .. 2^21 compressions";
      return FLOCK_E_SHAPE;
    }
  }
  bool rep = false;
  FrontRunner r{ctx, graphs};
  const int rc = r.front(jobs, graphs != nullptr, &rep);
  if (replayed) *replayed = rep;
  return rc;
}

struct BatchEngine {
  Ctx *ctx;  // the fronts' context: arena, job area, stream
  std::vector<std::unique_ptr<Ctx>> workers;
  FrontGraphs graphs;
  std::mutex mu;

  explicit BatchEngine(Ctx *c) : ctx(c) {}

  int ensure_workers(int n) {
    while ((int)workers.size() < n) {
      std::unique_ptr<Ctx> w(new Ctx());
      w->device = ctx->device;
      w->wait_timeout_s = ctx->wait_timeout_s;
      // several workers run at once: no graph captures on their streams
      // (a capture on one thread is disturbed by the others' runtime calls)
      w->zc_graph = false;
      if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess) {
        ctx->last_error = "batch: cannot create a worker stream";
        return FLOCK_E_DEVICE;
      }
      workers.push_back(std::move(w));
    }
    return FLOCK_OK;
  }
};

BatchEngine *batch_engine(Ctx *ctx) {
  if (!ctx->batch) ctx->batch = std::shared_ptr<BatchEngine>(new BatchEngine(ctx));
  return ctx->batch.get();
}

int prove_batch(Ctx *ctx, const std::vector<BatchItem> &items, const uint8_t *domain, size_t domain_len,
                int n_workers, bool use_graph, std::vector<std::vector<uint8_t>> &proofs,
                std::vector<double> &done_s, BatchStats *stats) {
  BatchEngine *E = batch_engine(ctx);
  std::lock_guard<std::mutex> lk(E->mu);
  const double t0 = now_s();
  if (items.empty()) return FLOCK_OK;
  if (n_workers < 1 || n_workers > 16) {
    ctx->last_error = "batch: n_workers must be in 1..16";
    return FLOCK_E_ARG;
  }
  if (int rc = E->ensure_workers(n_workers)) return rc;
  hipStream_t st = ctx->stream;
  if (int rc = ctx->wait_stream(st)) return rc;
  // shapes; small items get their front buffers in one arena (seeded inputs,
  // then z, a, b, codeword and tree)
  std::vector<Inst> in(items.size());
  std::vector<size_t> small, large;
  size_t total = 0;
  for (size_t i = 0; i < items.size(); ++i) {
    const BatchItem &it = items[i];
    const uint64_t n = it.blocks ? it.n_blocks : (1ull << it.log2);
    if (n == 0 || n > (1ull << 21)) {
      ctx->last_error = "batch: item " + std::to_string(i) + " must have 1..2^21 compressions";
      return FLOCK_E_SHAPE;
    }
    int n_log = 0;
    while ((1ull << n_log) < std::max<uint64_t>(n, 8)) ++n_log;
    Inst &x = in[i];
    x.n_log = n_log;
    x.m = r1cs::K_LOG + n_log;
    x.k = x.m - 12;
    x.n_real = n;
    x.blocks = it.blocks;
    if (x.m < 22) {
      ctx->last_error = "batch: item " + std::to_string(i) + " is below 256 compressions (no Ligerito configuration)";
      return FLOCK_E_SHAPE;
    }
    if (n_log > FUSE_MAX_LOG2) {
      large.push_back(i);
      continue;
    }
    small.push_back(i);
  }
  // Large items are taken largest first; small ones after their fused
  // fronts.
  std::stable_sort(large.begin(), large.end(), [&](size_t a, size_t b) { return in[a].n_log > in[b].n_log; });
  std::stable_sort(small.begin(), small.end(), [&](size_t a, size_t b) { return in[a].n_log > in[b].n_log; });
  // The small items' regions in the arena, in the order the fronts take
  // them: [compressions when seeded] z a b codeword tree, each 256-aligned.
  struct Region {
    size_t blocks, z, cw, tree, end;
  };
  std::vector<Region> reg(items.size());
  for (size_t i : small) {
    const Inst &x = in[i];
    Region &g = reg[i];
    g.blocks = total;
    if (!items[i].blocks) total = align256(total + x.n_real * sizeof(flock_compression));
    g.z = total;
    total = align256(total + 3 * ((size_t)8 << (x.m - 6)));  // z, a, b
    g.cw = total;
    total = align256(total + ((size_t)1024 << (x.m - 12)));
    g.tree = total;
    total = align256(total + (((size_t)2 << (x.m - 12)) - 1) * 32);
    g.end = total;
  }
  // Two work lists: the large items (whole proofs) and the small items'
  // rests once their fronts are done.  One worker at a time takes small
  // items while large ones remain: their proofs are latency-bound (short
  // launches, a Fiat-Shamir round trip each round) and overlap with the
  // large proofs' device-bound kernels, instead of all running at the end
  // on a device they cannot fill.
  std::mutex qmu;
  std::condition_variable qcv;
  std::deque<size_t> work(large.begin(), large.end()), work_small;
  int running_small = 0;
  bool fronts_done = small.empty();
  proofs.assign(in.size(), {});
  done_s.assign(in.size(), 0.0);
  std::vector<int> rcs(in.size(), FLOCK_OK);
  std::vector<std::string> errs(in.size());
  // first the small items' fronts, fused (a graph per shape), on this
  // thread alone: a graph capture must not race the workers' runtime calls
  bool replayed = false;
  double front_s = 0;
  int front_rc = FLOCK_OK;
  if (!small.empty()) {
    const double tf = now_s();
    std::vector<Inst> fused;
    uint8_t *arena = (uint8_t *)ctx->get_scratch("batch.arena", total);
    if (!arena) front_rc = FLOCK_E_OOM;
    size_t prev_end = 0;
    for (size_t q = 0; q < small.size() && !front_rc; ++q) {
      const size_t i = small[q];
      const Region &g = reg[i];
      // the regions follow each other inside the arena
      if (g.blocks != prev_end || g.end > total) {
        ctx->last_error = "batch: arena layout inconsistent";
        front_rc = FLOCK_E_SHAPE;
        break;
      }
      prev_end = g.end;
      Inst x = in[i];
      if (!items[i].blocks) {
        x.blocks = (const flock_compression *)(arena + g.blocks);
        front_rc = generate_compressions_device(ctx, st, items[i].log2, items[i].seed,
                                                (flock_compression *)(arena + g.blocks));
      }
      const size_t words = (size_t)1 << (x.m - 6);
      x.fb.z = arena + g.z;
      x.fb.a = arena + g.z + 8 * words;
      x.fb.b = arena + g.z + 16 * words;
      x.fb.codeword = arena + g.cw;
      x.fb.tree = arena + g.tree;
      fused.push_back(x);
    }
    if (!front_rc) front_rc = run_fronts(ctx, fused, use_graph ? &E->graphs : nullptr, &replayed);
    front_s = now_s() - tf;
    {
      std::lock_guard<std::mutex> ql(qmu);
      if (!front_rc)
        for (size_t q = 0; q < small.size(); ++q) {
          in[small[q]] = fused[q];
          work_small.push_back(small[q]);
        }
      fronts_done = true;
    }
    qcv.notify_all();
  }
  std::vector<std::thread> ts;
  for (int w = 0; w < n_workers; ++w)
    ts.emplace_back([&, w] {
      Ctx *wc = E->workers[w].get();
      hipSetDevice(wc->device);
      for (;;) {
        size_t i;
        bool is_small = false;
        {
          std::unique_lock<std::mutex> ql(qmu);
          qcv.wait(ql, [&] { return !work.empty() || !work_small.empty() || fronts_done; });
          if (!work_small.empty() && (running_small == 0 || work.empty())) {
            i = work_small.front();
            work_small.pop_front();
            is_small = true;
            ++running_small;
          } else if (!work.empty()) {
            i = work.front();
            work.pop_front();
          } else {
            return;
          }
        }
        const Inst &x = in[i];
        if (x.n_log > FUSE_MAX_LOG2) {
          // a whole proof on the worker's stream
          const flock_compression *blocks = x.blocks;
          if (!blocks) {
            void *d = wc->get_scratch("batch.blocks", x.n_real * sizeof(flock_compression));
            rcs[i] = d ? generate_compressions_device(wc, wc->stream, items[i].log2, items[i].seed,
                                                      (flock_compression *)d)
                       : FLOCK_E_OOM;
            blocks = (const flock_compression *)d;
          }
          if (!rcs[i]) rcs[i] = prove_blake3_device(wc, wc->stream, blocks, x.n_real, domain, domain_len, proofs[i]);
        } else {
          rcs[i] = prove_after_commit_device(wc, wc->stream, x.m, x.fb, x.root, domain, domain_len, proofs[i]);
        }
        if (rcs[i]) errs[i] = wc->last_error;
        done_s[i] = now_s() - t0;
        if (is_small) {
          std::lock_guard<std::mutex> ql(qmu);
          --running_small;
        }
      }
    });
  for (auto &t : ts) t.join();
  if (front_rc) return front_rc;
  for (size_t i = 0; i < in.size(); ++i)
    if (rcs[i]) {
      ctx->last_error = "batch: item " + std::to_string(i) + ": " + errs[i];
      for (auto &w : E->workers)
        if (w->timed_out) ctx->timed_out = true;
      return rcs[i];
    }
  if (stats) {
    stats->front_s = front_s;
    stats->graph_replayed = replayed;
    stats->fused = small.size();
  }
  return FLOCK_OK;
}

}  // namespace flock
