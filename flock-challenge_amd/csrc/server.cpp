// Proof server: continuous batching of independent BLAKE3 proofs on one GPU.
//
// Requests (compressions, or a (log2, seed) pair expanded on the device by
// generate_compressions, benchmark-tools/common/src/lib.rs:20-33) enter a FIFO
// queue.  A pool of worker threads, each with its own context and HIP
// stream, proves them, admitted FIFO by the HBM footprint of their buffers
// (Admission); one proof's kernels fill the GPU while another waits on its
// Fiat-Shamir round trip.  Small requests (at most 2^FUSE_MAX_LOG2
// compressions, whose short launches make launch overhead matter) are
// batched phase by phase across requests: one thread takes the small
// requests queued at the head, as many as fit the fronts' share of the
// budget, and runs their fronts (input upload or expansion, witness,
// codeword NTT, Merkle tree) fused, one launch per step for all of them
// (run_fronts, batch.cpp); a worker then runs the rest of each.  Larger
// requests go to the workers whole.
// Each proof is computed by the same kernels as a solo prove on private
// buffers, so its bytes do not depend on what else is running.
#include <hip/hip_runtime.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flock_gpu.h"
#include "admission.h"
#include "batch.h"
#include "ctx.h"
#include "ligerito_configs.h"
#include "prover.h"
#include "r1cs.h"
#include "stmt_digest.h"

namespace flock {

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int blocks_log2(uint64_t n) {
  int l = 0;
  while ((1ull << l) < (n < 8 ? 8 : n)) ++l;
  return l;
}

}  // namespace

// Device bytes of one proof: its front (compressions, witness z, a, b, the
// rate-1/2 codeword and its tree: about 5 packed bit-vectors of 2^(m-3)
// bytes), and the rest of the proof in a worker's context (zerocheck
// round-1 sums and tables, ring-switch and Ligerito tables: at most 11
// bit-vectors with room to spare, plus fixed tables).
uint64_t front_footprint_bytes(int n_log) {
  const int m = r1cs::K_LOG + n_log;
  return (1ull << n_log) * sizeof(flock_compression) + 5ull * (1ull << (m - 3)) + (64ull << (m - 12)) + (1u << 20);
}
uint64_t back_footprint_bytes(int n_log) {
  const int m = r1cs::K_LOG + n_log;
  return 11ull * (1ull << (m - 3)) + (256ull << 20);
}
uint64_t prove_footprint_bytes(int n_log) { return front_footprint_bytes(n_log) + back_footprint_bytes(n_log); }

struct Request {
  uint64_t id = 0;
  uint64_t n = 0;
  int log2 = 0;
  bool seeded = false;
  uint64_t seed = 0;
  std::vector<flock_compression> blocks;
  std::vector<uint8_t> domain;
  double t_submit = 0;
};

struct Result {
  uint64_t id = 0;
  int rc = 0;
  std::string error;
  std::vector<uint8_t> proof;
  double t_submit = 0, t_start = 0, t_done = 0;
};

// A request whose front has run: its buffers (one device allocation) and root.
struct Fronted {
  Request req;
  void *mem = nullptr;
  uint64_t bytes = 0;
  FrontJob job;
  double t_start = 0;
  int rc = 0;
  std::string error;
};

class Server {
 public:
  // Of the budget, fronts take FRONT_SHARE (at least one front of the
  // largest supported size when the budget allows), workers the rest.
  static constexpr double FRONT_SHARE = 0.35;
  static constexpr size_t MAX_FUSED = 32;  // requests per fused front

  Server(int device, int n_workers, uint64_t budget, int reserve_log2)
      : device_(device),
        reserve_log2_(reserve_log2),
        front_budget_((uint64_t)(budget * FRONT_SHARE)),
        adm_(budget - (uint64_t)(budget * FRONT_SHARE), n_workers),
        live_(n_workers) {
    front_thread_ = std::thread([this] { front_main(); });
    for (int i = 0; i < n_workers; ++i) workers_.emplace_back([this, i] { worker_main(i); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    front_thread_.join();
    for (auto &t : workers_) t.join();
    for (auto &f : ready_)
      if (f.mem) (void)hipFree(f.mem);
    for (auto &kv : pool_) (void)hipFree(kv.second);
  }

  int submit(Request &&r, uint64_t *id, std::string &why) {
    std::lock_guard<std::mutex> lk(mu_);
    const bool fits = r.log2 <= FUSE_MAX_LOG2
                          ? front_footprint_bytes(r.log2) <= front_budget_ && adm_.ever_fits(back_footprint_bytes(r.log2))
                          : adm_.ever_fits(prove_footprint_bytes(r.log2));
    if (!fits) {
      why = "request exceeds the HBM budget";
      return FLOCK_E_OOM;
    }
    if (live_ == 0 || front_dead_) {
      why = front_dead_ ? "the fronts' context timed out" : "no worker left: every worker's context timed out";
      return FLOCK_E_DEVICE;
    }
    r.id = next_id_++;
    r.t_submit = now_s();
    *id = r.id;
    queue_.push_back(std::move(r));
    cv_.notify_all();
    return FLOCK_OK;
  }

  // 1 when a result was returned, 0 on timeout.
  int poll(int wait_ms, Result &out) {
    std::unique_lock<std::mutex> lk(mu_);
    if (wait_ms != 0) {
      auto pred = [this] { return !done_.empty(); };
      if (wait_ms < 0)
        done_cv_.wait(lk, pred);
      else
        done_cv_.wait_for(lk, std::chrono::milliseconds(wait_ms), pred);
    }
    if (done_.empty()) return 0;
    out = std::move(done_.front());
    done_.pop_front();
    return 1;
  }

  size_t pending() {
    std::lock_guard<std::mutex> lk(mu_);
    return queue_.size() + fronting_ + ready_.size() + running_;
  }

  void set_wait_timeout(double s) {
    std::lock_guard<std::mutex> lk(mu_);
    wait_timeout_s_ = s;
  }

  // Requests whose fronts ran fused with others', and fused front launches.
  void fused_stats(uint64_t *requests, uint64_t *launches) {
    std::lock_guard<std::mutex> lk(mu_);
    *requests = fused_requests_;
    *launches = front_batches_;
  }

 private:
  // Fail a request without running it (under mu_).
  void fail_locked(const Request &r, int rc, const std::string &why, double t_start) {
    Result res;
    res.id = r.id;
    res.rc = rc;
    res.error = why;
    res.t_submit = r.t_submit;
    res.t_start = t_start ? t_start : r.t_submit;
    res.t_done = now_s();
    done_.push_back(std::move(res));
  }

  // ---- stage 1: fused fronts
  void front_main() {
    Ctx ctx;
    ctx.device = device_;
    (void)hipSetDevice(device_);
    if (hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking) != hipSuccess) {
      std::lock_guard<std::mutex> lk(mu_);
      front_dead_ = true;
      fail_all_locked("cannot create the fronts' stream");
      return;
    }
    for (;;) {
      std::vector<Fronted> batch;
      {
        std::unique_lock<std::mutex> lk(mu_);
        auto fusable = [&] {
          return !queue_.empty() && queue_.front().log2 <= FUSE_MAX_LOG2 &&
                 front_live_ + front_footprint_bytes(queue_.front().log2) <= front_budget_;
        };
        cv_.wait(lk, [&] { return stop_ || fusable(); });
        if (stop_) return;
        ctx.wait_timeout_s = wait_timeout_s_;
        const double t = now_s();
        while (batch.size() < MAX_FUSED && fusable()) {
          Fronted f;
          f.req = std::move(queue_.front());
          queue_.pop_front();
          f.bytes = front_footprint_bytes(f.req.log2);
          f.t_start = t;
          front_live_ += f.bytes;
          batch.push_back(std::move(f));
        }
        fronting_ += batch.size();
      }
      run_front_batch(ctx, batch);
      {
        std::lock_guard<std::mutex> lk(mu_);
        fronting_ -= batch.size();
        ++front_batches_;
        if (batch.size() > 1) fused_requests_ += batch.size();
        for (auto &f : batch) {
          if (f.rc) {
            fail_locked(f.req, f.rc, f.error, f.t_start);
            give_front_memory_locked(f.mem, f.bytes);
            front_live_ -= f.bytes;
          } else {
            ready_.push_back(std::move(f));
          }
        }
        if (ctx.timed_out) {
          front_dead_ = true;
          fail_all_locked("the fronts' context timed out");
        }
      }
      cv_.notify_all();
      done_cv_.notify_all();
      if (ctx.timed_out) return;
    }
  }

  // Allocate each request's front buffers, bring its compressions to the
  // device, then run every front fused.
  void run_front_batch(Ctx &ctx, std::vector<Fronted> &batch) {
    std::vector<FrontJob> jobs;
    std::vector<size_t> idx;
    for (size_t i = 0; i < batch.size(); ++i) {
      Fronted &f = batch[i];
      f.mem = take_front_memory(f.bytes);
      if (!f.mem) {
        f.rc = FLOCK_E_OOM;
        f.error = "cannot allocate the front buffers";
        continue;
      }
      const int m = r1cs::K_LOG + f.req.log2;
      uint8_t *p = (uint8_t *)f.mem;
      auto take = [&p](size_t n) {
        uint8_t *q = p;
        p += (n + 255) & ~(size_t)255;
        return q;
      };
      flock_compression *blocks = (flock_compression *)take((1ull << f.req.log2) * sizeof(flock_compression));
      FrontJob j;
      j.blocks = blocks;
      j.n_real = f.req.n;
      j.n_log = f.req.log2;
      const size_t words = (size_t)1 << (m - 6);
      j.fb.z = take(8 * words);
      j.fb.a = take(8 * words);
      j.fb.b = take(8 * words);
      j.fb.codeword = take((size_t)1024 << (m - 12));
      j.fb.tree = take((((size_t)2 << (m - 12)) - 1) * 32);
      int rc = FLOCK_OK;
      if (f.req.seeded) {
        rc = generate_compressions_device(&ctx, ctx.stream, (uint32_t)f.req.log2, f.req.seed, blocks);
      } else if (hipMemcpyAsync(blocks, f.req.blocks.data(), f.req.n * sizeof(flock_compression),
                                hipMemcpyHostToDevice, ctx.stream) != hipSuccess) {
        rc = FLOCK_E_DEVICE;
        ctx.last_error = "cannot upload the compressions";
      }
      if (rc) {
        f.rc = rc;
        f.error = ctx.last_error;
        continue;
      }
      jobs.push_back(j);
      idx.push_back(i);
    }
    if (jobs.empty()) return;
    const int rc = run_fronts(&ctx, jobs, nullptr, nullptr);
    for (size_t q = 0; q < jobs.size(); ++q) {
      Fronted &f = batch[idx[q]];
      f.job = jobs[q];
      if (rc) {
        f.rc = ctx.timed_out ? FLOCK_E_DEVICE : rc;
        f.error = ctx.last_error;
      }
    }
  }

  // Front buffers come from a pool of earlier ones of the same size (a
  // request's front is one allocation whose size depends only on its
  // log2), so steady serving neither allocates nor frees: hipFree waits for
  // the device, which would stall the workers' streams.  Pooled bytes count
  // against the fronts' budget; a new size frees pooled buffers of other
  // sizes when needed.
  void *take_front_memory(uint64_t bytes) {
    std::vector<void *> drop;
    {
      std::lock_guard<std::mutex> lk(mu_);
      auto it = pool_.find(bytes);
      if (it != pool_.end()) {
        void *p = it->second;
        pool_.erase(it);
        pooled_bytes_ -= bytes;
        return p;
      }
      // front_live_ already counts this request
      while (front_live_ + pooled_bytes_ > front_budget_ && !pool_.empty()) {
        auto last = std::prev(pool_.end());
        pooled_bytes_ -= last->first;
        drop.push_back(last->second);
        pool_.erase(last);
      }
    }
    for (void *p : drop) (void)hipFree(p);
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
  }
  void give_front_memory_locked(void *p, uint64_t bytes) {
    if (!p) return;
    pool_.emplace(bytes, p);
    pooled_bytes_ += bytes;
  }

  // Everything queued or fronted fails (under mu_).
  void fail_all_locked(const std::string &why) {
    while (!queue_.empty()) {
      fail_locked(queue_.front(), FLOCK_E_DEVICE, why, 0);
      queue_.pop_front();
    }
    while (!ready_.empty()) {
      fail_locked(ready_.front().req, FLOCK_E_DEVICE, why, ready_.front().t_start);
      give_front_memory_locked(ready_.front().mem, ready_.front().bytes);
      front_live_ -= ready_.front().bytes;
      ready_.pop_front();
    }
  }

  // ---- stage 2: the rest of each proof
  void worker_main(int w) {
    Ctx ctx;
    ctx.device = device_;
    // several workers run at once: no graph captures on their streams (a
    // capture on one thread is disturbed by the others' runtime calls)
    ctx.zc_graph = false;
    (void)hipSetDevice(device_);
    const bool stream_ok = hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking) == hipSuccess;
    {
      std::lock_guard<std::mutex> lk(mu_);
      ctx.wait_timeout_s = wait_timeout_s_;
    }
    if (stream_ok && reserve_log2_ > 0) {
      // Size the context's buffers once for the largest request (one whole
      // proof of generated compressions), so serving never reallocates.
      Request warm;
      warm.log2 = reserve_log2_;
      warm.n = 1ull << reserve_log2_;
      warm.seeded = true;
      {
        std::lock_guard<std::mutex> lk(mu_);
        adm_.reserve(w, (reserve_log2_ > FUSE_MAX_LOG2 ? prove_footprint_bytes : back_footprint_bytes)(reserve_log2_));
      }
      std::vector<uint8_t> scratch_proof;
      run_whole(ctx, warm, scratch_proof);
    }
    for (;;) {
      Fronted f;
      bool whole = false;
      {
        std::unique_lock<std::mutex> lk(mu_);
        if (ctx.timed_out || !stream_ok) {
          // This worker's context may still have work on the device: it
          // takes no more requests.  With no worker left, nothing can drain.
          --live_;
          if (live_ == 0) fail_all_locked("no worker left: every worker's context timed out");
          done_cv_.notify_all();
          cv_.notify_all();
          return;
        }
        // the earliest submitted of: the head of the fronted requests, the
        // head of the queue when it is large (a whole proof here)
        Admission::Action act = Admission::Wait;
        cv_.wait(lk, [&] {
          if (stop_) return true;
          const bool q_large = !queue_.empty() && queue_.front().log2 > FUSE_MAX_LOG2;
          if (ready_.empty() && !q_large) return false;
          whole = q_large && (ready_.empty() || queue_.front().t_submit < ready_.front().req.t_submit);
          act = whole ? adm_.decide(w, prove_footprint_bytes(queue_.front().log2))
                      : adm_.decide(w, back_footprint_bytes(ready_.front().req.log2));
          return act != Admission::Wait;
        });
        if (stop_) return;
        if (act == Admission::Release) {
          // The head request fits no idle worker: give up this worker's
          // buffers so that the reservations shrink until it does.
          adm_.release(w);
          lk.unlock();
          ctx.release_scratch();
          cv_.notify_all();
          continue;
        }
        if (whole) {
          f.req = std::move(queue_.front());
          queue_.pop_front();
          f.t_start = now_s();
          adm_.take(w, prove_footprint_bytes(f.req.log2));
        } else {
          f = std::move(ready_.front());
          ready_.pop_front();
          adm_.take(w, back_footprint_bytes(f.req.log2));
        }
        ctx.wait_timeout_s = wait_timeout_s_;
        ++running_;
      }
      Result res;
      res.id = f.req.id;
      res.t_submit = f.req.t_submit;
      res.t_start = f.t_start;
      (void)hipSetDevice(device_);
      res.rc = whole ? run_whole(ctx, f.req, res.proof)
                     : prove_after_commit_device(&ctx, ctx.stream, f.job.m, f.job.fb, f.job.root,
                                                 f.req.domain.data(), f.req.domain.size(), res.proof);
      if (res.rc) {
        if (ctx.timed_out) res.rc = FLOCK_E_DEVICE;
        res.error = ctx.last_error;
      }
      res.t_done = now_s();
      {
        std::lock_guard<std::mutex> lk(mu_);
        --running_;
        // the worker's stream has drained, so nothing reads the front any
        // more; after a timeout it may, so those buffers are not reused
        if (!ctx.timed_out) give_front_memory_locked(f.mem, f.bytes);
        front_live_ -= f.bytes;
        adm_.finish(w);
        done_.push_back(std::move(res));
      }
      done_cv_.notify_all();
      cv_.notify_all();
    }
  }

  // A whole proof on the worker's context and stream.
  int run_whole(Ctx &ctx, const Request &req, std::vector<uint8_t> &proof) {
    const size_t bytes = req.n * sizeof(flock_compression);
    void *d = ctx.get_scratch("server.blocks", bytes);
    if (!d) return FLOCK_E_OOM;
    int rc = FLOCK_OK;
    if (req.seeded) {
      rc = generate_compressions_device(&ctx, ctx.stream, (uint32_t)req.log2, req.seed, (flock_compression *)d);
    } else {
      FLOCK_HIP_CHECK(&ctx, hipMemcpyAsync(d, req.blocks.data(), bytes, hipMemcpyHostToDevice, ctx.stream));
    }
    if (rc) return rc;
    return prove_blake3_device(&ctx, ctx.stream, (const flock_compression *)d, req.n, req.domain.data(),
                               req.domain.size(), proof);
  }

  int device_;
  int reserve_log2_;
  uint64_t front_budget_;
  std::thread front_thread_;
  std::vector<std::thread> workers_;
  Admission adm_;  // guarded by mu_ (workers' buffers)
  int live_;       // workers whose context has not timed out
  bool front_dead_ = false;
  uint64_t front_live_ = 0;  // bytes of front buffers allocated or about to be
  std::multimap<uint64_t, void *> pool_;  // free front buffers by size
  uint64_t pooled_bytes_ = 0;
  double wait_timeout_s_ = 120.0;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Request> queue_;    // waiting for a front
  std::deque<Fronted> ready_;    // front done, waiting for a worker
  std::deque<Result> done_;
  size_t fronting_ = 0, running_ = 0;
  uint64_t fused_requests_ = 0, front_batches_ = 0;
  uint64_t next_id_ = 1;
  bool stop_ = false;
};

}  // namespace flock

// ------------------------------------------------------------------ C ABI
struct flock_server {
  flock::Server *s;
  std::string last_error;
};

static bool lig_config_exists(int n_log) {
  const int m = flock::r1cs::K_LOG + n_log;
  for (int i = 0; i < FLOCK_LIG_N_CONFIGS; ++i)
    if (FLOCK_LIG_CONFIGS[i].m == m && FLOCK_LIG_CONFIGS[i].profile == 0) return true;
  return false;
}

extern "C" int flock_server_create(int device, int n_workers, uint64_t hbm_budget_bytes, uint32_t reserve_log2,
                                   flock_server **out) {
  if (!out || n_workers < 1 || n_workers > 16 || reserve_log2 > 21) return FLOCK_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return FLOCK_E_DEVICE;
  if (hbm_budget_bytes == 0) {
    size_t free_b = 0, total_b = 0;
    hipSetDevice(device);
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return FLOCK_E_DEVICE;
    hbm_budget_bytes = free_b / 10 * 8;
  }
  flock::r1cs::matrices();  // build once before the workers start
  // statement digests of every supported size (setup work, as in the
  // reference): the spine from the persisted cache, else hashed on the GPU
  {
    flock::Ctx setup_ctx;
    setup_ctx.device = device;
    hipSetDevice(device);
    if (hipStreamCreateWithFlags(&setup_ctx.stream, hipStreamNonBlocking) == hipSuccess)
      flock::r1cs::spine(&setup_ctx);
  }
  flock::r1cs::precompute_digests(flock::r1cs::K_LOG + 8, flock::r1cs::K_LOG + 21);
  flock_server *s = new (std::nothrow) flock_server();
  if (!s) return FLOCK_E_OOM;
  const int rl = flock::blocks_log2(1ull << reserve_log2);
  const uint64_t per_worker =
      rl > flock::FUSE_MAX_LOG2 ? flock::prove_footprint_bytes(rl) : flock::back_footprint_bytes(rl);
  if (reserve_log2 && (uint64_t)n_workers * per_worker >
                          hbm_budget_bytes - (uint64_t)(hbm_budget_bytes * flock::Server::FRONT_SHARE)) {
    delete s;
    return FLOCK_E_OOM;
  }
  s->s = new (std::nothrow) flock::Server(device, n_workers, hbm_budget_bytes,
                                          reserve_log2 ? flock::blocks_log2(1ull << reserve_log2) : 0);
  if (!s->s) {
    delete s;
    return FLOCK_E_OOM;
  }
  *out = s;
  return FLOCK_OK;
}

extern "C" void flock_server_destroy(flock_server *s) {
  if (!s) return;
  delete s->s;
  delete s;
}

extern "C" const char *flock_server_last_error(const flock_server *s) {
  return s ? s->last_error.c_str() : "null server";
}

static int submit_common(flock_server *s, flock::Request &&r, uint64_t *id) {
  if (!lig_config_exists(r.log2)) {
    s->last_error = "no Ligerito configuration for this size (2^8 .. 2^21 compressions)";
    return FLOCK_E_SHAPE;
  }
  std::string why;
  int rc = s->s->submit(std::move(r), id, why);
  if (rc) s->last_error = why;
  return rc;
}

extern "C" int flock_server_submit(flock_server *s, const flock_compression *blocks, uint64_t n_blocks,
                                   const uint8_t *domain, size_t domain_len, uint64_t *request_id) {
  if (!s || !blocks || !request_id || n_blocks == 0 || (!domain && domain_len)) return FLOCK_E_ARG;
  flock::Request r;
  r.n = n_blocks;
  r.log2 = flock::blocks_log2(n_blocks);
  r.blocks.assign(blocks, blocks + n_blocks);
  r.domain.assign(domain, domain + domain_len);
  return submit_common(s, std::move(r), request_id);
}

extern "C" int flock_server_submit_seeded(flock_server *s, uint32_t log2, uint64_t seed, const uint8_t *domain,
                                          size_t domain_len, uint64_t *request_id) {
  if (!s || !request_id || log2 > 21 || (!domain && domain_len)) return FLOCK_E_ARG;
  flock::Request r;
  r.n = 1ull << log2;
  r.log2 = flock::blocks_log2(r.n);
  r.seeded = true;
  r.seed = seed;
  r.domain.assign(domain, domain + domain_len);
  return submit_common(s, std::move(r), request_id);
}

extern "C" int flock_server_poll(flock_server *s, int wait_ms, uint64_t *request_id, uint8_t *out, size_t cap,
                                 size_t *out_len, double *latency_s, double *queue_s) {
  if (!s || !request_id) return FLOCK_E_ARG;
  flock::Result r;
  if (!s->s->poll(wait_ms, r)) return 0;
  *request_id = r.id;
  if (latency_s) *latency_s = r.t_done - r.t_submit;
  if (queue_s) *queue_s = r.t_start - r.t_submit;
  if (out_len) *out_len = r.proof.size();
  if (r.rc) {
    s->last_error = r.error;
    return r.rc;
  }
  if (!out || cap < r.proof.size()) {
    s->last_error = "output buffer too small; the result is dropped";
    return FLOCK_E_CAP;
  }
  memcpy(out, r.proof.data(), r.proof.size());
  return 1;
}

extern "C" int flock_server_set_wait_timeout(flock_server *s, double seconds) {
  if (!s || !(seconds == seconds)) return FLOCK_E_ARG;
  s->s->set_wait_timeout(seconds);
  return FLOCK_OK;
}

extern "C" int flock_server_stats(flock_server *s, uint64_t *fused_requests, uint64_t *front_launches) {
  if (!s || !fused_requests || !front_launches) return FLOCK_E_ARG;
  s->s->fused_stats(fused_requests, front_launches);
  return FLOCK_OK;
}

extern "C" int flock_server_pending(flock_server *s, uint64_t *n) {
  if (!s || !n) return FLOCK_E_ARG;
  *n = s->s->pending();
  return FLOCK_OK;
}
