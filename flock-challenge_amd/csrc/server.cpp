// Proof server: continuous batching of independent BLAKE3 proofs on one GPU.
//
// Requests (compressions, or a (log2, seed) pair expanded on the device by
// generate_compressions, benchmark-tools/common/src/lib.rs:20-33) enter a FIFO
// queue.  A pool of worker threads, each with its own context (device
// buffers, HIP stream), admits the request at the head of the queue when a
// worker is free and the HBM footprint of all admitted work stays within the
// budget.  Proofs in different streams overlap on the device: one proof's
// kernels run while another waits on its Fiat-Shamir round trip, and small
// late-round kernels of one proof share the chip with large kernels of
// another.  Each proof is computed by the same code as a solo prove, on
// private buffers, so its bytes do not depend on what else is running.
#include <hip/hip_runtime.h>
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flock_gpu.h"
#include "admission.h"
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

// Device bytes a proof of 2^log2 compressions keeps in its context: at most
// 16 packed bit-vectors of 2^(m-3) bytes (witness 3, codeword 2, zerocheck
// round-1 sums and tables 4, ring-switch and Ligerito tables 3.5, with room
// to spare) plus fixed tables.
uint64_t prove_footprint_bytes(int n_log) {
  const int m = r1cs::K_LOG + n_log;
  return 16ull * (1ull << (m - 3)) + (256ull << 20);
}

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

class Server {
 public:
  Server(int device, int n_workers, uint64_t budget, int reserve_log2)
      : device_(device), reserve_log2_(reserve_log2), adm_(budget, n_workers), live_(n_workers) {
    for (int i = 0; i < n_workers; ++i) workers_.emplace_back([this, i] { worker_main(i); });
  }
  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
  }

  int submit(Request &&r, uint64_t *id, std::string &why) {
    std::lock_guard<std::mutex> lk(mu_);
    if (!adm_.ever_fits(prove_footprint_bytes(r.log2))) {
      why = "request exceeds the HBM budget";
      return FLOCK_E_OOM;
    }
    if (live_ == 0) {
      why = "no worker left: every worker's context timed out";
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
    return queue_.size() + running_;
  }

  void set_wait_timeout(double s) {
    std::lock_guard<std::mutex> lk(mu_);
    wait_timeout_s_ = s;
  }

 private:
  // Fail a request without running it (under mu_).
  void fail_locked(Request &&r, int rc, const std::string &why) {
    Result res;
    res.id = r.id;
    res.rc = rc;
    res.error = why;
    res.t_submit = res.t_start = r.t_submit;
    res.t_done = now_s();
    done_.push_back(std::move(res));
  }

  void worker_main(int w) {
    Ctx ctx;
    ctx.device = device_;
    hipSetDevice(device_);
    const bool stream_ok = hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking) == hipSuccess;
    {
      std::lock_guard<std::mutex> lk(mu_);
      ctx.wait_timeout_s = wait_timeout_s_;
    }
    if (stream_ok && reserve_log2_ > 0) {
      // Size every device buffer once for the largest request (one proof of
      // generated compressions), so serving never reallocates.
      Request warm;
      warm.log2 = reserve_log2_;
      warm.n = 1ull << reserve_log2_;
      warm.seeded = true;
      {
        std::lock_guard<std::mutex> lk(mu_);
        adm_.reserve(w, prove_footprint_bytes(warm.log2));
      }
      std::vector<uint8_t> scratch_proof;
      run(ctx, warm, scratch_proof);
    }
    for (;;) {
      Request req;
      {
        std::unique_lock<std::mutex> lk(mu_);
        if (ctx.timed_out) {
          // This worker's context may still have work on the device: it
          // takes no more requests.  With no worker left, the queue can
          // never drain, so its requests fail now.
          --live_;
          if (live_ == 0)
            while (!queue_.empty()) {
              fail_locked(std::move(queue_.front()), FLOCK_E_DEVICE, "no worker left: every worker's context timed out");
              queue_.pop_front();
            }
          done_cv_.notify_all();
          cv_.notify_all();
          return;
        }
        Admission::Action act = Admission::Wait;
        cv_.wait(lk, [&] {
          if (stop_) return true;
          if (queue_.empty()) return false;
          act = adm_.decide(w, prove_footprint_bytes(queue_.front().log2));
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
        req = std::move(queue_.front());
        queue_.pop_front();
        adm_.take(w, prove_footprint_bytes(req.log2));
        ctx.wait_timeout_s = wait_timeout_s_;
        ++running_;
      }
      Result res;
      res.id = req.id;
      res.t_submit = req.t_submit;
      res.t_start = now_s();
      res.rc = stream_ok ? run(ctx, req, res.proof) : FLOCK_E_DEVICE;
      if (res.rc) {
        if (ctx.timed_out) res.rc = FLOCK_E_DEVICE;
        res.error = ctx.last_error;
      }
      res.t_done = now_s();
      {
        std::lock_guard<std::mutex> lk(mu_);
        --running_;
        adm_.finish(w);
        done_.push_back(std::move(res));
      }
      done_cv_.notify_all();
      cv_.notify_all();
    }
  }

  int run(Ctx &ctx, const Request &req, std::vector<uint8_t> &proof) {
    hipSetDevice(device_);
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
  std::vector<std::thread> workers_;
  Admission adm_;  // guarded by mu_
  int live_;       // workers whose context has not timed out
  double wait_timeout_s_ = 120.0;
  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<Request> queue_;
  std::deque<Result> done_;
  size_t running_ = 0;
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
  if (reserve_log2 && (uint64_t)n_workers * flock::prove_footprint_bytes(flock::blocks_log2(1ull << reserve_log2)) >
                          hbm_budget_bytes) {
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

extern "C" int flock_server_pending(flock_server *s, uint64_t *n) {
  if (!s || !n) return FLOCK_E_ARG;
  *n = s->s->pending();
  return FLOCK_OK;
}
