#include "host_pool.h"

#include <chrono>

namespace flock {

HostPool::HostPool(int workers) {
  for (int i = 0; i < workers; ++i) threads_.emplace_back([this] { worker(); });
}

HostPool::~HostPool() {
  {
    std::lock_guard<std::mutex> g(mu_);
    stop_.store(true);
  }
  cv_.notify_all();
  for (std::thread &t : threads_) t.join();
}

void HostPool::run_chunks(Job *j) {
  for (;;) {
    const size_t c = j->next.fetch_add(1, std::memory_order_relaxed);
    if (c >= j->chunks) return;
    const size_t b = c * j->grain, e = b + j->grain < j->n ? b + j->grain : j->n;
    (*j->fn)(c, b, e);
    j->done.fetch_add(1, std::memory_order_release);
  }
}

void HostPool::worker() {
  // how long a worker polls for the next job before it sleeps
  constexpr auto kSpin = std::chrono::microseconds(300);
  uint64_t seen = 0;
  for (;;) {
    const auto t0 = std::chrono::steady_clock::now();
    while (gen_.load(std::memory_order_acquire) == seen && !stop_.load(std::memory_order_relaxed)) {
      if (std::chrono::steady_clock::now() - t0 > kSpin) {
        std::unique_lock<std::mutex> lk(mu_);
        sleepers_.fetch_add(1, std::memory_order_relaxed);
        cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
        sleepers_.fetch_sub(1, std::memory_order_relaxed);
        break;
      }
      __builtin_ia32_pause();
    }
    Job *j;
    {
      std::lock_guard<std::mutex> g(mu_);
      if (stop_.load()) return;
      seen = gen_.load();
      j = job_;
      if (!j) continue;
      // entered under the lock: the caller clears job_ under the same lock
      // before waiting for users to drop to zero
      j->users.fetch_add(1, std::memory_order_relaxed);
    }
    run_chunks(j);
    j->users.fetch_sub(1, std::memory_order_release);
  }
}

void HostPool::parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t, size_t)> &fn) {
  if (n == 0) return;
  if (grain == 0) grain = 1;
  Job j;
  j.fn = &fn;
  j.n = n;
  j.grain = grain;
  j.chunks = (n + grain - 1) / grain;
  if (threads_.empty() || j.chunks == 1) {
    run_chunks(&j);
    return;
  }
  bool wake;
  {
    std::lock_guard<std::mutex> g(mu_);
    job_ = &j;
    gen_.fetch_add(1, std::memory_order_release);
    wake = sleepers_.load(std::memory_order_relaxed) > 0;
  }
  if (wake) cv_.notify_all();
  run_chunks(&j);
  while (j.done.load(std::memory_order_acquire) < j.chunks) __builtin_ia32_pause();
  {
    std::lock_guard<std::mutex> g(mu_);
    job_ = nullptr;
  }
  while (j.users.load(std::memory_order_acquire) > 0) __builtin_ia32_pause();
}

}  // namespace flock
