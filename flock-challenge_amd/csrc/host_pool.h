// A small pool of host threads for the prover's host-side loops (the
// lincheck sumcheck rounds): parallel_for splits [0, n) into chunks that the
// workers and the calling thread claim from one counter.  After a job a
// worker polls for the next one for a short window (the rounds' jobs come a
// few microseconds apart, far less than a sleeping thread takes to wake),
// then sleeps on a condition variable.  One pool per context; calls on one
// pool must not overlap (a context's calls never do).
#pragma once
#include <stddef.h>

#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace flock {

class HostPool {
 public:
  explicit HostPool(int workers);
  ~HostPool();
  int size() const { return (int)threads_.size() + 1; }  // workers + the caller
  // fn(chunk, begin, end) for the chunks [c * grain, min((c + 1) * grain, n));
  // returns when every chunk is done.
  void parallel_for(size_t n, size_t grain, const std::function<void(size_t, size_t, size_t)> &fn);

 private:
  struct Job {
    const std::function<void(size_t, size_t, size_t)> *fn;
    size_t n, grain, chunks;
    std::atomic<size_t> next{0}, done{0};
    std::atomic<int> users{0};
  };
  void worker();
  static void run_chunks(Job *j);
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable cv_;
  std::atomic<uint64_t> gen_{0};  // changed under mu_
  std::atomic<int> sleepers_{0};  // workers waiting on cv_ (changed under mu_)
  std::atomic<bool> stop_{false};
  Job *job_ = nullptr;  // guarded by mu_
};

}  // namespace flock
