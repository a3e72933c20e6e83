// Exercises flock::HostPool (flock-challenge_amd/csrc/host_pool.cpp) on the
// CPU: every chunk of every job runs exactly once with the right bounds,
// over many back-to-back jobs (workers still polling), jobs after the
// workers have gone to sleep, single-chunk jobs on the caller, and pools of
// 0..7 workers created and destroyed.  Prints "ok" or the first failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

#include "host_pool.h"

using flock::HostPool;

static int check_job(HostPool &pool, size_t n, size_t grain) {
  const size_t chunks = (n + grain - 1) / grain;
  std::vector<std::atomic<int>> hits(chunks);
  for (auto &h : hits) h.store(0);
  std::atomic<int> bad{0};
  pool.parallel_for(n, grain, [&](size_t c, size_t b, size_t e) {
    if (c >= chunks || b != c * grain || e != (b + grain < n ? b + grain : n)) bad.fetch_add(1);
    else hits[c].fetch_add(1);
  });
  if (bad.load()) return 1;
  for (auto &h : hits)
    if (h.load() != 1) return 2;
  return 0;
}

int main() {
  for (int workers = 0; workers <= 7; ++workers) {
    HostPool pool(workers);
    if (pool.size() != workers + 1) {
      printf("size %d for %d workers\n", pool.size(), workers);
      return 1;
    }
    const size_t shapes[][2] = {{1, 1}, {1000, 1000}, {1000, 64}, {8192, 512}, {8191, 512}, {100000, 7}, {5, 1}};
    for (int rep = 0; rep < 200; ++rep)
      for (auto &sh : shapes)
        if (int rc = check_job(pool, sh[0], sh[1])) {
          printf("workers %d job (%zu, %zu) rep %d failed (%d)\n", workers, sh[0], sh[1], rep, rc);
          return 1;
        }
    // after the workers' polling window they sleep; jobs must still complete
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    for (auto &sh : shapes)
      if (int rc = check_job(pool, sh[0], sh[1])) {
        printf("workers %d job (%zu, %zu) after sleep failed (%d)\n", workers, sh[0], sh[1], rc);
        return 1;
      }
  }
  printf("ok\n");
  return 0;
}
