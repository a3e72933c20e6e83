// CPU checks of two pieces of host logic that have no GPU dependency:
//  - flock::bounded_wait / wait_flag (flock-challenge_amd/csrc/host_wait.h):
//    a flag the host never sets ends in TimedOut after about the deadline;
//    a flag another thread sets ends in Done; a probe reporting failure ends
//    in Failed at once; a probe reporting "finished" without the flag ends in
//    Failed (the stream drained without publishing).
//  - flock::Admission (flock-challenge_amd/csrc/admission.h): the proof
//    server's admission by HBM footprint, with uneven reservations that used
//    to stall the queue.
// Prints "ok" or the first failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>
#include <vector>

#include "admission.h"
#include "host_wait.h"

using namespace flock;

#define CHECK(c)                                             \
  do {                                                       \
    if (!(c)) {                                              \
      std::printf("FAIL line %d: %s\n", __LINE__, #c);       \
      return 1;                                              \
    }                                                        \
  } while (0)

static double secs_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

static int waits() {
  alignas(64) uint32_t flag = 0;
  // never set: the deadline ends the wait
  auto t0 = std::chrono::steady_clock::now();
  WaitEnd e = wait_flag(&flag, 7, [] { return 0; }, 0.05);
  double el = secs_since(t0);
  CHECK(e == WaitEnd::TimedOut);
  CHECK(el >= 0.05 && el < 2.0);
  // set by another thread after 10 ms
  std::thread setter([&] {
    std::this_thread::sleep_for(std::chrono::milliseconds(10));
    __atomic_store_n(&flag, 7u, __ATOMIC_RELEASE);
  });
  t0 = std::chrono::steady_clock::now();
  e = wait_flag(&flag, 7, [] { return 0; }, 5.0);
  el = secs_since(t0);
  setter.join();
  CHECK(e == WaitEnd::Done);
  CHECK(el < 2.0);
  // already set: done without probing
  int probes = 0;
  e = wait_flag(&flag, 7, [&] { return ++probes, 0; }, 0.01);
  CHECK(e == WaitEnd::Done && probes == 0);
  // a failing probe ends the wait at once
  t0 = std::chrono::steady_clock::now();
  e = wait_flag(&flag, 9, [] { return -1; }, 10.0);
  CHECK(e == WaitEnd::Failed && secs_since(t0) < 1.0);
  // "finished" without the flag: failed, not done
  e = wait_flag(&flag, 9, [] { return 1; }, 10.0);
  CHECK(e == WaitEnd::Failed);
  // no limit (timeout <= 0): still ends when the flag appears
  __atomic_store_n(&flag, 0u, __ATOMIC_RELEASE);
  std::thread setter2([&] {
    std::this_thread::sleep_for(std::chrono::milliseconds(5));
    __atomic_store_n(&flag, 3u, __ATOMIC_RELEASE);
  });
  e = wait_flag(&flag, 3, [] { return 0; }, 0.0);
  setter2.join();
  CHECK(e == WaitEnd::Done);
  return 0;
}

static int admission() {
  const uint64_t GiB = 1ull << 30;
  const uint64_t r18 = 8 * GiB + GiB / 4, r19 = 16 * GiB + GiB / 4;
  // Budget 20 GiB; two idle workers have each served a 2^18 proof (8.25 GiB
  // reserved); the head request needs 16.25 GiB.  Neither can take it as is
  // (16.25 + 8.25 > 20), so each would release; once one has, the other
  // takes it.
  {
    Admission a(20 * GiB, 2);
    a.take(0, r18);
    a.finish(0);
    a.take(1, r18);
    a.finish(1);
    CHECK(a.ever_fits(r19) && !a.ever_fits(21 * GiB));
    CHECK(a.decide(0, r19) == Admission::Release);
    CHECK(a.decide(1, r19) == Admission::Release);
    a.release(0);
    CHECK(a.decide(0, r19) == Admission::Wait);
    CHECK(a.decide(1, r19) == Admission::Take);
    a.take(1, r19);
    CHECK(a.reserved(1) == r19 && a.busy(1));
  }
  // A busy worker holds 8.25 GiB: the idle one releases, then waits; once
  // the busy one finishes it takes the request.
  {
    Admission a(20 * GiB, 2);
    a.take(0, r18);
    a.take(1, r18);
    a.finish(1);
    CHECK(a.decide(1, r19) == Admission::Release);
    a.release(1);
    CHECK(a.decide(1, r19) == Admission::Wait);
    a.finish(0);
    CHECK(a.decide(0, r19) == Admission::Take);
  }
  // An idle worker that can take the request exists: the other keeps its
  // buffers and waits.
  {
    Admission a(20 * GiB, 2);
    a.reserve(0, r19);
    a.reserve(1, 2 * GiB);
    CHECK(a.decide(1, r19) == Admission::Wait);
    CHECK(a.decide(0, r19) == Admission::Take);
  }
  // Random schedules: whenever the queue is not empty and every worker is
  // idle, some worker acts (takes or releases), so the queue always drains.
  uint64_t st = 0x1234567;
  auto rnd = [&] {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return st;
  };
  for (int trial = 0; trial < 200; ++trial) {
    const int W = 1 + (int)(rnd() % 5);
    const uint64_t budget = (4 + rnd() % 60) * GiB;
    Admission a(budget, W);
    std::vector<uint64_t> q;
    for (int i = 0; i < 40; ++i) q.push_back(GiB / 2 + rnd() % budget);
    for (auto &x : q) x = x > budget ? budget : x;
    size_t head = 0;
    int steps = 0;
    while (head < q.size()) {
      CHECK(++steps < 100000);
      bool acted = false;
      for (int w = 0; w < W && head < q.size(); ++w) {
        if (a.busy(w)) {
          if (rnd() % 3 == 0) a.finish(w);
          continue;
        }
        const Admission::Action act = a.decide(w, q[head]);
        if (act == Admission::Take) {
          a.take(w, q[head++]);
          acted = true;
        } else if (act == Admission::Release) {
          a.release(w);
          acted = true;
        }
      }
      bool all_idle = true;
      for (int w = 0; w < W; ++w) all_idle = all_idle && !a.busy(w);
      uint64_t total = 0;
      for (int w = 0; w < W; ++w) total += a.reserved(w);
      CHECK(total <= budget);
      if (all_idle && head < q.size() && !acted) {
        // a full pass over idle workers must have done something
        bool any = false;
        for (int w = 0; w < W; ++w) any = any || a.decide(w, q[head]) != Admission::Wait;
        CHECK(any);
      }
    }
  }
  return 0;
}

int main() {
  if (int r = waits()) return r;
  if (int r = admission()) return r;
  std::printf("ok\n");
  return 0;
}
