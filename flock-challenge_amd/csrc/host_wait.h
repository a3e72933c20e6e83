// Bounded host waits.  Every place the host waits for the device (a sequence
// number a kernel publishes in mapped memory, an event, a whole stream)
// goes through these loops, which give up after a wall-clock deadline
// instead of spinning forever on work that never finishes.  No HIP types:
// the device side is reached through the `probe` callback, so the loops
// also run (and are tested) on a host without a GPU.
#pragma once
#include <stdint.h>

#include <chrono>
#include <thread>

namespace flock {

enum class WaitEnd { Done, Failed, TimedOut };

// Spin until `done()` holds.  Every `probe_every` polls, `probe()` is asked
// about the device: 0 = keep waiting, > 0 = finished (then done() is
// checked once more), < 0 = failed.  After `timeout_s` seconds of wall time
// the wait ends with TimedOut (timeout_s <= 0: no limit).  After the first
// millisecond the loop yields between polls, so a long wait does not hold a
// core the other provers' threads need.
template <class Done, class Probe>
WaitEnd bounded_wait(Done &&done, Probe &&probe, double timeout_s, uint32_t probe_every = 4096) {
  using clock = std::chrono::steady_clock;
  const clock::time_point t0 = clock::now();
  bool yielding = false;
  for (uint32_t spins = 1;; ++spins) {
    if (done()) return WaitEnd::Done;
    if (spins % probe_every == 0) {
      const int p = probe();
      if (p < 0) return WaitEnd::Failed;
      if (p > 0) return done() ? WaitEnd::Done : WaitEnd::Failed;
      const double el = std::chrono::duration<double>(clock::now() - t0).count();
      if (timeout_s > 0 && el > timeout_s) return WaitEnd::TimedOut;
      if (el > 1e-3) yielding = true;
    }
    if (yielding)
      std::this_thread::yield();
    else
      __builtin_ia32_pause();
  }
}

// Poll a 32-bit flag in host memory until it holds `seq`.
template <class Probe>
WaitEnd wait_flag(const uint32_t *flag, uint32_t seq, Probe &&probe, double timeout_s) {
  return bounded_wait([&] { return __atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq; }, probe, timeout_s);
}

}  // namespace flock
