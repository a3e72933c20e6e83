// Admission by HBM footprint for the proof server (server.cpp).  Pure host
// logic with no HIP types, so it is tested on a host without a GPU.
//
// Each worker keeps the device buffers of the largest proof it has run
// (grow-only, so a worker never reallocates while serving), and the sum of
// those reservations stays within the budget.  A worker may take the head
// request when the sum, with its own reservation raised to the request's
// footprint, still fits.  When the head request fits no idle worker that
// way, idle workers that hold a reservation give it up (free their buffers),
// one after another, until it fits: every request is at most the budget
// (checked at submit), so once the busy workers finish and the idle ones
// release, it always fits, and the FIFO queue cannot stall behind it.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <vector>

namespace flock {

class Admission {
 public:
  enum Action { Wait, Take, Release };

  Admission(uint64_t budget, int workers) : budget_(budget), reserved_(workers, 0), busy_(workers, false) {}

  uint64_t budget() const { return budget_; }
  uint64_t reserved(int w) const { return reserved_[w]; }
  bool busy(int w) const { return busy_[w]; }
  // A request larger than the budget could never be admitted.
  bool ever_fits(uint64_t need) const { return need <= budget_; }

  // Would the reservations fit with worker w's raised to `need`?
  bool fits(int w, uint64_t need) const {
    uint64_t total = 0;
    for (size_t i = 0; i < reserved_.size(); ++i) total += (int)i == w ? std::max(reserved_[i], need) : reserved_[i];
    return total <= budget_;
  }

  // What idle worker w does about a head request of `need` bytes: take it;
  // give up its reservation because no idle worker can take it otherwise;
  // or wait (for another worker to take it, or for busy workers to finish).
  Action decide(int w, uint64_t need) const {
    if (fits(w, need)) return Take;
    if (reserved_[w] == 0) return Wait;
    for (size_t i = 0; i < reserved_.size(); ++i)
      if ((int)i != w && !busy_[i] && fits((int)i, need)) return Wait;
    return Release;
  }

  void take(int w, uint64_t need) {
    reserved_[w] = std::max(reserved_[w], need);
    busy_[w] = true;
  }
  void finish(int w) { busy_[w] = false; }
  void release(int w) { reserved_[w] = 0; }
  // A worker that holds buffers without a request (the start-up reservation).
  void reserve(int w, uint64_t bytes) { reserved_[w] = std::max(reserved_[w], bytes); }

 private:
  uint64_t budget_;
  std::vector<uint64_t> reserved_;
  std::vector<bool> busy_;
};

}  // namespace flock
