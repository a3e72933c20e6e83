// Per-device prover context: stream, error text, cached constant tables and
// a grow-only device arena.  One context per GPU; calls on distinct contexts
// are independent, calls on one context are serialized by the caller.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "host_field.h"
#include "host_pool.h"
#include "host_wait.h"

namespace flock {

struct BatchEngine;

struct DeviceBuffer {
  void *ptr = nullptr;
  size_t bytes = 0;
};

// Device-side phase timing with HIP events on the launching stream.
struct PhaseEvent {
  std::string name;
  hipEvent_t begin, end;
};

struct Ctx {
  int device = 0;
  bool timing = false;
  // When non-empty, only these phases are timed (fewer events on the stream).
  std::set<std::string> timing_only;
  bool timed(const char *name) const { return timing && (timing_only.empty() || timing_only.count(name)); }
  std::vector<hipEvent_t> event_pool;
  hipEvent_t take_event();
  std::vector<PhaseEvent> phase_events;
  std::vector<size_t> open_phases;
  std::map<std::string, std::pair<double, long>> phase_totals;  // ms, count
  void phase_begin(const std::string &name, hipStream_t st);
  void phase_end(hipStream_t st);
  void resolve_phases();  // after the stream has been synchronized
  hipStream_t stream = nullptr;
  std::string last_error;
  // Every host wait for the device gives up after this many seconds
  // (<= 0: no limit).  A wait that gives up marks the context timed out:
  // its work may still be running, so every later call on it fails, and
  // destroying it leaks its device memory instead of waiting on that work.
  double wait_timeout_s = 120.0;
  bool timed_out = false;
  // Run the zerocheck rounds' Fiat-Shamir steps on the device
  // (transcript_dev.h) instead of a host round trip per round; the proof
  // bytes are the same.  Off by default: measured slower at every size.
  bool device_transcript = false;
  // Capture the zerocheck rounds (device transcript only) in a HIP graph per
  // m and buffer set, replayed by later proofs.
  bool zc_graph = true;
  std::map<std::vector<const void *>, hipGraphExec_t> zc_graphs;
  // Bounded waits: until a sequence number appears at `flag` (published by
  // the device in the mapped readback area), until the stream has drained,
  // until the event has completed.  FLOCK_OK, or FLOCK_E_DEVICE with
  // last_error set.
  int wait_flag_seq(const uint32_t *flag, uint32_t seq, hipStream_t st);
  int wait_stream(hipStream_t st);
  int wait_event(hipEvent_t e);
  int wait_failed(WaitEnd how, hipError_t err, const char *what);
  // Breadth-first NTT twiddle table per domain dimension (device copy).
  std::map<int, F128 *> twiddles_dev;
  std::map<int, std::vector<F128>> twiddles_host;
  // Named grow-only scratch buffers.
  std::map<std::string, DeviceBuffer> scratch;
  // The R1CS transposes live in named scratch buffers once uploaded.
  bool csc_uploaded = false;
  // The zerocheck round-1 constant tables live in a named scratch buffer once uploaded.
  bool zc_tables_uploaded = false;
  // Bytes of the last proof made on this context.
  std::vector<uint8_t> last_proof;
  // Pinned host staging: an upload ring (asynchronous host-to-device copies
  // read from it; reset only after the stream has drained) and a landing
  // area for small device-to-host reads.
  uint8_t *pinned = nullptr;
  size_t pinned_bytes = 0, pinned_off = 0;
  uint8_t *landing = nullptr;
  size_t landing_bytes = 0;
  // Pinned host area for Merkle openings copied back without waiting (read
  // only when the proof is serialized); open_area_off is reset per proof.
  uint8_t *open_area = nullptr;
  size_t open_area_bytes = 0, open_area_off = 0;
  // Coherent pinned host memory mapped into the device: a kernel writes a
  // small result (a sumcheck message) straight into it, and the host reads it
  // after synchronizing the stream, with no copy kernel in between.
  uint8_t *readback = nullptr;
  void *readback_dev = nullptr;
  static constexpr size_t READBACK_BYTES = 4096;
  // A sequence number the device publishes in the readback area (at this
  // offset) once the work queued before it has finished: the host polls it
  // instead of synchronizing the stream, which wakes several microseconds later.
  // Kernels that write a result into the readback area publish the number
  // themselves after their writes (the reductions and the grinding take);
  // otherwise a one-lane signal kernel does.
  static constexpr size_t READBACK_FLAG = READBACK_BYTES - 64;
  uint32_t readback_seq = 0;
  // The next sequence number to publish; never 0, which kernels read as
  // "publish nothing".
  uint32_t next_readback_seq() {
    if (++readback_seq == 0) ++readback_seq;
    return readback_seq;
  }
  uint32_t *readback_flag_dev() const { return (uint32_t *)((uint8_t *)readback_dev + READBACK_FLAG); }
  uint8_t *pinned_alloc(size_t n);  // nullptr when the ring is full (or unavailable)
  void pinned_reset() { pinned_off = 0; }

  void *get_scratch(const std::string &name, size_t bytes);
  // Free every named scratch buffer (after the stream has drained); the
  // next prove allocates again.  FLOCK_OK or the wait's error.
  int release_scratch();
  // Free one named scratch buffer (after the stream has drained), if any.
  void free_scratch(const std::string &name);
  const std::vector<F128> &host_twiddles(int dim);
  const F128 *device_twiddles(int dim);
  // Host threads for the prover's host-side loops, started on first use.
  std::unique_ptr<HostPool> pool;
  // Batched proving state (worker contexts, captured graphs), on first use.
  std::shared_ptr<BatchEngine> batch;
  HostPool &host_pool();
  ~Ctx();
};

struct ScopedPhase {
  Ctx *ctx;
  hipStream_t st;
  bool open;
  ScopedPhase(Ctx *c, const char *name, hipStream_t s) : ctx(c), st(s), open(c->timed(name)) {
    if (open) ctx->phase_begin(name, st);
  }
  void end() {
    if (open) ctx->phase_end(st);
    open = false;
  }
  ~ScopedPhase() { end(); }
};

// Host wall time of a section, reported as "host:<name>" when timing is on.
struct HostPhase {
  Ctx *ctx;
  const char *name;
  std::chrono::steady_clock::time_point t0;
  bool on;
  HostPhase(Ctx *c, const char *n) : ctx(c), name(n), on(c->timing && c->timing_only.empty()) {
    if (on) t0 = std::chrono::steady_clock::now();
  }
  ~HostPhase() {
    if (!on) return;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    auto &t = ctx->phase_totals[std::string("host:") + name];
    t.first += ms;
    t.second += 1;
  }
};

// Error plumbing: every HIP failure is turned into a return code plus text.
struct Error {
  int code;
  std::string msg;
};

#define FLOCK_HIP_CHECK(ctx, expr)                                                          \
  do {                                                                                      \
    hipError_t _e = (expr);                                                                 \
    if (_e != hipSuccess) {                                                                 \
      (ctx)->last_error = std::string(#expr) + ": " + hipGetErrorString(_e);                \
      return FLOCK_E_DEVICE;                                                                \
    }                                                                                       \
  } while (0)

// Standard-basis additive-NTT twiddles for domain dimension `dim`
// (crates/flock-core/src/ntt/additive_ntt_f128.rs:57-128): layer l, block b at
// index (1 << l) - 1 + b.
std::vector<F128> ntt_twiddles(int dim);
// The normalized subspace-polynomial rows the table is spanned from:
// twiddle (l, b) = sum over the set bits i of b of rows[dim - l - 1][1 + i].
std::vector<std::vector<F128>> ntt_subspace_rows(int dim);

}  // namespace flock
