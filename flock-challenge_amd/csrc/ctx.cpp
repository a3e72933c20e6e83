// Context, scratch arena and constant tables.
#include "ctx.h"

#include "../../include/flock_gpu.h"

namespace flock {

const Phi8 &phi8_table() {
  static const Phi8 t;
  return t;
}

// Normalized subspace-polynomial rows evals[i][k] = W_i(b_{i+k}) / W_i(b_i)
// for the standard basis b_i = x^i, then the breadth-first span table
// (ntt/additive_ntt_f128.rs:57-89, 106-128).
std::vector<std::vector<F128>> ntt_subspace_rows(int dim) {
  std::vector<std::vector<F128>> ev(dim);
  ev[0].resize(dim);
  for (int i = 0; i < dim; ++i) ev[0][i] = i < 64 ? F128(1ull << i, 0) : F128(0, 1ull << (i - 64));
  for (int i = 1; i < dim; ++i) {
    ev[i].resize(dim - i);
    for (int k = 1; k <= dim - i; ++k) {
      F128 p = ev[i - 1][k];
      ev[i][k - 1] = p * (p + ev[i - 1][0]);
    }
  }
  for (int i = 0; i < dim; ++i) {
    F128 iv = inv(ev[i][0]);
    for (auto &x : ev[i]) x = x * iv;
  }
  return ev;
}

std::vector<F128> ntt_twiddles(int dim) {
  const std::vector<std::vector<F128>> ev = ntt_subspace_rows(dim);
  std::vector<F128> tw(dim ? ((size_t)1 << dim) - 1 : 0);
  for (int layer = 0; layer < dim; ++layer) {
    size_t start = ((size_t)1 << layer) - 1;
    const auto &row = ev[dim - layer - 1];
    tw[start] = F128();
    for (int bit = 0; bit < layer; ++bit) {
      size_t half = (size_t)1 << bit;
      for (size_t b = 0; b < half; ++b) tw[start + half + b] = tw[start + b] + row[1 + bit];
    }
  }
  return tw;
}

const std::vector<F128> &Ctx::host_twiddles(int dim) {
  auto it = twiddles_host.find(dim);
  if (it != twiddles_host.end()) return it->second;
  return twiddles_host[dim] = ntt_twiddles(dim);
}

const F128 *Ctx::device_twiddles(int dim) {
  auto it = twiddles_dev.find(dim);
  if (it != twiddles_dev.end()) return it->second;
  const auto &h = host_twiddles(dim);
  F128 *d = nullptr;
  size_t bytes = std::max<size_t>(1, h.size()) * sizeof(F128);
  if (hipMalloc(&d, bytes) != hipSuccess) {
    last_error = "hipMalloc twiddles failed";
    return nullptr;
  }
  if (!h.empty() && hipMemcpy(d, h.data(), h.size() * sizeof(F128), hipMemcpyHostToDevice) != hipSuccess) {
    last_error = "hipMemcpy twiddles failed";
    hipFree(d);
    return nullptr;
  }
  twiddles_dev[dim] = d;
  return d;
}

int Ctx::wait_failed(WaitEnd how, hipError_t err, const char *what) {
  if (how == WaitEnd::TimedOut) {
    timed_out = true;
    last_error = std::string(what) + ": the device did not finish within " + std::to_string(wait_timeout_s) +
                 " s; the context is unusable";
  } else if (err != hipSuccess) {
    last_error = std::string(what) + ": " + hipGetErrorString(err);
  } else {
    last_error = std::string(what) + ": the stream finished without publishing its signal";
  }
  return FLOCK_E_DEVICE;
}

int Ctx::wait_flag_seq(const uint32_t *flag, uint32_t seq, hipStream_t st) {
  hipError_t err = hipSuccess;
  // hipStreamQuery now and then: a failed kernel, or a stream that drained
  // without publishing, ends the wait
  const WaitEnd how = wait_flag(
      flag, seq,
      [&] {
        const hipError_t q = hipStreamQuery(st);
        if (q == hipErrorNotReady) return 0;
        if (q == hipSuccess) return 1;
        err = q;
        return -1;
      },
      wait_timeout_s);
  return how == WaitEnd::Done ? FLOCK_OK : wait_failed(how, err, "readback");
}

int Ctx::wait_stream(hipStream_t st) {
  hipError_t err = hipSuccess;
  bool fin = false;
  const WaitEnd how = bounded_wait([&] { return fin; },
                                   [&] {
                                     const hipError_t q = hipStreamQuery(st);
                                     if (q == hipErrorNotReady) return 0;
                                     if (q == hipSuccess) return (int)(fin = true);
                                     err = q;
                                     return -1;
                                   },
                                   wait_timeout_s, 16);
  return how == WaitEnd::Done ? FLOCK_OK : wait_failed(how, err, "stream wait");
}

int Ctx::wait_event(hipEvent_t e) {
  hipError_t err = hipSuccess;
  bool fin = false;
  const WaitEnd how = bounded_wait([&] { return fin; },
                                   [&] {
                                     const hipError_t q = hipEventQuery(e);
                                     if (q == hipErrorNotReady) return 0;
                                     if (q == hipSuccess) return (int)(fin = true);
                                     err = q;
                                     return -1;
                                   },
                                   wait_timeout_s, 16);
  return how == WaitEnd::Done ? FLOCK_OK : wait_failed(how, err, "event wait");
}

void *Ctx::get_scratch(const std::string &name, size_t bytes) {
  auto &b = scratch[name];
  if (b.bytes >= bytes && b.ptr) return b.ptr;
  if (b.ptr) {
    // this context's queued kernels may still use the old buffer
    if (wait_stream(stream)) return nullptr;
    hipFree(b.ptr);
  }
  b.ptr = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.ptr, bytes) != hipSuccess) {
    b.ptr = nullptr;
    last_error = "hipMalloc scratch '" + name + "' failed";
    return nullptr;
  }
  b.bytes = bytes;
  return b.ptr;
}

int Ctx::release_scratch() {
  if (int rc = wait_stream(stream)) return rc;
  for (auto &kv : scratch)
    if (kv.second.ptr) hipFree(kv.second.ptr);
  scratch.clear();
  // the tables these flags stand for lived in scratch buffers
  csc_uploaded = false;
  zc_tables_uploaded = false;
  return FLOCK_OK;
}

void Ctx::free_scratch(const std::string &name) {
  auto it = scratch.find(name);
  if (it == scratch.end()) return;
  if (wait_stream(stream)) return;
  if (it->second.ptr) hipFree(it->second.ptr);
  scratch.erase(it);
}

hipEvent_t Ctx::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

void Ctx::phase_begin(const std::string &name, hipStream_t st) {
  PhaseEvent e;
  e.name = name;
  e.begin = take_event();
  e.end = take_event();
  hipEventRecord(e.begin, st);
  open_phases.push_back(phase_events.size());
  phase_events.push_back(e);
}

void Ctx::phase_end(hipStream_t st) {
  if (open_phases.empty()) return;
  hipEventRecord(phase_events[open_phases.back()].end, st);
  open_phases.pop_back();
}

void Ctx::resolve_phases() {
  for (auto &e : phase_events) {
    hipEventSynchronize(e.end);
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.begin, e.end) == hipSuccess) {
      auto &t = phase_totals[e.name];
      t.first += ms;
      t.second += 1;
    }
    event_pool.push_back(e.begin);
    event_pool.push_back(e.end);
  }
  phase_events.clear();
}

uint8_t *Ctx::pinned_alloc(size_t n) {
  if (!pinned) {
    const size_t want = 64ull << 20;
    void *p = nullptr;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
    pinned = (uint8_t *)p;
    pinned_bytes = want;
    pinned_off = 0;
  }
  const size_t at = (pinned_off + 255) & ~(size_t)255;
  if (at + n > pinned_bytes) return nullptr;
  pinned_off = at + n;
  return pinned + at;
}

HostPool &Ctx::host_pool() {
  if (!pool) {
    const unsigned hw = std::thread::hardware_concurrency();
    pool.reset(new HostPool(hw >= 8 ? 7 : (hw > 1 ? (int)hw - 1 : 0)));
  }
  return *pool;
}

Ctx::~Ctx() {
  // Work that timed out may still be running: leave its memory and stream
  // alone (held until the process ends) rather than block on it.
  if (timed_out) return;
  hipSetDevice(device);
  for (auto &e : phase_events) {
    hipEventDestroy(e.begin);
    hipEventDestroy(e.end);
  }
  for (hipEvent_t e : event_pool) hipEventDestroy(e);
  for (auto &kv : zc_graphs) (void)hipGraphExecDestroy(kv.second);
  if (pinned) hipHostFree(pinned);
  if (landing) hipHostFree(landing);
  if (open_area) hipHostFree(open_area);
  if (readback) hipHostFree(readback);
  for (auto &kv : twiddles_dev) hipFree(kv.second);
  for (auto &kv : scratch)
    if (kv.second.ptr) hipFree(kv.second.ptr);
  if (stream) hipStreamDestroy(stream);
}

}  // namespace flock
