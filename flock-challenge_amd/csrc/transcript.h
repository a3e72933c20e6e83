// Fiat-Shamir transcript over BLAKE3: the reference FsChallenger
// (crates/flock-core/src/challenger.rs:246-437) with its op/kind framing:
// DOMAIN=1 LABEL=2 OBSERVE=3 SQUEEZE=4 BYTES=5, SCALAR=1 SLICE=2, u64le
// lengths, squeeze = XOF of a clone of the state, squeezed bytes re-absorbed.
#pragma once
#include <stdint.h>
#include <string.h>

#include <vector>

#include "host_blake3.h"
#include "host_field.h"

namespace flock {

class Transcript {
 public:
  explicit Transcript(const void *domain, size_t len) {
    u8(1);
    u64(len);
    h_.update(domain, len);
  }
  void label(const char *s) {
    size_t n = strlen(s);
    u8(2);
    u64(n);
    h_.update(s, n);
  }
  void observe(F128 v) {
    uint8_t t[2] = {3, 1};
    h_.update(t, 2);
    h_.update(&v.lo, 8);
    h_.update(&v.hi, 8);
  }
  void observe_slice(const F128 *v, size_t n) {
    uint8_t t[2] = {3, 2};
    h_.update(t, 2);
    u64(n);
    h_.update(v, 16 * n);
  }
  void observe_bytes(const void *b, size_t n) {
    u8(5);
    u64(n);
    h_.update(b, n);
  }
  F128 sample() {
    uint8_t t[2] = {4, 1};
    h_.update(t, 2);
    F128 v;
    h_.finalize_xof((uint8_t *)&v, 16);
    h_.update(&v, 16);
    return v;
  }
  std::vector<F128> sample_vec(size_t n) {
    uint8_t t[2] = {4, 2};
    h_.update(t, 2);
    u64(n);
    std::vector<F128> v(n);
    if (n) h_.finalize_xof((uint8_t *)v.data(), 16 * n);
    h_.update(v.data(), 16 * n);
    return v;
  }
  // PoW base: 32-byte root output of the current state (challenger.rs:440).
  void state_digest(uint8_t d[32]) const { h_.finalize_xof(d, 32); }
  // The hasher state as bytes (the device transcript, fs_dev.hip, keeps the
  // same layout), to move the transcript between host and device.
  static constexpr size_t STATE_BYTES = sizeof(hb3::Hasher);
  void export_state(void *dst) const { memcpy(dst, &h_, sizeof h_); }
  void import_state(const void *src) { memcpy(&h_, src, sizeof h_); }

 private:
  void u8(uint8_t v) { h_.update(&v, 1); }
  void u64(uint64_t v) { h_.update(&v, 8); }
  hb3::Hasher h_;
};

}  // namespace flock
