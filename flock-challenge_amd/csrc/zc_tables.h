// Constant tables of zerocheck round 1, built on the host (prover.cpp) and
// copied to the device once per context.
#pragma once
#include <stdint.h>

namespace flock {

struct ZcTables {
  // S -> Lambda extension by byte: the 64 Lambda bytes of the contribution
  // of byte value v at byte position p of a row sit at byte offset
  // (v * 8 + p) * 64 (16-byte chunk q4 = lambda / 16 at (v * 8 + p) * 4 + q4).
  alignas(16) uint32_t ext8[256 * 8 * 16];
  uint16_t log8[256];          // discrete log base 3; log8[0] = 1024 (zero sentinel)
  uint8_t exp8[3 * 1024 + 4];  // exp8[i] = 3^(i mod 255) for i < 1024, else 0
  uint16_t log_eps[8];         // log of eps(x_lo)
  alignas(16) uint32_t pt[16][8][4];  // E4(x_mid) * phi8(2^t), F128 as 4 LE dwords
};

}  // namespace flock
