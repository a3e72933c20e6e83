// Constant tables of zerocheck round 1, built on the host (prover.cpp) and
// copied to the device once per context.
#pragma once
#include <stdint.h>

namespace flock {

struct ZcTables {
  // Bitsliced S -> Lambda extension.  For byte value v at byte position p of
  // a row, half h of the 64 lambdas and plane group k, uint4 number
  // k * 4096 + v * 16 + 2 p + h holds 4 dwords: dword d is bit 4 k + d of
  // the contribution's values at lambdas 32 h .. 32 h + 31 (bit b of the
  // dword <-> lambda 32 h + b).  Its 16-byte bank group is 2 p + h.
  alignas(16) uint32_t extb[2 * 4096 * 4];
  alignas(16) uint32_t pt[16][8][4];  // E4(x_mid) * phi8(2^t), F128 as 4 LE dwords
};

}  // namespace flock
