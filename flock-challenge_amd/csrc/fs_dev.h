// The Fiat-Shamir transcript on the device (transcript_dev.h, fs_dev.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace flock {

// Bytes of the transcript state, the layout of the host's hb3::Hasher.
constexpr size_t FS_STATE_BYTES = 1848;

// Byte tables of rho * ws[q] for the 64 values ws (device): out[v * 8 + p] =
// sum over the set bits t of v of rho ws[8 p + t] (2048 F128).
int rho_bytetab_device(Ctx *ctx, hipStream_t st, const void *d_ws, const void *d_rho, void *d_out);

}  // namespace flock
