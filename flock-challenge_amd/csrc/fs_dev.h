// The Fiat-Shamir transcript on the device (fs_dev.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ctx.h"

namespace flock {

// Bytes of the transcript state, the layout of the host's hb3::Hasher.
constexpr size_t FS_STATE_BYTES = 1848;

// Observe n_msg F128 values at d_msg (device) into the transcript state
// d_state (device), then sample one challenge into d_rho (device), as
// Transcript::observe and Transcript::sample do on the host.
int fs_round_device(Ctx *ctx, hipStream_t st, void *d_state, const void *d_msg, int n_msg, void *d_rho);
// Byte tables of rho * ws[q] for the 64 values ws (device): out[v * 8 + p] =
// sum over the set bits t of v of rho ws[8 p + t] (2048 F128).
int rho_bytetab_device(Ctx *ctx, hipStream_t st, const void *d_ws, const void *d_rho, void *d_out);

}  // namespace flock
