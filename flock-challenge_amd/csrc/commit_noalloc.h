// The commit that allocates nothing (commit_noalloc.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flock_gpu.h"
#include "ctx.h"

namespace flock {

// The twiddle table is built in a fixed per-device array, so k_code is at
// most COMMIT_NOALLOC_MAX_K (2^20 twiddles, 16 MiB per GPU).
constexpr int COMMIT_NOALLOC_MAX_K = 20;
// Look up this GPU's copy of the array (once per device; the context's
// device is made current).
int commit_noalloc_init(Ctx *ctx);
// commit_device with the twiddles in that array; no device allocation.
int commit_device_noalloc(Ctx *ctx, const flock_pcs_params *p, const void *d_z, void *d_codeword, void *d_tree,
                          uint8_t root[32], hipStream_t st);

}  // namespace flock
