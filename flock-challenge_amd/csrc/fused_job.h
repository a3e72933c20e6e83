// Fused multi-instance launches: one grid covers the workgroups of several
// instances, each described by a job whose wg_begin is its first workgroup.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flock {

// The job a workgroup of a fused launch belongs to: the last one whose first
// workgroup is at most blockIdx.x (jobs sorted by wg_begin).
template <class Job>
__device__ __forceinline__ const Job &job_of(const Job *__restrict__ jobs, uint32_t n_jobs) {
  uint32_t j = 0;
  for (uint32_t i = 1; i < n_jobs; ++i)
    if (jobs[i].wg_begin <= blockIdx.x) j = i;
  return jobs[j];
}

}  // namespace flock
