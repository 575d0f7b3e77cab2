// Workgroup barrier that orders LDS only.  __syncthreads() is a
// workgroup-scope fence on every address space, which makes each wave
// wait for all of its outstanding global loads and stores before the
// barrier; the transform and statistics kernels exchange data between
// threads through LDS alone, so their barriers fence LDS and leave
// global traffic (the next row's prefetch, the last row's stores) in
// flight.
#pragma once
#include <hip/hip_runtime.h>

namespace wct {

__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace wct
