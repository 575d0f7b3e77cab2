// Global-memory views of generic pointers.  A pointer read out of a job
// table is generic (flat) to the compiler, and flat loads and stores
// count against the LDS counter as well as the vector-memory counter: an
// LDS wait or a workgroup barrier then also waits for every outstanding
// global access (a row prefetch, the last row's stores).  Every pointer
// a kernel takes from a job table and dereferences goes through G(),
// which types it as global memory (address space 1), so the accesses are
// global_load / global_store and count against the vector-memory counter
// only.  The pointers are device allocations (hipMalloc or torch CUDA
// tensors), which always live in global memory.
#pragma once
#include <hip/hip_runtime.h>

namespace wct {

// The host pass of a HIP translation unit also checks kernel bodies; it
// has no global address space, so there the view is the plain pointer.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
#else
template <typename T>
using gptr = T*;
#endif

template <typename T>
__device__ __forceinline__ gptr<T> G(T* p)
{
    return (gptr<T>)p;
}

}  // namespace wct
