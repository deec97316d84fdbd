// Work-item and workgroup ids straight from the AMDGPU builtins.
//
// The kernels are built with -mno-amdgpu-ieee; the device library's accessors
// behind threadIdx / blockIdx are compiled with IEEE mode on, and the inliner
// refuses to inline across that difference, so every threadIdx.x would be a
// real function call (s_swappc) with its stack set-up.  These helpers read
// the same hardware values without the call.
#pragma once
#include <hip/hip_runtime.h>

namespace simd2 {

__device__ __forceinline__ int tid_x() { return (int)__builtin_amdgcn_workitem_id_x(); }
__device__ __forceinline__ int bid_x() { return (int)__builtin_amdgcn_workgroup_id_x(); }
__device__ __forceinline__ int bid_y() { return (int)__builtin_amdgcn_workgroup_id_y(); }

}  // namespace simd2
