// Wave64 reductions for gfx950.  Header-only HIP.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {
namespace dev {

// Sum over all 64 work-items of the wave; every work-item gets the total.
template <typename T>
__device__ __forceinline__ T wave_sum(T x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Sum over the 32-lane half-wave that owns one FastLanes vector.
template <typename T>
__device__ __forceinline__ T half_sum(T x) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ __forceinline__ bool is_wave_leader() { return (threadIdx.x & 63u) == 0u; }

}  // namespace dev
}  // namespace fls
