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

// Inclusive prefix sum over the 64 work-items of the wave (work-item i gets
// x_0 + ... + x_i), with DPP: shifts by 1, 2, 4 and 8 within each 16-lane row
// (lanes shifted in from outside the row read 0), then lane 15 of a row added
// to rows 1 and 3, then lane 31 to rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);  // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);  // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);  // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);  // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

}  // namespace dev
}  // namespace fls
