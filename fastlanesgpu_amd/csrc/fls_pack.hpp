// Packing of one lane's 32 rows into FastLanes words (the inverse of
// fls::dev::unpack_lane), shared by the GPU encoders.  Internal.
#pragma once

#include "fls/device/unpack.hpp"

namespace fls {

// rows (row order) -> the lane's BW words at vec_out + lane + 32*k
template <int BW>
__device__ __forceinline__ void pack_lane(const uint32_t (&x)[32], uint32_t* __restrict vec_out) {
    if constexpr (BW == 0) {
        return;
    } else {
        uint32_t w[BW];
#pragma unroll
        for (int k = 0; k < BW; ++k) w[k] = 0u;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const uint32_t v = x[r] & dev::Mask<BW>::value;
            const int bit = r * BW, k = bit >> 5, s = bit & 31;
            w[k] |= v << s;
            if (s + BW > 32) w[k + 1] |= v >> (32 - s);
        }
        uint32_t* o = vec_out + dev::lane32();
#pragma unroll
        for (int k = 0; k < BW; ++k) o[32 * k] = w[k];
    }
}

// run-time width
__device__ __forceinline__ void pack_rows(const uint32_t (&x)[32], uint32_t* __restrict vec_out, uint32_t bw) {
    switch (bw) {
#define FLS_PK(B) \
    case B: pack_lane<B>(x, vec_out); break;
        FLS_PK(0) FLS_PK(1) FLS_PK(2) FLS_PK(3) FLS_PK(4) FLS_PK(5) FLS_PK(6) FLS_PK(7) FLS_PK(8) FLS_PK(9)
        FLS_PK(10) FLS_PK(11) FLS_PK(12) FLS_PK(13) FLS_PK(14) FLS_PK(15) FLS_PK(16) FLS_PK(17) FLS_PK(18)
        FLS_PK(19) FLS_PK(20) FLS_PK(21) FLS_PK(22) FLS_PK(23) FLS_PK(24) FLS_PK(25) FLS_PK(26) FLS_PK(27)
        FLS_PK(28) FLS_PK(29) FLS_PK(30) FLS_PK(31) FLS_PK(32)
#undef FLS_PK
        default: break;
    }
}

}  // namespace fls
