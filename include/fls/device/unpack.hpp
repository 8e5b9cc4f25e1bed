// Device-side FastLanes decode for gfx950 (CDNA4, wave64).  Header-only HIP.
//
// Layout (fastlanes/src/pack.cpp:2970-3123): a 1024-value vector is 32 lanes x
// 32 rows; lane l's word w is packed[32*w + l]; row r of lane l is bits
// [r*bw, r*bw+bw) of the lane's LSB-first stream.
//
// Mapping on CDNA4: a wave of 64 work-items decodes two vectors.  Work-item t
// serves lane (t % 32) of vector (t / 32) of its wave, so every load
// instruction reads two fully used 128-byte runs (one per vector) and the 32
// decoded rows land in the work-item's registers -- the "fused" register form
// of the reference (unpack_fused.cuh:876-905), i.e. Crystal's striped tile
// where item j is tile index lane + 32*j.
//
// Width is a template parameter so that every shift, mask and word index is an
// immediate: rows inside one word become v_bfe_u32, rows that straddle two
// words become v_alignbit_b32 (funnel shift) + mask.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {
namespace dev {

__device__ __forceinline__ uint32_t lane32() { return threadIdx.x & 31u; }

template <int BW>
struct Mask {
    static constexpr uint32_t value = BW >= 32 ? 0xffffffffu : ((1u << BW) - 1u);
};

// Extract row R of a lane given the lane's words in registers.
template <int BW, int R>
__device__ __forceinline__ uint32_t extract_row(const uint32_t (&w)[BW > 0 ? BW : 1]) {
    if constexpr (BW == 0) {
        return 0u;
    } else if constexpr (BW == 32) {
        return w[R];
    } else {
        constexpr int bit = R * BW;
        constexpr int k = bit >> 5;
        constexpr int s = bit & 31;
        if constexpr (s + BW <= 32) {
            if constexpr (s == 0) return w[k] & Mask<BW>::value;
            else if constexpr (s + BW == 32) return w[k] >> s;
            else return __builtin_amdgcn_ubfe(w[k], s, BW);
        } else {
            return __builtin_amdgcn_alignbit(w[k + 1], w[k], s) & Mask<BW>::value;
        }
    }
}

// Same extraction with the row as an ordinary argument: inside a fully
// unrolled row loop r is a constant and this folds to one v_bfe_u32 or one
// v_alignbit_b32 + mask.
template <int BW>
__device__ __forceinline__ uint32_t extract_at(const uint32_t (&w)[BW > 0 ? BW : 1], int r) {
    if constexpr (BW == 0) {
        return 0u;
    } else if constexpr (BW == 32) {
        return w[r];
    } else {
        const int bit = r * BW, k = bit >> 5, s = bit & 31;
        if (s + BW <= 32) return __builtin_amdgcn_ubfe(w[k], s, BW);
        return __builtin_amdgcn_alignbit(w[k + 1], w[k], s) & Mask<BW>::value;
    }
}

template <int BW, int R = 0>
__device__ __forceinline__ void extract_rows(const uint32_t (&w)[BW > 0 ? BW : 1], uint32_t (&out)[32]) {
    if constexpr (R < 32) {
        out[R] = extract_row<BW, R>(w);
        extract_rows<BW, R + 1>(w, out);
    }
}

// Load the BW words of one lane: `lane_in` points at word 0 of the lane
// (vector base + lane).  Word k sits 32*k words further.
template <int BW>
__device__ __forceinline__ void load_lane(const uint32_t* __restrict lane_in, uint32_t (&w)[BW > 0 ? BW : 1]) {
    if constexpr (BW == 0) {
        w[0] = 0u;
    } else {
#pragma unroll
        for (int k = 0; k < BW; ++k) w[k] = __builtin_nontemporal_load(lane_in + 32 * k);
    }
}

// Decode one lane of one vector into 32 registers (row r -> out[r]).
// `vec_in` is the vector's first packed word; the lane comes from threadIdx.
// Counterpart of generated::unpack::cuda::fused::unpack_<bw>bw_32ow_32crw_1uf
// (fastlanes/src/include/fls_gen/unpack/unpack_fused.cuh:7-3328).
template <int BW>
__device__ __forceinline__ void unpack_lane(const uint32_t* __restrict vec_in, uint32_t (&out)[32]) {
    uint32_t w[BW > 0 ? BW : 1];
    load_lane<BW>(vec_in + lane32(), w);
    extract_rows<BW>(w, out);
}

// Runtime-width dispatch, the reference's unpack_device(in, out, bw)
// (unpack_fused.cuh:3455-3462).  Work-items of one half-wave must share bw for
// best speed; divergent widths are correct but serialise.
template <typename T>
__device__ __forceinline__ void unpack_device(const T* __restrict vec_in, T (&out_t)[32], uint32_t bw) {
    const uint32_t* in = reinterpret_cast<const uint32_t*>(vec_in);
    uint32_t(&out)[32] = reinterpret_cast<uint32_t(&)[32]>(out_t);
    switch (bw) {
#define FLS_UNPACK_CASE(B) \
    case B: unpack_lane<B>(in, out); break;
        FLS_UNPACK_CASE(0) FLS_UNPACK_CASE(1) FLS_UNPACK_CASE(2) FLS_UNPACK_CASE(3) FLS_UNPACK_CASE(4)
        FLS_UNPACK_CASE(5) FLS_UNPACK_CASE(6) FLS_UNPACK_CASE(7) FLS_UNPACK_CASE(8) FLS_UNPACK_CASE(9)
        FLS_UNPACK_CASE(10) FLS_UNPACK_CASE(11) FLS_UNPACK_CASE(12) FLS_UNPACK_CASE(13) FLS_UNPACK_CASE(14)
        FLS_UNPACK_CASE(15) FLS_UNPACK_CASE(16) FLS_UNPACK_CASE(17) FLS_UNPACK_CASE(18) FLS_UNPACK_CASE(19)
        FLS_UNPACK_CASE(20) FLS_UNPACK_CASE(21) FLS_UNPACK_CASE(22) FLS_UNPACK_CASE(23) FLS_UNPACK_CASE(24)
        FLS_UNPACK_CASE(25) FLS_UNPACK_CASE(26) FLS_UNPACK_CASE(27) FLS_UNPACK_CASE(28) FLS_UNPACK_CASE(29)
        FLS_UNPACK_CASE(30) FLS_UNPACK_CASE(31) FLS_UNPACK_CASE(32)
#undef FLS_UNPACK_CASE
    default:
#pragma unroll
        for (int r = 0; r < 32; ++r) out[r] = 0u;
        break;
    }
}

// Chain order of the FastLanes delta (rsum.cuh:8-106): lane l visits
// positions l + O_j with O_j = 128*(j%8) + {0,64,32,96}[j/8], i.e. register
// rows r_j = 4*(j%8) + {0,2,1,3}[j/8] of the striped tile.
__host__ __device__ constexpr int rsum_row(int j) {
    return 4 * (j % 8) + (j / 8 == 0 ? 0 : j / 8 == 1 ? 2 : j / 8 == 2 ? 1 : 3);
}

// In-register d_rsum_32 (rsum.cuh:3-107): the lane's 32 deltas become running
// sums seeded by the lane's base.  No LDS is needed: every position a lane
// touches is one of its own striped rows.
__device__ __forceinline__ void rsum_lane(uint32_t (&v)[32], uint32_t base) {
    uint32_t acc = base;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        acc += v[rsum_row(j)];
        v[rsum_row(j)] = acc;
    }
}

// Transposed position of original index k = 64a + 8b + c inside the vector:
// transpose_i's k -> a + 16*rev3(b) + 128*c (rev3 reverses 3 bits).
__host__ __device__ constexpr uint32_t transpose_index(uint32_t k) {
    uint32_t a = k >> 6, b = (k >> 3) & 7u, c = k & 7u;
    uint32_t rb = ((b & 1u) << 2) | (b & 2u) | ((b >> 2) & 1u);
    return a + 16u * rb + 128u * c;
}

// Original (pre-transpose) index inside the vector of transposed position
// p = lane + 32*r: inverse of transpose_i's k -> a + 16*rev3(b) + 128*c.
__host__ __device__ constexpr uint32_t untranspose_index(uint32_t p) {
    uint32_t a = p & 15u, rb = (p >> 4) & 7u, c = p >> 7;
    uint32_t b = ((rb & 1u) << 2) | (rb & 2u) | ((rb >> 2) & 1u);
    return 64u * a + 8u * b + c;
}

}  // namespace dev
}  // namespace fls
