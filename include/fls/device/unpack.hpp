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

// ======================================================================
// The reference's decode entry points under their own names
// (fastlanes/src/include/fls_gen/unpack/{unpack,unpack_fused,hardcoded_16}.cuh,
// fastlanes/src/include/fls_gen/rsum/rsum.cuh), so a reference kernel body
// compiles against this header as written.  Every one works on the vector (or
// quarter vector) whose first packed word `in` points at, for the lane
// threadIdx.x % 32: launched the reference's way (32 work-items per vector)
// the lane is threadIdx.x itself; a 64-wide wave may serve two vectors, each
// half-wave passing its own `in`.
// ======================================================================

namespace fls {
namespace dev {

// Rows 0..7 of one lane of a quarter vector.  Rows 8q..8q+7 of a vector packed
// at a width that is a multiple of 4 occupy lane words [q*bw/4, (q+1)*bw/4),
// so the quarter is itself a 256-value FastLanes block of 32 lanes x 8 rows
// starting at word 8*bw*q of the vector (hardcoded_16.cuh:436-786).
template <int BW>
__device__ __forceinline__ void unpack_quarter_lane(const uint32_t* __restrict q_in, uint32_t (&out)[8]) {
    static_assert(BW % 4 == 0 && BW >= 0 && BW <= 32, "a quarter vector needs a width that is a multiple of 4");
    if constexpr (BW == 0) {
#pragma unroll
        for (int r = 0; r < 8; ++r) out[r] = 0u;
    } else {
        uint32_t w[BW] = {};
        const uint32_t* lane_in = q_in + lane32();
#pragma unroll
        for (int k = 0; k < BW / 4; ++k) w[k] = lane_in[32 * k];
#pragma unroll
        for (int r = 0; r < 8; ++r) out[r] = extract_at<BW>(w, r);
    }
}

// Decode one lane into memory in the reference's "normal" order:
// out[32*r + lane] = row r.
template <int BW>
__device__ __forceinline__ void unpack_lane_to(const uint32_t* __restrict vec_in, uint32_t* __restrict out) {
    uint32_t v[32];
    unpack_lane<BW>(vec_in, v);
    const uint32_t l = lane32();
#pragma unroll
    for (int r = 0; r < 32; ++r) out[32 * r + l] = v[r];
}

}  // namespace dev
}  // namespace fls

#define FLS_FOR_EACH_BW(X)                                                                                       \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) \
        X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)
#define FLS_FOR_EACH_BW4(X) X(0) X(4) X(8) X(12) X(16) X(20) X(24) X(28) X(32)

namespace generated {
namespace unpack {
namespace cuda {

// unpack_fused.cuh:7-3433: one lane's 32 rows into registers, out[r] = row r.
namespace fused {
#define FLS_FUSED_NAME(B)                                                                                     \
    __device__ __forceinline__ void unpack_##B##bw_32ow_32crw_1uf(const uint32_t* __restrict in,              \
                                                                  uint32_t* __restrict out) {                 \
        fls::dev::unpack_lane<B>(in, *reinterpret_cast<uint32_t(*)[32]>(out));                                \
    }
FLS_FOR_EACH_BW(FLS_FUSED_NAME)
#undef FLS_FUSED_NAME

// A width outside 0..32 (which the reference has no case for) decodes as zeros,
// here and in normal::unpack.
__device__ __forceinline__ void unpack(const uint32_t* __restrict in, uint32_t* __restrict out, uint8_t bw) {
    fls::dev::unpack_device(in, *reinterpret_cast<uint32_t(*)[32]>(out), bw);
}
}  // namespace fused

// unpack.cuh:7-3434: one lane's 32 rows into memory, out[32*r + lane] = row r.
namespace normal {
#define FLS_NORMAL_NAME(B)                                                                                    \
    __device__ __forceinline__ void unpack_##B##bw_32ow_32crw_1uf(const uint32_t* __restrict in,              \
                                                                  uint32_t* __restrict out) {                 \
        fls::dev::unpack_lane_to<B>(in, out);                                                                 \
    }
FLS_FOR_EACH_BW(FLS_NORMAL_NAME)
#undef FLS_NORMAL_NAME

__device__ __forceinline__ void unpack(const uint32_t* __restrict in, uint32_t* __restrict out, uint8_t bw) {
    switch (bw) {
#define FLS_NORMAL_CASE(B) \
    case B: unpack_##B##bw_32ow_32crw_1uf(in, out); break;
        FLS_FOR_EACH_BW(FLS_NORMAL_CASE)
#undef FLS_NORMAL_CASE
    default:
#pragma unroll
        for (int r = 0; r < 32; ++r) out[32 * r + fls::dev::lane32()] = 0u;
        break;
    }
}
}  // namespace normal

}  // namespace cuda
}  // namespace unpack
}  // namespace generated

// unpack_fused.cuh:3455-3462 / unpack.cuh:3444-3451.  The reference picks the
// register ("fused") or the memory ("normal") decode by which header it
// includes; here the destination's type picks it: a 32-element array is one
// lane's registers (out[r] = row r), a 1024-element array is the whole vector
// in memory, e.g. in LDS (out[32*r + lane] = row r).
template <typename T>
__device__ __forceinline__ void unpack_device(const T* __restrict in, T (&out)[32], uint32_t bw) {
    fls::dev::unpack_device(in, out, bw);
}

template <typename T>
__device__ __forceinline__ void unpack_device(const T* __restrict in, T (&out)[1024], uint32_t bw) {
    generated::unpack::cuda::normal::unpack(reinterpret_cast<const uint32_t*>(in), reinterpret_cast<uint32_t*>(out),
                                            (uint8_t)bw);
}

// unpack_fused.cuh:3436-3442: out[j*BLOCK_THREADS + i] = registers[j].
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void load_registers(int i, T* out, const T* registers) {
#pragma unroll
    for (int j = 0; j < ITEMS_PER_THREAD; ++j) out[j * BLOCK_THREADS + i] = registers[j];
}

// unpack_fused.cuh:3444-3453: decode whole vectors of a fixed-width column,
// vector v from in + 32*bw*v to out + 1024*v.  The reference launches one
// 32-work-item block per vector; any block size that is a multiple of 32
// works here (each half-wave takes the next vector), so a grid of
// n_vec * 32 / blockDim.x blocks covers n_vec vectors.  A template with a
// defaulted parameter so the header can be included by several translation
// units; `unpack_global<<<grid, block>>>(in, out, bw)` launches it.
template <int = 0>
__global__ void unpack_global(const uint32_t* __restrict in, uint32_t* __restrict out, uint8_t bw) {
    const uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / 32u) + threadIdx.x / 32u;
    uint32_t registers[32];
    fls::dev::unpack_device(in + v * 32u * bw, registers, bw);
    load_registers<uint32_t, 32, 32>((int)fls::dev::lane32(), out + (v << 10), registers);
}

// hardcoded_16.cuh:436-786: quarter-vector decode, 8 rows per lane
// (out[r] = row r of the quarter, i.e. row 8q + r of the vector).  The
// reference defines widths 0, 4, ..., 24 and traps on others; this defines
// every multiple of 4 up to 32 and leaves out zero for any other width (a
// quarter of such a vector does not start on a word boundary).
namespace unpack_8_at_a_time {
#define FLS_QUARTER_NAME(B)                                                                                   \
    __device__ __forceinline__ void unpack_##B##bw_32ow_32crw_1uf(const uint32_t* __restrict in,              \
                                                                  uint32_t* __restrict out) {                 \
        fls::dev::unpack_quarter_lane<B>(in, *reinterpret_cast<uint32_t(*)[8]>(out));                         \
    }
FLS_FOR_EACH_BW4(FLS_QUARTER_NAME)
#undef FLS_QUARTER_NAME

__device__ __forceinline__ void unpack(const uint32_t* __restrict in, uint32_t* __restrict out, uint8_t bw) {
    switch (bw) {
#define FLS_QUARTER_CASE(B) \
    case B: unpack_##B##bw_32ow_32crw_1uf(in, out); break;
        FLS_FOR_EACH_BW4(FLS_QUARTER_CASE)
#undef FLS_QUARTER_CASE
    default:
#pragma unroll
        for (int r = 0; r < 8; ++r) out[r] = 0u;
        break;
    }
}

__device__ __forceinline__ void unpack_device(const uint32_t* __restrict in, uint32_t* __restrict out, uint8_t bw) {
    unpack(in, out, bw);
}

__device__ __forceinline__ void unpack_device(const int32_t* __restrict in, int32_t* __restrict out, uint8_t bw) {
    unpack(reinterpret_cast<const uint32_t*>(in), reinterpret_cast<uint32_t*>(out), bw);
}
}  // namespace unpack_8_at_a_time

// rsum.cuh:3-107: FastLanes delta decode of one vector in memory.  Lane l
// walks its 32 positions l + O_j (O_j = 128*(j%8) + {0,64,32,96}[j/8]) in
// chain order, out[p] = base[l] + in[first] + ... + in[p]; in and out hold the
// vector's 1024 values in transposed order (in LDS or global memory).
__device__ __forceinline__ void d_rsum_32(const uint32_t* in, uint32_t* out, const uint32_t* base) {
    const uint32_t l = fls::dev::lane32();
    uint32_t acc = base[l];
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const uint32_t p = l + 32u * (uint32_t)fls::dev::rsum_row(j);
        acc += in[p];
        out[p] = acc;
    }
}

#undef FLS_FOR_EACH_BW
#undef FLS_FOR_EACH_BW4
