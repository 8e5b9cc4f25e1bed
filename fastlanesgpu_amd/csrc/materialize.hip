// Untranspose and decode-to-materialize on the GPU.
//
// transpose_i / transpose_o (fastlanes/src/transpose.cpp:4110-5135 and the
// inverse that transpose.cpp:5136-6161 is meant to be) over whole 1024-value
// vectors: each vector is staged in LDS with coalesced loads and written back
// through the permutation with coalesced stores.
//
// fls_decode_column_u32 turns a compressed column of any encoding into plain
// values for a downstream operator (the role of Crystal's BlockStore,
// crystal-opt/store.cuh:57-69, after a decode): in the order the column is
// stored, or in original row order, which untransposes DELTA / RLE / MIXED /
// transposed BITPACK columns on the fly so they line up with plain columns.
#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

constexpr int kTrBlock = 256;                 // 4 waves, one vector each
constexpr int kTrVecs = kTrBlock / 64;
constexpr uint32_t kTrPad = 1024u + 64u;      // words per staged vector

// LDS slot of index k: four padding words per 64 keep 16-byte groups aligned
// and the strided gathers of the permutation on distinct banks.
__device__ __forceinline__ uint32_t tr_slot(uint32_t k) { return k + 4u * (k >> 6); }

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// A wave per vector: four 16-byte loads per work-item (a work-item holds
// words 4t + 256j .. +3) into LDS in stored order, then four 16-byte stores
// of four consecutive output words, each gathered from LDS through the
// permutation.  (One word per load and per store, 16 each, took 205 us over
// 117,188 vectors.)
template <bool INVERSE>
__global__ __launch_bounds__(kTrBlock) void transpose_kernel(const uint32_t* __restrict in, uint64_t n_vec,
                                                             uint32_t* __restrict out) {
    __shared__ __align__(16) uint32_t stage[kTrVecs][kTrPad];
    const uint32_t w = threadIdx.x >> 6, t = threadIdx.x & 63u;
    uint32_t* st = stage[w];
    for (uint64_t g = blockIdx.x; g * kTrVecs < n_vec; g += gridDim.x) {
        const uint64_t v = g * kTrVecs + w;
        if (v < n_vec) {  // wave-uniform: one vector per wave
            const u32x4* src = reinterpret_cast<const u32x4*>(in + v * 1024u);
            u32x4 x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) x[j] = __builtin_nontemporal_load(src + t + 64u * j);
#pragma unroll
            for (int j = 0; j < 4; ++j) *reinterpret_cast<u32x4*>(st + tr_slot(4u * t + 256u * j)) = x[j];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (v < n_vec) {
            u32x4* dst = reinterpret_cast<u32x4*>(out + v * 1024u);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                u32x4 y;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const uint32_t p = 4u * t + 256u * j + i;
                    // transpose_i: position p holds original index untranspose_index(p);
                    // transpose_o: original index p sits at position transpose_index(p)
                    y[i] = st[tr_slot(INVERSE ? dev::transpose_index(p) : dev::untranspose_index(p))];
                }
                __builtin_nontemporal_store(y, dst + t + 64u * j);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Both kernels launch a grid of every vector group (over the resident set
// with a grid-stride loop they measured 8-10 % slower); materialize's stores
// are non-temporal (2-3 % faster there, 13 % slower in the transpose).
constexpr int kMatBlock = 256;          // 4 waves, two vectors per wave
constexpr int kMatVecs = kMatBlock / 32;
constexpr uint32_t kRowPad = 33u;       // words per 32 rows of a staged vector

// One vector per 32 lanes: decode (any encoding) into the lane's 32 rows,
// then store.  ROWS with a transposed column: row r of lane l is original row
// untranspose_index(l + 32 r); the vector is staged in LDS by original row
// (33 words per 32 rows) and read back by rows 32 r + l, so the stores are
// coalesced.  ROWS writes rows < n_rows only; stored order writes whole
// vectors.  VEC (a 16-byte aligned output): the vector goes through the LDS
// stage in either order (stored order: word 32 r + l at its own index) and
// each work-item stores four consecutive words at a time.
template <bool ROWS, bool VEC = false>
__global__ __launch_bounds__(kMatBlock) void materialize_kernel(const ColDev c, uint64_t n_rows, uint64_t n_vec,
                                                                uint32_t* __restrict out) {
    constexpr bool kStage = ROWS || VEC;
    __shared__ uint32_t stage[kStage ? kMatVecs : 1][kStage ? 32u * kRowPad : 1u];
    const uint32_t l = dev::lane32();
    const uint32_t slot = threadIdx.x >> 5;
    for (uint64_t g = blockIdx.x; g * kMatVecs < n_vec; g += gridDim.x) {
        const uint64_t v = g * kMatVecs + slot;
        const bool live = v < n_vec;
        uint32_t x[32];
        if (live) decode_regs(c, v, x);
        uint32_t* dst = out + v * 1024u;
        if constexpr (!kStage) {
            if (live) {
#pragma unroll
                for (int r = 0; r < 32; ++r) __builtin_nontemporal_store(x[r], dst + 32u * r + l);
            }
        } else {
            const uint32_t rows = !live ? 0u : ROWS ? (uint32_t)min<uint64_t>(1024u, n_rows - v * 1024u) : 1024u;
            if (ROWS && !c.transposed) {
#pragma unroll
                for (int r = 0; r < 32; ++r)
                    if (32u * r + l < rows) __builtin_nontemporal_store(x[r], dst + 32u * r + l);
            } else {
                uint32_t* st = stage[slot];
                if (live) {
#pragma unroll
                    for (int r = 0; r < 32; ++r) {
                        const uint32_t i = ROWS ? dev::untranspose_index(l + 32u * r) : l + 32u * r;
                        st[i + (i >> 5)] = x[r];
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if constexpr (VEC) {
                    // rows 4l + 128j .. +3 per work-item: four LDS reads (on
                    // distinct banks across the half-wave) and one 16-byte store
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const uint32_t r0 = 4u * l + 128u * (uint32_t)j;
                        const uint32_t* sp = st + r0 + (r0 >> 5);
                        if (r0 + 4u <= rows) {
                            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                            __builtin_nontemporal_store(u32x4{sp[0], sp[1], sp[2], sp[3]},
                                                        reinterpret_cast<u32x4*>(dst + r0));
                        } else {
                            for (uint32_t i = 0; r0 + i < rows; ++i) dst[r0 + i] = sp[i];
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 32; ++r)
                        if (32u * r + l < rows) __builtin_nontemporal_store(st[kRowPad * r + l], dst + 32u * r + l);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
    }
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_transpose_u32(const uint32_t* d_in, uint64_t n_vec, uint32_t inverse, uint32_t* d_out, void* stream) {
    using namespace fls;
    if (n_vec == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_out, FLS_ERR_NULL);
    FLS_RETURN_IF(d_in == d_out, FLS_ERR_ARG);
    FLS_RETURN_IF((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) % 16u, FLS_ERR_ARG);
    uint32_t grid = 0;
    const uint64_t work = ceil_div(n_vec, kTrVecs);  // a grid of every group: the grid-stride loop runs once
    FLS_RETURN_IF(work > 0x7fffffffu, FLS_ERR_SIZE);
    if (inverse) {
        grid = (uint32_t)work;
        hipLaunchKernelGGL(transpose_kernel<true>, dim3(grid), dim3(kTrBlock), 0, as_stream(stream), d_in, n_vec,
                           d_out);
    } else {
        grid = (uint32_t)work;
        hipLaunchKernelGGL(transpose_kernel<false>, dim3(grid), dim3(kTrBlock), 0, as_stream(stream), d_in, n_vec,
                           d_out);
    }
    return check_launch();
}

int fls_decode_column_u32(const fls_column* col, uint64_t n_rows, uint32_t order, uint32_t* d_out, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!col, FLS_ERR_NULL);
    FLS_RETURN_IF(order != FLS_ORDER_STORED && order != FLS_ORDER_ROWS, FLS_ERR_ARG);
    ColDev c{};
    if (const int st = to_dev_column(*col, n_rows, &c, true)) return st;
    if (n_rows == 0) return FLS_OK;
    FLS_RETURN_IF(!d_out, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n_rows, 1024);
    const uint64_t work = ceil_div(n_vec, kMatVecs);  // a grid of every group: the grid-stride loop runs once
    FLS_RETURN_IF(work > 0x7fffffffu, FLS_ERR_SIZE);
    uint32_t grid = 0;
    if (order == FLS_ORDER_ROWS) {
        grid = (uint32_t)work;
        // 16-byte stores where the output allows them
        auto k = reinterpret_cast<uintptr_t>(d_out) % 16u ? materialize_kernel<true, false>
                                                           : materialize_kernel<true, true>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kMatBlock), 0, as_stream(stream), c, n_rows, n_vec, d_out);
    } else {
        grid = (uint32_t)work;
        auto k = reinterpret_cast<uintptr_t>(d_out) % 16u ? materialize_kernel<false, false>
                                                           : materialize_kernel<false, true>;
        hipLaunchKernelGGL(k, dim3(grid), dim3(kMatBlock), 0, as_stream(stream), c, n_rows, n_vec, d_out);
    }
    return check_launch();
}

}  // extern "C"
