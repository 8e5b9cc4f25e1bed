// Grid-level FastLanes decode kernels (plain bit-unpack and FastLanes delta)
// and their C-ABI launchers.
//
// One wave decodes two vectors (work-item t: lane t%32 of vector t/32 of the
// wave); a 256-work-item block covers 8 vectors.  Each work-item issues bw
// coalesced dword loads (two 128-byte runs per wave instruction) and 32
// coalesced dword stores.
#include <type_traits>

#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

constexpr int kBlock = 256;
constexpr int kVecPerBlock = kBlock / 32;

// Checksum reduction in one launch, with no zero-fill before it.  The
// caller's buffer holds the result (word 0) and the reduction's state: per
// group g < kSumGroups of blocks (block b is in group b % kSumGroups) a
// running sum and a completion count on their own 128-byte line (words
// 16(g+1), 16(g+1)+1), and the same for the groups (words 16(kSumGroups+1)
// and +1).  A block adds its sum to its group's and counts itself; the
// group's last block moves the group sum up and counts the group; the last
// group's block writes the result.  Each atomic waits for its predecessor to
// be performed (returning form), so a count never runs ahead of its sum, and
// the last one to touch a sum or a count resets it: the state is zero again
// when the launch ends.  Spreading the sums over 64 lines keeps any one
// address to about 2 x 32 atomics (one per block on a single address
// serialises: 2,048 blocks measured 15 us).
constexpr uint32_t kSumGroups = 64;
static_assert(FLS_CHECKSUM_WORDS == 16u * (kSumGroups + 2u), "checksum buffer layout");

__device__ __forceinline__ void checksum_finish(unsigned long long t, uint64_t* __restrict ck) {
    const uint32_t grid = gridDim.x, g = blockIdx.x % kSumGroups;
    const uint32_t n_g = (grid - 1u - g) / kSumGroups + 1u;          // blocks of group g
    const uint32_t n_groups = grid < kSumGroups ? grid : kSumGroups;  // groups with a block
    unsigned long long* line = reinterpret_cast<unsigned long long*>(ck) + 16u * (g + 1u);
    unsigned long long prev = atomicAdd(line, t);
    asm volatile("" : : "v"(prev));  // performed before the count below
    if (atomicInc(reinterpret_cast<unsigned int*>(line + 1), n_g - 1u) != n_g - 1u) return;
    const unsigned long long s = atomicExch(line, 0ull);
    unsigned long long* top = reinterpret_cast<unsigned long long*>(ck) + 16u * (kSumGroups + 1u);
    prev = atomicAdd(top, s);
    asm volatile("" : : "v"(prev));
    if (atomicInc(reinterpret_cast<unsigned int*>(top + 1), n_groups - 1u) != n_groups - 1u) return;
    ck[0] = atomicExch(top, 0ull);
}

// Decode [+ write] [+ checksum sum_i (i+1)*v_i] of fixed-width vectors, one
// group of 8 vectors per block (a grid of every group: a grid-stride loop
// over the resident set measured 15-20 % slower at 2^24 values).  Writes to a
// 16-byte aligned output go eight rows at a time through an 8.4 KB LDS stage
// as 16-byte stores (2.3 % faster than one word per store at 2^24 values; a
// whole-vector stage, 33.8 KB per block, halved the resident blocks and was
// 5 % slower).  The block's
// checksum is reduced before its stores are issued: the reduction's atomics
// return a value, and on gfx950 waiting for one also waits for the wave's
// outstanding stores.
template <int BW, bool WRITE, bool CHECKSUM>
__global__ __launch_bounds__(kBlock) void unpack_global_kernel(const uint32_t* __restrict in,
                                                               uint32_t* __restrict out, uint64_t n_vec,
                                                               uint64_t n_values, uint64_t* __restrict checksum) {
    const uint32_t l = dev::lane32();
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const bool live = v < n_vec;
    uint32_t vals[32];
    if (live) dev::unpack_lane<BW>(in + v * (32u * BW), vals);
    if constexpr (CHECKSUM) {
        // sum_r (first + 32r) * v_r with first = v*1024 + l + 1
        //   = first * S0 + 32 * S1,  S0 = sum v_r,  S1 = sum r*v_r   (mod 2^64);
        // below 24 bits S0 < 2^31 and S1 <= 496 (2^BW - 1) < 2^32 fit 32 bits
        using Acc = typename std::conditional<(BW <= 23), uint32_t, uint64_t>::type;
        uint64_t acc = 0;
        if (live) {
            const uint64_t first = v * 1024u + l + 1u;
            Acc s0 = 0, s1 = 0;
            if ((v + 1) * 1024u <= n_values) {
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    s0 += vals[r];
                    s1 += (Acc)vals[r] * (uint32_t)r;
                }
            } else {
#pragma unroll
                for (int r = 0; r < 32; ++r) {
                    const uint32_t x = first + 32u * r <= n_values ? vals[r] : 0u;
                    s0 += x;
                    s1 += (Acc)x * (uint32_t)r;
                }
            }
            acc = first * (uint64_t)s0 + 32u * (uint64_t)s1;
        }
        __shared__ uint64_t s_part[kBlock / 64];
        acc = dev::wave_sum(acc);
        if (dev::is_wave_leader()) s_part[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t t = 0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) t += s_part[w];
            checksum_finish(t, checksum);
        }
    }
    if constexpr (WRITE) {
        if ((reinterpret_cast<uintptr_t>(out) & 15u) == 0u) {
            // eight rows at a time through a small LDS stage (33 words per 32
            // values), then two 16-byte stores of four consecutive words per
            // work-item
            __shared__ uint32_t stage[kVecPerBlock][8u * 33u];
            uint32_t* st = stage[threadIdx.x >> 5];
            typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
            uint32_t* o = out + v * 1024u;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                if (live) {
#pragma unroll
                    for (int r = 0; r < 8; ++r) st[33u * r + l] = vals[8 * c + r];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                if (live) {
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const uint32_t p0 = 4u * l + 128u * (uint32_t)j;
                        const uint32_t* sp = st + p0 + (p0 >> 5);
                        __builtin_nontemporal_store(u32x4{sp[0], sp[1], sp[2], sp[3]},
                                                    reinterpret_cast<u32x4*>(o + 256u * c + p0));
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();  // the stage is rewritten next
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        } else if (live) {
            uint32_t* o = out + v * 1024u + l;
#pragma unroll
            for (int r = 0; r < 32; ++r) __builtin_nontemporal_store(vals[r], o + 32 * r);
        }
    }
}

// FastLanes delta decode: per-vector (offset, bw), 32 bases per vector.
__global__ __launch_bounds__(kBlock) void delta_decode_kernel(const uint32_t* __restrict words,
                                                              const fls_vmeta* __restrict vmeta,
                                                              const uint32_t* __restrict bases, uint64_t n_vec,
                                                              uint32_t* __restrict out) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const uint32_t l = dev::lane32();
    if (v >= n_vec) return;
    const fls_vmeta m = vmeta[v];
    uint32_t vals[32];
    dev::unpack_device(words + FLS_VMETA_OFFSET(m), vals, FLS_VMETA_BW(m));
    dev::rsum_lane(vals, bases[v * 32u + l]);
    uint32_t* o = out + v * 1024u + l;
#pragma unroll
    for (int r = 0; r < 32; ++r) __builtin_nontemporal_store(vals[r], o + 32 * r);
}

// FastLanes RLE decode (run values staged per vector in LDS), transposed order.
__global__ __launch_bounds__(kBlock) void rle_decode_kernel(const uint32_t* __restrict words,
                                                            const fls_vmeta* __restrict vmeta,
                                                            const fls_rle_meta* __restrict rmeta,
                                                            const uint32_t* __restrict bases, uint64_t n_vec,
                                                            uint32_t* __restrict out) {
    __shared__ uint32_t stage[kVecPerBlock][1024];
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const bool live = v < n_vec;
    const ColDev c{FLS_ENC_RLE, 0, 0, 1, words, vmeta, bases, nullptr};
    uint32_t x[32];
    decode_rle(c, rmeta, v, stage[threadIdx.x >> 5], live, x);
    if (!live) return;
    uint32_t* o = out + v * 1024u + dev::lane32();
#pragma unroll
    for (int r = 0; r < 32; ++r) __builtin_nontemporal_store(x[r], o + 32 * r);
}

template <bool WRITE, bool CHECKSUM>
struct UnpackTable {
    using Fn = void (*)(const uint32_t*, uint32_t*, uint64_t, uint64_t, uint64_t*);
    template <int... B>
    static constexpr auto make(std::integer_sequence<int, B...>) {
        return std::array<Fn, sizeof...(B)>{&unpack_global_kernel<B, WRITE, CHECKSUM>...};
    }
    static constexpr auto table = make(std::make_integer_sequence<int, 33>{});
};

template <bool WRITE, bool CHECKSUM>
int launch_unpack(const uint32_t* in, uint32_t* out, uint64_t n_vec, uint64_t n_values, uint32_t bw,
                  uint64_t* checksum, hipStream_t stream) {
    if (n_vec == 0) return FLS_OK;
    const uint64_t blocks = ceil_div(n_vec, kVecPerBlock);
    FLS_RETURN_IF(blocks > 0x7fffffffu, FLS_ERR_SIZE);
    auto fn = UnpackTable<WRITE, CHECKSUM>::table[bw];
    hipLaunchKernelGGL(fn, dim3((uint32_t)blocks), dim3(kBlock), 0, stream, in, out, n_vec, n_values, checksum);
    return check_launch();
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_unpack_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n_vec, uint32_t bw, void* stream) {
    FLS_RETURN_IF(bw > 32, FLS_ERR_BW);
    FLS_RETURN_IF(n_vec && (!d_out || (bw && !d_in)), FLS_ERR_NULL);
    return fls::launch_unpack<true, false>(d_in, d_out, n_vec, n_vec * 1024u, bw, nullptr, fls::as_stream(stream));
}

int fls_unpack_u32_checksum(const uint32_t* d_in, uint32_t* d_out, uint64_t n_values, uint32_t bw,
                            uint64_t* d_checksum, void* stream) {
    FLS_RETURN_IF(bw > 32, FLS_ERR_BW);
    FLS_RETURN_IF(!d_checksum, FLS_ERR_NULL);
    FLS_RETURN_IF(n_values && bw && !d_in, FLS_ERR_NULL);
    hipStream_t s = fls::as_stream(stream);
    const uint64_t n_vec = fls::ceil_div(n_values, 1024);
    if (n_vec == 0) return fls::zero_u64(d_checksum, 1, s);  // the empty sum (the reduction state is untouched)
    if (d_out) return fls::launch_unpack<true, true>(d_in, d_out, n_vec, n_values, bw, d_checksum, s);
    return fls::launch_unpack<false, true>(d_in, nullptr, n_vec, n_values, bw, d_checksum, s);
}

int fls_delta_decode_u32(const uint32_t* d_words, const fls_vmeta* d_vmeta, const uint32_t* d_bases,
                         uint64_t n_vec, uint32_t* d_out, void* stream) {
    if (n_vec == 0) return FLS_OK;
    FLS_RETURN_IF(!d_words || !d_vmeta || !d_bases || !d_out, FLS_ERR_NULL);
    const uint64_t blocks = fls::ceil_div(n_vec, fls::kVecPerBlock);
    FLS_RETURN_IF(blocks > 0x7fffffffu, FLS_ERR_SIZE);
    hipLaunchKernelGGL(fls::delta_decode_kernel, dim3((uint32_t)blocks), dim3(fls::kBlock), 0,
                       fls::as_stream(stream), d_words, d_vmeta, d_bases, n_vec, d_out);
    return fls::check_launch();
}

int fls_rle_decode_u32(const uint32_t* d_words, const fls_vmeta* d_vmeta, const fls_rle_meta* d_rmeta,
                       const uint32_t* d_bases, uint64_t n_vec, uint32_t* d_out, void* stream) {
    if (n_vec == 0) return FLS_OK;
    FLS_RETURN_IF(!d_words || !d_vmeta || !d_rmeta || !d_bases || !d_out, FLS_ERR_NULL);
    const uint64_t blocks = fls::ceil_div(n_vec, fls::kVecPerBlock);
    FLS_RETURN_IF(blocks > 0x7fffffffu, FLS_ERR_SIZE);
    hipLaunchKernelGGL(fls::rle_decode_kernel, dim3((uint32_t)blocks), dim3(fls::kBlock), 0, fls::as_stream(stream),
                       d_words, d_vmeta, d_rmeta, d_bases, n_vec, d_out);
    return fls::check_launch();
}

}  // extern "C"
