// GPU encoders: FastLanes bit-packing (plain or transposed), FastLanes delta
// and FastLanes RLE, each as plan -> scan -> pack over device buffers.
//
// Work-item t of a 256-work-item block handles lane t%32 of vector
// blockIdx.x*8 + t/32, like the decoders.  Lane l's FastLanes delta chain
// visits the 32 consecutive original rows starting at 32*(2*(l%16) + l/16)
// (transpose_i maps that run to positions l + 32*rsum_row(j), rsum.cuh:8-106),
// so delta and RLE planning read contiguous rows and need no transposition
// buffer.
#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_encode.hpp"
#include "fls_pack.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

constexpr int kBlock = 256;
constexpr int kVecPerBlock = kBlock / 32;

__device__ __forceinline__ uint32_t chain_first_row(uint32_t l) { return 32u * (2u * (l & 15u) + (l >> 4)); }

__device__ __forceinline__ uint32_t width_of(uint32_t x) { return x ? 32u - __clz(x) : 0u; }

// OR / MIN / MAX over the 32 lanes of a half-wave
__device__ __forceinline__ uint32_t half_or(uint32_t x) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) x |= __shfl_xor(x, off, 64);
    return x;
}
__device__ __forceinline__ uint32_t half_min(uint32_t x) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) x = min(x, (uint32_t)__shfl_xor(x, off, 64));
    return x;
}
__device__ __forceinline__ uint32_t half_max(uint32_t x) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) x = max(x, (uint32_t)__shfl_xor(x, off, 64));
    return x;
}

// ------------------------------------------------------------------ bitpack
template <int BW>
__global__ __launch_bounds__(kBlock) void pack_kernel(const uint32_t* __restrict in, uint64_t n, uint32_t transposed,
                                                       uint32_t base, uint32_t* __restrict out) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    if (v * 1024u >= n) return;
    const uint32_t l = dev::lane32();
    uint32_t x[32];
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        const uint32_t p = l + 32u * r;
        const uint64_t i = v * 1024u + (transposed ? dev::untranspose_index(p) : p);
        x[r] = i < n ? in[i] - base : 0u;
    }
    pack_lane<BW>(x, out + v * (32u * BW));
}

// Value of original row k of the column, padding past n with the last value.
__device__ __forceinline__ uint32_t padded(const uint32_t* __restrict in, uint64_t n, uint64_t k) {
    return in[k < n ? k : n - 1];
}

// ------------------------------------------------------------------- delta
// Load lane l's chain (32 consecutive rows) and return it in row order (the
// row of chain step j is rsum_row(j)); also its differences, chain start 0.
__device__ __forceinline__ void chain_deltas(const uint32_t* __restrict in, uint64_t n, uint64_t v,
                                             uint32_t (&d)[32], uint32_t& base) {
    const uint64_t c0 = v * 1024u + chain_first_row(dev::lane32());
    uint32_t prev = padded(in, n, c0);
    base = prev;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const uint32_t x = j ? padded(in, n, c0 + j) : prev;
        d[dev::rsum_row(j)] = x - prev;
        prev = x;
    }
}

__global__ __launch_bounds__(kBlock) void delta_plan_kernel(const uint32_t* __restrict in, uint64_t n,
                                                            uint64_t n_vec, uint32_t* __restrict bws,
                                                            uint32_t* __restrict bases) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    if (v >= n_vec) return;
    uint32_t d[32], base;
    chain_deltas(in, n, v, d, base);
    uint32_t acc = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) acc |= d[r];
    const uint32_t bw = width_of(half_or(acc));
    bases[v * 32u + dev::lane32()] = base;
    if (dev::lane32() == 0) bws[v] = bw;
}

__global__ __launch_bounds__(kBlock) void delta_pack_kernel(const uint32_t* __restrict in, uint64_t n,
                                                            uint64_t n_vec, const fls_vmeta* __restrict vmeta,
                                                            uint32_t* __restrict out) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    if (v >= n_vec) return;
    uint32_t d[32], base;
    chain_deltas(in, n, v, d, base);
    const fls_vmeta m = vmeta[v];
    pack_rows(d, out + FLS_VMETA_OFFSET(m), FLS_VMETA_BW(m));
}

// --------------------------------------------------------------------- RLE
// Per vector: run starts in original order, run index per row, compacted run
// values.  Done per half-wave with LDS for the cross-chain prefix and the
// run-value table.
struct RlePlanLds {
    uint32_t cnt[kVecPerBlock][32];     // starts per chain slot
    uint32_t vals[kVecPerBlock][1024];  // run values by run index
};

// Phase A (no barrier): lane's chain values x[j] (chain order), start flags
// as the run-index deltas in row order d[rsum_row(j)] (chain start included
// for now), and the chain's start count published to lds.cnt.
__device__ __forceinline__ void rle_chain_local(const uint32_t* __restrict in, uint64_t n, uint64_t v,
                                                RlePlanLds& lds, int slot, uint32_t (&d)[32], uint32_t (&x)[32]) {
    const uint32_t l = dev::lane32();
    const uint32_t c = 2u * (l & 15u) + (l >> 4);  // which run of 32 rows this lane's chain covers
    const uint64_t c0 = v * 1024u + 32u * c;
    uint32_t prev = (c == 0) ? 0u : padded(in, n, c0 - 1);
    uint32_t starts = 0;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const uint32_t xv = padded(in, n, c0 + j);
        const uint32_t st = (c == 0 && j == 0) ? 1u : (xv != prev ? 1u : 0u);
        x[j] = xv;
        d[dev::rsum_row(j)] = st;
        starts += st;
        prev = xv;
    }
    lds.cnt[slot][c] = starts;
}

// Phase B (after a block barrier): run index of the chain's first row and the
// vector's run count from the per-chain start counts; the chain start's delta
// is set to 0 (unrsum convention).
__device__ __forceinline__ void rle_chain_prefix(const RlePlanLds& lds, int slot, uint32_t (&d)[32],
                                                 uint32_t& first_idx, uint32_t& n_runs) {
    const uint32_t l = dev::lane32();
    const uint32_t c = 2u * (l & 15u) + (l >> 4);
    uint32_t before = 0, total = 0;
    for (uint32_t k = 0; k < 32; ++k) {
        const uint32_t cnt = lds.cnt[slot][k];
        before += k < c ? cnt : 0u;
        total += cnt;
    }
    n_runs = total;
    first_idx = before + d[0] - 1u;  // d[0] is the chain start's flag (rsum_row(0) == 0)
    d[0] = 0u;
}

__global__ __launch_bounds__(kBlock) void rle_plan_kernel(const uint32_t* __restrict in, uint64_t n, uint64_t n_vec,
                                                          uint32_t* __restrict words, fls_rle_meta* __restrict rmeta,
                                                          uint32_t* __restrict bases, uint32_t* __restrict ibws) {
    __shared__ RlePlanLds lds;
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const int slot = threadIdx.x >> 5;
    const bool live = v < n_vec;
    uint32_t d[32], x[32], first_idx = 0, n_runs = 0;
    if (live) rle_chain_local(in, n, v, lds, slot, d, x);
    __syncthreads();  // every work-item passes exactly this one barrier
    if (!live) return;
    rle_chain_prefix(lds, slot, d, first_idx, n_runs);
    uint32_t vmin = 0xffffffffu, vmax = 0u, dmax = 0u;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const uint32_t st = j ? d[dev::rsum_row(j)] : 0u;
        dmax |= st;
        if (j == 0 || st) {  // every run's value is seen at its start or at a chain start inside it
            vmin = min(vmin, x[j]);
            vmax = max(vmax, x[j]);
        }
    }
    vmin = half_min(vmin);
    vmax = half_max(vmax);
    const uint32_t ibw = width_of(half_or(dmax));
    const uint32_t vbw = width_of(vmax - vmin);
    bases[v * 32u + dev::lane32()] = first_idx;
    if (dev::lane32() == 0) {
        rmeta[v] = fls_rle_meta{vmin, (uint16_t)n_runs, (uint8_t)vbw, (uint8_t)ibw};
        words[v] = 32u * ibw + (n_runs * vbw + 31u) / 32u;
        ibws[v] = ibw;
    }
}

__global__ __launch_bounds__(kBlock) void rle_pack_kernel(const uint32_t* __restrict in, uint64_t n, uint64_t n_vec,
                                                          const fls_vmeta* __restrict vmeta,
                                                          const fls_rle_meta* __restrict rmeta,
                                                          uint32_t* __restrict out) {
    __shared__ RlePlanLds lds;
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const int slot = threadIdx.x >> 5;
    const bool live = v < n_vec;
    uint32_t d[32], x[32], first_idx = 0, n_runs = 0;
    if (live) rle_chain_local(in, n, v, lds, slot, d, x);
    __syncthreads();  // barrier 1 of 2, unconditional
    if (live) {
        rle_chain_prefix(lds, slot, d, first_idx, n_runs);
        // run values into LDS by run index; a chain's first row writes its
        // run's value too, which equals what the run's start row writes
        uint32_t idx = first_idx;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const uint32_t st = j ? d[dev::rsum_row(j)] : 0u;
            idx += st;
            if (j == 0 || st) lds.vals[slot][idx] = x[j];
        }
    }
    __syncthreads();  // barrier 2 of 2, unconditional
    if (!live) return;
    const fls_vmeta m = vmeta[v];
    const fls_rle_meta rm = rmeta[v];
    uint32_t* vec_out = out + FLS_VMETA_OFFSET(m);
    const uint32_t ibw = FLS_VMETA_BW(m);
    pack_rows(d, vec_out, ibw);
    // linear LSB-first packing of (value - vbase) at vbw bits after the index
    // words; lane l builds words l, l+32, ...
    uint32_t* vals_out = vec_out + 32u * ibw;
    const uint32_t vbw = rm.vbw, nr = rm.n_runs;
    const uint32_t n_words = (nr * vbw + 31u) / 32u;
    for (uint32_t w = dev::lane32(); w < n_words; w += 32u) {
        uint32_t word = 0;
        const uint32_t bit0 = 32u * w;
        const uint32_t k0 = bit0 / vbw, k1 = min(nr - 1u, (bit0 + 31u) / vbw);
        for (uint32_t k = k0; k <= k1; ++k) {
            const uint32_t kb = k * vbw;  // first bit of value k
            const uint32_t val = lds.vals[slot][k] - rm.vbase;
            if (kb >= bit0) {
                if (kb - bit0 < 32u) word |= val << (kb - bit0);
            } else {
                word |= val >> (bit0 - kb);
            }
        }
        vals_out[w] = word;
    }
}

// -------------------------------------------------------------------- scan
// offsets[i] = sum_{k<i} words[i]; vmeta[i] = offsets[i] << 8 | bw[i];
// *total = sum of all.  Three phases over blocks of 1024 entries.
constexpr int kScanBlock = 1024;

constexpr int kScanWaves = kScanBlock / 64;

// Exclusive scan of one 32-bit value per work-item over the block: a DPP
// prefix sum per wave, then the 16 wave totals read by every work-item (two
// barriers; a block's values here are vector word counts, at most about 2^20
// in total).  sh: kScanWaves words of LDS.
__device__ __forceinline__ uint32_t block_exclusive_scan32(uint32_t x, uint32_t* sh, uint32_t& total) {
    const uint32_t incl = dev::wave_inclusive_sum(x);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63u) == 63u) sh[w] = incl;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kScanWaves; ++k) {
        const uint32_t t = sh[k];
        before += k < w ? t : 0u;
        tot += t;
    }
    total = tot;
    __syncthreads();  // sh is written again by the next scan
    return before + incl - x;
}

// The same for 64-bit values (the block sums): a shuffle prefix sum per wave.
__device__ __forceinline__ uint64_t block_exclusive_scan64(uint64_t x, uint64_t* sh, uint64_t& total) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint64_t y = __shfl_up(incl, off, 64);
        if (l >= off) incl += y;
    }
    if (l == 63) sh[w] = incl;
    __syncthreads();
    uint64_t before = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kScanWaves; ++k) {
        const uint64_t t = sh[k];
        before += k < w ? t : 0u;
        tot += t;
    }
    total = tot;
    __syncthreads();
    return before + incl - x;
}

__global__ __launch_bounds__(kScanBlock) void scan_reduce_kernel(const uint32_t* __restrict words, uint64_t n,
                                                                  uint64_t* __restrict block_sums) {
    __shared__ uint32_t sh[kScanWaves];
    const uint64_t i = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t total;
    block_exclusive_scan32(i < n ? words[i] : 0u, sh, total);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void scan_blocks_kernel(uint64_t* __restrict block_sums, uint64_t n_blocks,
                                                                  uint64_t* __restrict total_out) {
    __shared__ uint64_t sh[kScanWaves];
    uint64_t carry = 0;
    for (uint64_t b0 = 0; b0 < n_blocks; b0 += kScanBlock) {
        const uint64_t i = b0 + threadIdx.x;
        const uint64_t x = i < n_blocks ? block_sums[i] : 0u;
        uint64_t total;
        const uint64_t ex = block_exclusive_scan64(x, sh, total);
        if (i < n_blocks) block_sums[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanBlock) void scan_apply_kernel(const uint32_t* __restrict words,
                                                                 const uint32_t* __restrict bws, uint64_t n,
                                                                 const uint64_t* __restrict block_sums,
                                                                 fls_vmeta* __restrict vmeta) {
    __shared__ uint32_t sh[kScanWaves];
    const uint64_t i = (uint64_t)blockIdx.x * kScanBlock + threadIdx.x;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan32(i < n ? words[i] : 0u, sh, total);
    if (i < n) vmeta[i] = FLS_VMETA(block_sums[blockIdx.x] + ex, bws[i]);
}

// words[i] = 32 * bws[i] (delta)
__global__ void delta_words_kernel(const uint32_t* __restrict bws, uint64_t n, uint32_t* __restrict words) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        words[i] = 32u * bws[i];
}

template <bool ALLOWED>
struct PackTable {
    using Fn = void (*)(const uint32_t*, uint64_t, uint32_t, uint32_t, uint32_t*);
    template <int... B>
    static constexpr auto make(std::integer_sequence<int, B...>) {
        return std::array<Fn, sizeof...(B)>{&pack_kernel<B>...};
    }
    static constexpr auto table = make(std::make_integer_sequence<int, 33>{});
};

}  // namespace

uint64_t scratch_words(uint64_t n_vec) {
    // bws[n_vec] + words[n_vec] (u32) + block sums (u64), 256-byte aligned
    const uint64_t nb = ceil_div(n_vec ? n_vec : 1, kScanBlock);
    return ((2 * n_vec * 4 + 255) / 256 * 256 + nb * 8 + 255) / 4;
}

int scan_offsets(const uint32_t* words, const uint32_t* bws, uint64_t n_vec, uint64_t* block_sums, fls_vmeta* vmeta,
                 uint64_t* d_total, hipStream_t s) {
    const uint64_t nb = ceil_div(n_vec, kScanBlock);
    FLS_RETURN_IF(nb > 0x7fffffffu, FLS_ERR_SIZE);
    hipLaunchKernelGGL(scan_reduce_kernel, dim3((uint32_t)nb), dim3(kScanBlock), 0, s, words, n_vec, block_sums);
    hipLaunchKernelGGL(scan_blocks_kernel, dim3(1), dim3(kScanBlock), 0, s, block_sums, nb, d_total);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((uint32_t)nb), dim3(kScanBlock), 0, s, words, bws, n_vec, block_sums,
                       vmeta);
    return check_launch();
}

}  // namespace fls

extern "C" {

uint64_t fls_encode_scratch_bytes(uint64_t n) { return 4u * fls::scratch_words(fls::ceil_div(n, 1024)); }

int fls_pack_column_device(const uint32_t* d_in, uint64_t n, uint32_t bw, uint32_t base, uint32_t flags,
                           uint32_t* d_out, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(bw > 32, FLS_ERR_BW);
    if (n == 0 || bw == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_out, FLS_ERR_NULL);
    const uint64_t blocks = ceil_div(ceil_div(n, 1024), kVecPerBlock);
    FLS_RETURN_IF(blocks > 0x7fffffffu, FLS_ERR_SIZE);
    hipLaunchKernelGGL(PackTable<true>::table[bw], dim3((uint32_t)blocks), dim3(kBlock), 0, as_stream(stream), d_in,
                       n, (uint32_t)((flags & FLS_ENC_FLAG_TRANSPOSE) != 0), base, d_out);
    return check_launch();
}

int fls_delta_plan_device(const uint32_t* d_in, uint64_t n, fls_vmeta* d_vmeta, uint32_t* d_bases,
                          uint64_t* d_total_words, void* d_scratch, void* stream) {
    using namespace fls;
    if (n == 0) return zero_u64(d_total_words, 1, as_stream(stream));
    FLS_RETURN_IF(!d_in || !d_vmeta || !d_bases || !d_total_words || !d_scratch, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    uint32_t* bws = static_cast<uint32_t*>(d_scratch);
    uint32_t* words = bws + n_vec;
    uint64_t* bsum = reinterpret_cast<uint64_t*>(static_cast<char*>(d_scratch) + (2 * n_vec * 4 + 255) / 256 * 256);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(delta_plan_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0, s, d_in, n,
                       n_vec, bws, d_bases);
    hipLaunchKernelGGL(delta_words_kernel, dim3(256), dim3(256), 0, s, bws, n_vec, words);
    return scan_offsets(words, bws, n_vec, bsum, d_vmeta, d_total_words, s);
}

int fls_delta_pack_device(const uint32_t* d_in, uint64_t n, const fls_vmeta* d_vmeta, uint32_t* d_out,
                          void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_vmeta || !d_out, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    hipLaunchKernelGGL(delta_pack_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0,
                       as_stream(stream), d_in, n, n_vec, d_vmeta, d_out);
    return check_launch();
}

int fls_rle_plan_device(const uint32_t* d_in, uint64_t n, fls_vmeta* d_vmeta, fls_rle_meta* d_rmeta,
                        uint32_t* d_bases, uint64_t* d_total_words, void* d_scratch, void* stream) {
    using namespace fls;
    if (n == 0) return zero_u64(d_total_words, 1, as_stream(stream));
    FLS_RETURN_IF(!d_in || !d_vmeta || !d_rmeta || !d_bases || !d_total_words || !d_scratch, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    uint32_t* ibws = static_cast<uint32_t*>(d_scratch);
    uint32_t* words = ibws + n_vec;
    uint64_t* bsum = reinterpret_cast<uint64_t*>(static_cast<char*>(d_scratch) + (2 * n_vec * 4 + 255) / 256 * 256);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(rle_plan_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0, s, d_in, n,
                       n_vec, words, d_rmeta, d_bases, ibws);
    return scan_offsets(words, ibws, n_vec, bsum, d_vmeta, d_total_words, s);
}

int fls_rle_pack_device(const uint32_t* d_in, uint64_t n, const fls_vmeta* d_vmeta, const fls_rle_meta* d_rmeta,
                        uint32_t* d_out, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_vmeta || !d_rmeta || !d_out, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    hipLaunchKernelGGL(rle_pack_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0,
                       as_stream(stream), d_in, n, n_vec, d_vmeta, d_rmeta, d_out);
    return check_launch();
}

}  // extern "C"
