// MI355X decoders of the tile-based competitor formats GPU-FOR, GPU-DFOR and
// GPU-RFOR (Shanbhag et al.; the reference's LoadBinPack / LoadDBinPack /
// LoadRBinPack, tile_based/src/include/*.cuh, run 128-thread blocks with CUB
// block scans).  Here one wave64 decodes one 512-value tile from the column
// (no staging copy: the tile's words are read through the cache), the prefix
// sums of DFOR and RFOR are wave scans, and the RLE expansion marks run starts
// in a per-wave LDS slice.  Each decoder optionally writes the values and/or
// adds their 64-bit sum (one atomic per wave at the end).
#include "fls/device/wave.hpp"
#include "fls_runtime.hpp"
#include "fls_columns.hpp"

namespace fls {
namespace {

constexpr int kTileBlock = 256;  // 4 waves, one tile each
constexpr int kTileWaves = kTileBlock / 64;
constexpr uint32_t kTile = 512;

__device__ __forceinline__ uint32_t lane64() { return threadIdx.x & 63u; }

// Value idx of a linear LSB-first run at width bw (bw <= 32) starting at w.
// The second word is read unconditionally (the same word when not needed), so
// no read goes past the run's last word.
__device__ __forceinline__ uint32_t run_value(const uint32_t* __restrict w, uint32_t idx, uint32_t bw) {
    const uint32_t bit = idx * bw, i = bit >> 5, s = bit & 31u;
    const uint32_t lo = w[i];
    const uint32_t hi = w[(s + bw > 32u) ? i + 1u : i];
    const uint32_t mask = bw >= 32u ? 0xffffffffu : ((1u << bw) - 1u);
    return bw ? (__builtin_amdgcn_alignbit(hi, lo, s) & mask) : 0u;
}

// Element e (< 128) of a FOR block {ref, widths, miniblocks}: miniblock m
// starts at the sum of the widths before it, the byte-wise prefix sums of the
// widths word (each width <= 32, so no byte overflows).
__device__ __forceinline__ uint32_t for_element(const uint32_t* __restrict blk, uint32_t e) {
    const uint32_t m = e >> 5, widths = blk[1];
    const uint32_t bw = (widths >> (8u * m)) & 255u;
    const uint32_t off = ((widths * 0x01010100u) >> (8u * m)) & 255u;
    return blk[0] + run_value(blk + 2u + off, e & 31u, bw);
}

// Inclusive scan over the wave (64 lanes).
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    const uint32_t l = lane64();
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (l >= (uint32_t)d) x += y;
    }
    return x;
}

__device__ __forceinline__ void wave_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void add_sum(uint64_t* sum, uint64_t s) {
    s = dev::wave_sum(s);
    if (lane64() == 0 && s) atomicAdd(reinterpret_cast<unsigned long long*>(sum), (unsigned long long)s);
}

// GPU-FOR: lane l decodes elements l + 64 k (k < 8) of its tile, so each
// step's block (k / 2) is the same across the wave and the stores coalesce.
template <bool OUT, bool SUM>
__global__ __launch_bounds__(kTileBlock) void tile_for_kernel(const uint32_t* __restrict data,
                                                              const uint32_t* __restrict starts, uint64_t n,
                                                              uint32_t* __restrict out, uint64_t* sum) {
    const uint64_t n_tiles = (n + kTile - 1) / kTile, n_blocks = (n + 127) / 128;
    const uint32_t l = lane64();
    uint64_t acc = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kTileWaves + (threadIdx.x >> 6); t < n_tiles;
         t += (uint64_t)gridDim.x * kTileWaves) {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint64_t b = 4 * t + (k >> 1);
            if (b >= n_blocks) break;  // wave-uniform
            const uint64_t e = kTile * t + l + 64u * k;
            const uint32_t v = for_element(data + starts[b], l + 64u * (k & 1));
            if (e < n) {
                if constexpr (OUT) out[e] = v;
                if constexpr (SUM) acc += v;
            }
        }
    }
    if constexpr (SUM) add_sum(sum, acc);
}

// GPU-DFOR: lane l decodes the 8 consecutive deltas 8l .. 8l+7 of its tile
// (one miniblock), element 0 being the tile's first value; a lane prefix plus
// a wave scan give the values.
template <bool OUT, bool SUM>
__global__ __launch_bounds__(kTileBlock) void tile_dfor_kernel(const uint32_t* __restrict data,
                                                               const uint32_t* __restrict starts, uint64_t n,
                                                               uint32_t* __restrict out, uint64_t* sum) {
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    const uint32_t l = lane64();
    uint64_t acc = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kTileWaves + (threadIdx.x >> 6); t < n_tiles;
         t += (uint64_t)gridDim.x * kTileWaves) {
        const uint64_t base = kTile * t;
        const uint32_t live = (uint32_t)min<uint64_t>(kTile, n - base);
        const uint32_t* blk = data + starts[4 * t + (l >> 4)];  // tiles always hold 4 blocks
        const uint32_t e0 = 8u * (l & 15u);
        uint32_t x[8], run = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            x[i] = for_element(blk, e0 + i);
            if (l == 0 && i == 0) x[0] = data[starts[4 * t] - 1u];
            run += x[i];
            x[i] = run;
        }
        const uint32_t before = wave_inclusive_scan(run) - run;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t p = 8u * l + i;
            const uint32_t v = x[i] + before;
            if (p < live) {
                if constexpr (OUT) out[base + p] = v;
                if constexpr (SUM) acc += v;
            }
        }
    }
    if constexpr (SUM) add_sum(sum, acc);
}

// GPU-RFOR: lane l decodes runs 8l .. 8l+7 of its tile (value, length); a
// wave scan of the lengths gives each run's first position.  With OUT the
// run values and run-start marks go to the wave's LDS slice and a scan of
// the marks over positions 8l .. 8l+7 gives each position's run.  The sum is
// value x length over the runs, clipped to the column's end.
template <bool OUT, bool SUM>
__global__ __launch_bounds__(kTileBlock) void tile_rfor_kernel(const uint32_t* __restrict values,
                                                               const uint32_t* __restrict val_starts,
                                                               const uint32_t* __restrict run_lengths,
                                                               const uint32_t* __restrict rl_starts, uint64_t n,
                                                               uint32_t* __restrict out, uint64_t* sum) {
    __shared__ uint32_t lds[kTileWaves][2 * kTile];
    uint32_t* runval = lds[threadIdx.x >> 6];
    uint32_t* mark = runval + kTile;
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    const uint32_t l = lane64();
    uint64_t acc = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kTileWaves + (threadIdx.x >> 6); t < n_tiles;
         t += (uint64_t)gridDim.x * kTileWaves) {
        const uint64_t base = kTile * t;
        const uint32_t live = (uint32_t)min<uint64_t>(kTile, n - base);
        const uint32_t* vb = values + val_starts[t];
        const uint32_t* rb = run_lengths + rl_starts[t];
        const uint32_t count = min(vb[2], kTile), vbw = vb[1] & 255u, rbw = rb[1] & 255u;
        uint32_t v[8], len[8], run = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t r = 8u * l + i;
            v[i] = r < count ? vb[0] + run_value(vb + 3, r, vbw) : 0u;
            len[i] = r < count ? rb[0] + run_value(rb + 3, r, rbw) : 0u;
            run += len[i];
        }
        const uint32_t first = wave_inclusive_scan(run) - run;  // position of run 8l
        if constexpr (SUM) {
            uint32_t p = first;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t end = min(p + len[i], live);
                if (end > p) acc += (uint64_t)v[i] * (end - p);
                p += len[i];
            }
        }
        if constexpr (OUT) {
#pragma unroll
            for (int i = 0; i < 8; ++i) mark[8u * l + i] = 0u;
            wave_fence();
            uint32_t p = first;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t r = 8u * l + i;
                if (r < count) {
                    runval[r] = v[i];
                    if (p < kTile) mark[p] = 1u;  // lengths past the tile end are clipped
                }
                p += len[i];
            }
            wave_fence();
            uint32_t idx[8], c = 0;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                c += mark[8u * l + i];
                idx[i] = c;
            }
            const uint32_t runs_before = wave_inclusive_scan(c) - c;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const uint32_t p8 = 8u * l + i;
                const uint32_t r = idx[i] + runs_before;  // runs started at or before p8
                if (p8 < live) out[base + p8] = runval[r ? min(r, count) - 1u : 0u];
            }
            wave_fence();
        }
    }
    if constexpr (SUM) add_sum(sum, acc);
}

template <typename Kernel, typename... Args>
int launch_tiles(Kernel k, uint64_t n, hipStream_t s, Args... args) {
    uint32_t grid = 0;
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    if (const int st = resident_grid(k, kTileBlock, (n_tiles + kTileWaves - 1) / kTileWaves, &grid)) return st;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kTileBlock), 0, s, args...);
    return check_launch();
}

int prepare(uint64_t n, uint32_t* d_out, uint64_t* d_sum, hipStream_t s) {
    if (!d_out && !d_sum) return FLS_ERR_NULL;
    if (d_sum) return zero_u64(d_sum, 1, s);
    return FLS_OK;
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_tile_for_decode_u32(const uint32_t* d_data, const uint32_t* d_starts, uint64_t n, uint32_t* d_out,
                            uint64_t* d_sum, void* stream) {
    using namespace fls;
    hipStream_t s = as_stream(stream);
    if (const int st = prepare(n, d_out, d_sum, s)) return st;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_data || !d_starts, FLS_ERR_NULL);
    if (d_out && d_sum) return launch_tiles(tile_for_kernel<true, true>, n, s, d_data, d_starts, n, d_out, d_sum);
    if (d_out) return launch_tiles(tile_for_kernel<true, false>, n, s, d_data, d_starts, n, d_out, d_sum);
    return launch_tiles(tile_for_kernel<false, true>, n, s, d_data, d_starts, n, d_out, d_sum);
}

int fls_tile_dfor_decode_u32(const uint32_t* d_data, const uint32_t* d_starts, uint64_t n, uint32_t* d_out,
                             uint64_t* d_sum, void* stream) {
    using namespace fls;
    hipStream_t s = as_stream(stream);
    if (const int st = prepare(n, d_out, d_sum, s)) return st;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_data || !d_starts, FLS_ERR_NULL);
    if (d_out && d_sum) return launch_tiles(tile_dfor_kernel<true, true>, n, s, d_data, d_starts, n, d_out, d_sum);
    if (d_out) return launch_tiles(tile_dfor_kernel<true, false>, n, s, d_data, d_starts, n, d_out, d_sum);
    return launch_tiles(tile_dfor_kernel<false, true>, n, s, d_data, d_starts, n, d_out, d_sum);
}

int fls_tile_rfor_decode_u32(const uint32_t* d_values, const uint32_t* d_val_starts, const uint32_t* d_run_lengths,
                             const uint32_t* d_rl_starts, uint64_t n, uint32_t* d_out, uint64_t* d_sum,
                             void* stream) {
    using namespace fls;
    hipStream_t s = as_stream(stream);
    if (const int st = prepare(n, d_out, d_sum, s)) return st;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_values || !d_val_starts || !d_run_lengths || !d_rl_starts, FLS_ERR_NULL);
    if (d_out && d_sum)
        return launch_tiles(tile_rfor_kernel<true, true>, n, s, d_values, d_val_starts, d_run_lengths, d_rl_starts,
                            n, d_out, d_sum);
    if (d_out)
        return launch_tiles(tile_rfor_kernel<true, false>, n, s, d_values, d_val_starts, d_run_lengths, d_rl_starts,
                            n, d_out, d_sum);
    return launch_tiles(tile_rfor_kernel<false, true>, n, s, d_values, d_val_starts, d_run_lengths, d_rl_starts, n,
                        d_out, d_sum);
}

}  // extern "C"
