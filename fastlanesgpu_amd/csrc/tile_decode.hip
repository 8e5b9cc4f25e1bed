// MI355X decoders of the tile-based competitor formats GPU-FOR, GPU-DFOR and
// GPU-RFOR (Shanbhag et al.; the reference's LoadBinPack / LoadDBinPack /
// LoadRBinPack, tile_based/src/include/*.cuh, run 128-thread blocks with CUB
// block scans).  Here one wave64 decodes one 512-value tile: the tile's words
// are copied into the wave's LDS slice with coalesced loads (one memory round
// trip per tile), elements are decoded from LDS, the prefix sums of DFOR and
// RFOR are wave scans, and the RLE expansion marks run starts in LDS.  Each
// decoder optionally writes the values and/or adds their 64-bit sum (one
// atomic per wave at the end).
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

// Words of a tile: 4 FOR blocks of at most 2 + 4 x 32 words (DFOR: plus the
// tile's first value); RFOR: a run-value and a run-length block of at most
// 3 + 512 words each.
constexpr uint32_t kForTileWords = 4u * (2u + 4u * 32u) + 1u;
constexpr uint32_t kRforBlockWords = 3u + kTile;

// Copy the first min(count, CAP) words of src into st, coalesced: all of a
// lane's loads are issued before any is stored.
template <uint32_t CAP>
__device__ __forceinline__ void stage_words(uint32_t* st, const uint32_t* __restrict src, uint32_t count) {
    constexpr int kSteps = (CAP + 63) / 64;
    count = min(count, CAP);
    const uint32_t l = lane64();
    uint32_t w[kSteps];
#pragma unroll
    for (int k = 0; k < kSteps; ++k) {
        const uint32_t i = l + 64u * k;
        w[k] = i < count ? src[i] : 0u;
    }
#pragma unroll
    for (int k = 0; k < kSteps; ++k) {
        const uint32_t i = l + 64u * k;
        if (i < count) st[i] = w[k];
    }
}

// GPU-FOR: lane l decodes elements l + 64 k (k < 8) of its tile, so each
// step's block (k / 2) is the same across the wave and the stores coalesce.
template <bool OUT, bool SUM>
__global__ __launch_bounds__(kTileBlock) void tile_for_kernel(const uint32_t* __restrict data,
                                                              const uint32_t* __restrict starts, uint64_t n,
                                                              uint32_t* __restrict out, uint64_t* sum) {
    __shared__ uint32_t lds[kTileWaves][kForTileWords];
    uint32_t* st = lds[threadIdx.x >> 6];
    const uint64_t n_tiles = (n + kTile - 1) / kTile, n_blocks = (n + 127) / 128;
    const uint32_t l = lane64();
    uint64_t acc = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kTileWaves + (threadIdx.x >> 6); t < n_tiles;
         t += (uint64_t)gridDim.x * kTileWaves) {
        const uint64_t b0 = 4 * t, nb = min<uint64_t>(4, n_blocks - b0);
        const uint32_t w0 = starts[b0];
        stage_words<kForTileWords - 1u>(st, data + w0, starts[b0 + nb] - w0);
        wave_fence();
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if ((uint64_t)(k >> 1) >= nb) break;  // wave-uniform
            const uint64_t e = kTile * t + l + 64u * k;
            const uint32_t v = for_element(st + (starts[b0 + (k >> 1)] - w0), l + 64u * (k & 1));
            if (e < n) {
                if constexpr (OUT) out[e] = v;
                if constexpr (SUM) acc += v;
            }
        }
        wave_fence();
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
    __shared__ uint32_t lds[kTileWaves][kForTileWords];
    uint32_t* st = lds[threadIdx.x >> 6];
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    const uint32_t l = lane64();
    uint64_t acc = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kTileWaves + (threadIdx.x >> 6); t < n_tiles;
         t += (uint64_t)gridDim.x * kTileWaves) {
        const uint64_t base = kTile * t;
        const uint32_t live = (uint32_t)min<uint64_t>(kTile, n - base);
        const uint32_t w0 = starts[4 * t] - 1u;  // the tile's first value, then its 4 blocks
        stage_words<kForTileWords>(st, data + w0, starts[4 * t + 4] - w0);
        wave_fence();
        const uint32_t* blk = st + (starts[4 * t + (l >> 4)] - w0);
        const uint32_t e0 = 8u * (l & 15u);
        uint32_t x[8], run = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            x[i] = for_element(blk, e0 + i);
            if (l == 0 && i == 0) x[0] = st[0];
            run += x[i];
            x[i] = run;
        }
        wave_fence();
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
    __shared__ uint32_t lds[kTileWaves][2 * kRforBlockWords];
    uint32_t* runval = lds[threadIdx.x >> 6];
    uint32_t* mark = runval + kTile;
    uint32_t* vst = lds[threadIdx.x >> 6];
    uint32_t* rst = vst + kRforBlockWords;
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    const uint32_t l = lane64();
    uint64_t acc = 0;
    for (uint64_t t = (uint64_t)blockIdx.x * kTileWaves + (threadIdx.x >> 6); t < n_tiles;
         t += (uint64_t)gridDim.x * kTileWaves) {
        const uint64_t base = kTile * t;
        const uint32_t live = (uint32_t)min<uint64_t>(kTile, n - base);
        const uint32_t vs = val_starts[t], rs = rl_starts[t];
        stage_words<kRforBlockWords>(vst, values + vs, val_starts[t + 1] - vs);
        stage_words<kRforBlockWords>(rst, run_lengths + rs, rl_starts[t + 1] - rs);
        wave_fence();
        const uint32_t count = min(vst[2], kTile), vbw = vst[1] & 255u, rbw = rst[1] & 255u;
        uint32_t v[8], len[8], run = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const uint32_t r = 8u * l + i;
            v[i] = r < count ? vst[0] + run_value(vst + 3, r, vbw) : 0u;
            len[i] = r < count ? rst[0] + run_value(rst + 3, r, rbw) : 0u;
            run += len[i];
        }
        wave_fence();  // the slice is reused for the run values and marks below
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

int prepare(uint32_t* d_out, uint64_t* d_sum, hipStream_t s) {
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
    if (const int st = prepare(d_out, d_sum, s)) return st;
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
    if (const int st = prepare(d_out, d_sum, s)) return st;
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
    if (const int st = prepare(d_out, d_sum, s)) return st;
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
