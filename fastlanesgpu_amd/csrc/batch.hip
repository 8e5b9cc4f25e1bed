// Batched decode + range filter over many compressed segments of uneven
// length, bit width and encoding, in one persistent launch.
//
// Segments are cut into tiles of `tile_vectors` FastLanes vectors (or a tile
// size per encoding and width class); a tile never spans two segments, so every block-wide branch on encoding and width
// is uniform.  A persistent grid (every block resident) walks the tile list
// either round-robin (static) or through an atomic work queue (dynamic).  Per
// tile the block counts the rows with lo <= value <= hi and adds the count to
// the segment's counter with one 64-bit atomic.  The launch (plus the kernels
// that zero the counters and the queue) can be captured once in a hipGraph
// and replayed; the baseline is one launch per segment.  An autotuner times
// tile sizes x schedule x block size on the batch and caches the winner on
// disk.
//
// No reference counterpart: every reference launch is one kernel per query
// over one column set (fls_q11.cu:137-171, the query driver).
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <string>
#include <vector>

#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_encode.hpp"
#include "fls_rng.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

struct SegDev {
    uint32_t enc, bw, transposed, base;  // transposed: rows in FastLanes transposed order; base: BITPACK FOR
    uint64_t n_values;
    const uint32_t* words;
    const fls_vmeta* vmeta;
    const uint32_t* bases;
    const fls_rle_meta* rmeta;
    const uint64_t* aux;  // MIXED: per-vector aux words
    uint32_t lo, hi;
};

struct TileDev {
    uint32_t seg, v0, nv, pad;
};

// ------------------------------------------------------------ range counts
// Every path counts the rows of a vector with lo <= value <= hi as
// (value - lo) <= span (unsigned), so one compare decides each row.
//
// Fast paths (full vectors of BITPACK / DELTA / RLE segments): each compare
// adds its carry to a per-vector counter in the work-item; a half-wave whose
// slot is past the tile's vectors discards its counter once per vector, not
// once per row.  (A wave ballot per row with a scalar popcount measured
// slower: the compiler turns the predicate into a 0/1 value and back, and the
// CU's one scalar unit serialises the counts of all four SIMDs.)  At narrow
// widths a half-wave takes K vectors per step and issues all their loads
// before any extraction (the packed words of K vectors fit the registers of
// one 32-bit vector), so a narrow segment keeps as many bytes in flight as a
// wide one.

// Vectors per half-wave per step at width BW: K * BW <= 32 words per lane.
template <int BW>
constexpr int batch_k() {
    return BW <= 4 ? 8 : BW <= 8 ? 4 : BW <= 16 ? 2 : 1;
}

// Widths that divide 32 (1, 2, 4, 8) are counted a packed word at a time
// (SWAR): the word's fields split into even and odd halves (each field then
// has BW free bits above it), and for an interval [a, b] of field values
//   field + (2^BW - a)  has bit BW set iff field >= a,
//   (2^BW + b) - field  has bit BW set iff field <= b,
// with no carry or borrow between fields; a popcount of the AND of both at
// the fields' bit BW counts the half.  About 13 operations per word of
// 32 / BW values, against about 4 per value one row at a time.
template <int BW>
struct Swar {
    static constexpr bool kOn = BW == 1 || BW == 2 || BW == 4 || BW == 8;
    static constexpr uint32_t kG = BW == 1 ? 0x55555555u : BW == 2 ? 0x11111111u : BW == 4 ? 0x01010101u : 0x00010001u;
    static constexpr uint32_t kE = kG * ((1u << (BW & 31)) - 1u);  // even fields
    static constexpr uint32_t kH = kG << (BW & 31);                // bit BW above each even field
    // count of fields of w in [a, b] (0 <= a <= b < 2^BW); ca = (2^BW - a) * kG, cb = (2^BW + b) * kG
    static __device__ __forceinline__ uint32_t count(uint32_t w, uint32_t ca, uint32_t cb) {
        const uint32_t e = w & kE, o = (w >> BW) & kE;
        return __builtin_popcount((e + ca) & (cb - e) & kH) + __builtin_popcount((o + ca) & (cb - o) & kH);
    }
};

// The values x of a BW-bit field with (x - lo_adj) mod 2^32 <= span, as at
// most two intervals [a, b] inside [0, 2^BW), in Swar's constant form.
template <int BW>
struct SwarRange {
    uint32_t n, ca[2], cb[2];
    __device__ __forceinline__ SwarRange(uint32_t lo_adj, uint32_t span) : n(0), ca{0, 0}, cb{0, 0} {
        const uint32_t m = (1u << BW) - 1u;
        const uint64_t end = (uint64_t)lo_adj + span;
        const bool first = lo_adj <= m, wraps = (end >> 32) != 0;  // wrapping adds [0, end - 2^32]
        const uint64_t e2 = end - (1ull << 32);
        // the first interval in slot 0, the wrapped part in slot n (compile-time indices only)
        const uint32_t a0 = first ? lo_adj : 0u, b0 = first ? (end < m ? (uint32_t)end : m) : (e2 < m ? (uint32_t)e2 : m);
        ca[0] = ((1u << BW) - a0) * Swar<BW>::kG;
        cb[0] = ((1u << BW) + b0) * Swar<BW>::kG;
        ca[1] = (1u << BW) * Swar<BW>::kG;  // a = 0
        cb[1] = ((1u << BW) + (e2 < m ? (uint32_t)e2 : m)) * Swar<BW>::kG;
        n = (first ? 1u : 0u) + (wraps ? 1u : 0u);
    }
};

// BITPACK (+ FOR): vectors [0, nfull) of the segment's tile range starting at
// v0; `lo_adj` = lo - base.  Returns this work-item's count.
template <int BW, int VPB>
__device__ __forceinline__ uint32_t count_fixed_full(const uint32_t* __restrict words, uint64_t v0, uint32_t nfull,
                                                     uint32_t lo_adj, uint32_t span) {
    constexpr int K = batch_k<BW>();
    const uint32_t hw = threadIdx.x >> 5, l = dev::lane32();
    if constexpr (BW == 0) {
        return 0u;  // every value is the base: the caller counts the whole range at once
    } else if constexpr (Swar<BW>::kOn) {
        const SwarRange<BW> rg(lo_adj, span);
        uint32_t cnt = 0;
        for (uint32_t vb = 0; vb < nfull; vb += VPB * K) {
            uint32_t w[K][BW];
            bool live[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t idx = vb + hw * K + k;
                live[k] = idx < nfull;
                const uint64_t v = v0 + (live[k] ? idx : nfull - 1u);  // a dead slot re-reads a live vector
                load_lane_g<BW>(words + v * (32u * BW) + l, w[k]);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                uint32_t ck = 0;
                if (rg.n) {
#pragma unroll
                    for (int j = 0; j < BW; ++j) ck += Swar<BW>::count(w[k][j], rg.ca[0], rg.cb[0]);
                }
                if (rg.n > 1) {  // uniform: a second interval only where the range wraps
#pragma unroll
                    for (int j = 0; j < BW; ++j) ck += Swar<BW>::count(w[k][j], rg.ca[1], rg.cb[1]);
                }
                cnt += live[k] ? ck : 0u;
            }
        }
        return cnt;
    } else {
        uint32_t cnt = 0;
        for (uint32_t vb = 0; vb < nfull; vb += VPB * K) {
            uint32_t w[K][BW];
            bool live[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t idx = vb + hw * K + k;
                live[k] = idx < nfull;
                const uint64_t v = v0 + (live[k] ? idx : nfull - 1u);  // a dead slot re-reads a live vector
                load_lane_g<BW>(words + v * (32u * BW) + l, w[k]);
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                uint32_t ck = 0;
#pragma unroll
                for (int r = 0; r < 32; ++r) ck += (uint32_t)(dev::extract_at<BW>(w[k], r) - lo_adj <= span);
                cnt += live[k] ? ck : 0u;
            }
        }
        return cnt;
    }
}

// DELTA with one width for every vector of the tile range: the chain walk of
// rsum.cuh with the running sum kept as value - lo.  Returns this work-item's
// count.
template <int BW, int VPB>
__device__ __forceinline__ uint32_t count_delta_full(const uint32_t* __restrict words,
                                                     const fls_vmeta* __restrict vmeta,
                                                     const uint32_t* __restrict bases, uint64_t v0, uint32_t nfull,
                                                     uint32_t lo, uint32_t span) {
    constexpr int K = batch_k<BW>();
    const uint32_t hw = threadIdx.x >> 5, l = dev::lane32();
    uint32_t cnt = 0;
    for (uint32_t vb = 0; vb < nfull; vb += VPB * K) {
        uint32_t w[K][BW > 0 ? BW : 1];
        uint32_t acc[K];
        bool live[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t idx = vb + hw * K + k;
            live[k] = idx < nfull;
            const uint64_t v = v0 + (live[k] ? idx : nfull - 1u);
            acc[k] = ldg(bases + v * 32u + l) - lo;
            if constexpr (BW > 0) load_lane_g<BW>(words + FLS_VMETA_OFFSET(ldg(vmeta + v)) + l, w[k]);
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            uint32_t ck = 0;
            if constexpr (BW == 0) {
                ck = 32u * (uint32_t)(acc[k] <= span);
            } else {
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    acc[k] += dev::extract_at<BW>(w[k], dev::rsum_row(j));
                    ck += (uint32_t)(acc[k] <= span);
                }
            }
            cnt += live[k] ? ck : 0u;
        }
    }
    return cnt;
}

// Flag words of a vector slot: bit k of word k/32 = run k's value in range.
constexpr uint32_t kFlagWords = 33;  // 1024 runs + one word the window may read past the last
constexpr int kRleK = 2;             // RLE vectors per half-wave per step

// RLE: each run value is tested once, by the 32 work-items of the vector's
// half-wave together (a ballot per 32 runs), into 32 LDS flag words; each
// work-item then reads the 32 flags from its first run on (two LDS words and
// a funnel shift) and walks its 32 rows in chain order: the run index advances
// at each run start (bit r of the lane's index word), and the row counts when
// the flag at that run offset is set.  No per-row LDS access or gather.  A
// half-wave takes K vectors per step; every load that does not depend on a
// run value (metadata, the lane's first run and index word) is issued before
// the run values are tested, so a step waits for two rounds of loads.
template <int VPB, int K>
__device__ __forceinline__ uint32_t count_rle_full(const SegDev& sg, uint64_t v0, uint32_t nfull, uint32_t lo,
                                                   uint32_t span, uint32_t* __restrict flags_lds) {
    const uint32_t hw = threadIdx.x >> 5, l = dev::lane32(), half = (threadIdx.x >> 5) & 1u;
    uint32_t c = 0;
    for (uint32_t vb = 0; vb < nfull; vb += VPB * K) {
        bool live[K];
        const uint32_t* vals[K];
        uint32_t nr[K], vlo[K], vbw[K], first[K], starts[K], ibw[K];
        uint32_t steps_here = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const uint32_t idx = vb + hw * K + k;
            live[k] = idx < nfull;
            const uint64_t v = v0 + (live[k] ? idx : nfull - 1u);
            const fls_vmeta m = ldg(sg.vmeta + v);
            const fls_rle_meta rm = ldg(sg.rmeta + v);
            first[k] = ldg(sg.bases + v * 32u + l);
            ibw[k] = FLS_VMETA_BW(m) ? 1u : 0u;
            const uint32_t* p = sg.words + FLS_VMETA_OFFSET(m);
            starts[k] = ibw[k] ? ldg_nt(p + l) : 0u;
            vals[k] = p + 32u * ibw[k];
            nr[k] = min((uint32_t)rm.n_runs, 1024u);
            vlo[k] = lo - rm.vbase;  // run value - vbase compared against lo - vbase
            vbw[k] = rm.vbw;
            steps_here = max(steps_here, (nr[k] + 31u) >> 5);
        }
        // the wave's half-waves may need different numbers of ballots: run to the larger
        const uint32_t steps = max(__builtin_amdgcn_readlane(steps_here, 0), __builtin_amdgcn_readlane(steps_here, 32));
        for (uint32_t i = 0; i < steps; ++i) {
            // the K vectors' value words are loaded before any of their ballots
            // (1.87 -> 1.82-1.83 ms over the 4096-segment workload against a
            // load, test and ballot per vector in turn; batching two rounds as
            // well measured 1.833, four rounds with both words always loaded
            // 2.01)
            const uint32_t r = 32u * i + l;
            uint32_t wlo[K], whi[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t bit = (r < nr[k] ? r : 0u) * vbw[k], wi = bit >> 5;  // a dead run reads run 0
                wlo[k] = vbw[k] ? ldg(vals[k] + wi) : 0u;
                whi[k] = (bit & 31u) + vbw[k] > 32u ? ldg(vals[k] + wi + 1u) : 0u;
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const uint32_t sh = ((r < nr[k] ? r : 0u) * vbw[k]) & 31u;
                const uint32_t mask = vbw[k] >= 32u ? 0xffffffffu : ((1u << vbw[k]) - 1u);
                const uint32_t x = __builtin_amdgcn_alignbit(whi[k], wlo[k], sh) & mask;
                const uint64_t b = __builtin_amdgcn_ballot_w64(r < nr[k] && x - vlo[k] <= span);
                if (l == 0u) flags_lds[(hw * K + k) * kFlagWords + i] = (uint32_t)(b >> (32u * half));
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (!live[k]) continue;
            const uint32_t* flags = flags_lds + (hw * K + k) * kFlagWords;
            const uint32_t F = __builtin_amdgcn_alignbit(flags[(first[k] >> 5) + 1u], flags[first[k] >> 5],
                                                         first[k] & 31u);
            if (ibw[k] == 0) {
                c += 32u * (F & 1u);
            } else {
                uint32_t off = 0;
#pragma unroll
                for (int j = 0; j < 32; ++j) {
                    off += (starts[k] >> dev::rsum_row(j)) & 1u;
                    c += (F >> off) & 1u;
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();  // the flag words are rewritten next step
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    return c;
}

// Slow path, any encoding, for a segment's partial last vector and for MIXED
// segments: one lane, rows walked with run-time width (two word loads per
// row, cache hits after the first) in a rolled loop, counting only the rows
// set in `rows` (register rows of this lane).  Few registers, so the fast
// paths keep the kernel's occupancy.
__device__ __forceinline__ uint32_t lane_extract(const uint32_t* __restrict lane_in, uint32_t bw, uint32_t r) {
    if (bw == 0) return 0u;
    const uint32_t bit = r * bw, k = bit >> 5, s = bit & 31u;
    const uint32_t a = ldg(lane_in + 32u * k);
    const uint32_t b = (s + bw > 32u) ? ldg(lane_in + 32u * (k + 1u)) : 0u;
    const uint32_t mask = bw >= 32 ? 0xffffffffu : ((1u << bw) - 1u);
    return __builtin_amdgcn_alignbit(b, a, s) & mask;
}

__device__ __forceinline__ uint32_t count_vector_slow(const SegDev& sg, uint64_t v, uint32_t rows, uint32_t lo,
                                                   uint32_t span) {
    const uint32_t l = dev::lane32();
    const uint32_t* p = sg.words;
    const uint32_t* vals = nullptr;
    uint32_t w = sg.bw, add = sg.base, seed = 0, gbase = 0, gbw = 0, nr = 1;
    bool chain = false;
    if (sg.enc == FLS_ENC_BITPACK) {
        p = sg.words + v * (32u * sg.bw);
    } else if (sg.enc == FLS_ENC_MIXED) {
        const fls_vmeta m = ldg(sg.vmeta + v);
        const uint32_t enc = (uint32_t)(m >> 6) & 3u;
        const uint64_t ax = ldg(sg.aux + v);
        p = sg.words + FLS_VMETA_OFFSET(m);
        w = (uint32_t)m & 63u;
        add = enc == FLS_MIX_FOR ? (uint32_t)ax : 0u;
        if (enc == FLS_MIX_DELTA || enc == FLS_MIX_RLE) {
            chain = true;
            seed = ldg(p + l);
            p += 32u;
            if (enc == FLS_MIX_RLE) {
                const uint32_t ibw = (uint32_t)(ax >> 48) & 1u;
                vals = p + 32u * ibw;
                gbase = (uint32_t)ax;
                gbw = w;
                w = ibw;
                nr = min((uint32_t)(ax >> 32) & 0xffffu, 1024u);
            }
        }
    } else {  // DELTA, RLE
        const fls_vmeta m = ldg(sg.vmeta + v);
        p = sg.words + FLS_VMETA_OFFSET(m);
        w = FLS_VMETA_BW(m);
        add = 0u;
        chain = true;
        seed = ldg(sg.bases + v * 32u + l);
        if (sg.enc == FLS_ENC_RLE) {
            const fls_rle_meta rm = ldg(sg.rmeta + v);
            vals = p + 32u * w;
            gbase = rm.vbase;
            gbw = rm.vbw;
            nr = min((uint32_t)rm.n_runs, 1024u);
        }
    }
    const uint32_t last = nr ? nr - 1u : 0u;
    uint32_t acc = seed, c = 0;
#pragma unroll 1
    for (uint32_t j = 0; j < 32; ++j) {
        const uint32_t r = chain ? (uint32_t)dev::rsum_row((int)j) : j;
        uint32_t x = lane_extract(p + l, w, r);
        if (chain) {
            acc += x;
            x = acc;
        }
        x = vals ? gbase + linear_extract(vals, min(x, last), gbw) : x + add;
        c += ((rows >> r) & 1u) & (uint32_t)(x - lo <= span);
    }
    return c;
}

// ---------------------------------------------------------------- kernel
// Eight waves per SIMD (64 VGPRs): the fast paths are latency-bound, and
// measured 2.32 ms at five waves against 1.99 ms at eight over the
// 4096-segment batch (a few loop-invariant values spill; no spill in a row loop).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(8, 8))) void batch_kernel(const SegDev* __restrict segs, const TileDev* __restrict tiles,
                                                      uint32_t n_tiles, uint64_t* __restrict counts,
                                                      uint32_t* __restrict queue, uint32_t dynamic) {
    constexpr int VPB = BLOCK / 32;
    constexpr int WAVES = BLOCK / 64;
    __shared__ uint32_t flags_lds[VPB * kRleK * kFlagWords];
    __shared__ uint32_t s_part[WAVES];
    __shared__ uint32_t s_mm[2 * WAVES];
    __shared__ uint32_t s_tile;
    const uint32_t hw = threadIdx.x >> 5, wave = threadIdx.x >> 6;
    uint32_t iter = 0;
    for (;;) {
        uint32_t t;
        if (dynamic) {
            if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1u);
            __syncthreads();
            t = s_tile;
        } else {
            t = blockIdx.x + iter * gridDim.x;
            ++iter;
        }
        if (t >= n_tiles) break;  // uniform over the block
        const TileDev tile = tiles[t];
        const SegDev sg = segs[tile.seg];
        const uint32_t lo = sg.lo, span = sg.hi - sg.lo;
        const uint64_t seg_vecs = ceil_div(sg.n_values, 1024);
        const bool has_partial = (sg.n_values & 1023u) != 0 && tile.v0 + tile.nv == seg_vecs;
        const uint32_t nfull = tile.nv - (has_partial ? 1u : 0u);
        uint32_t lane_cnt = 0;  // per work-item
        const bool none = sg.hi < sg.lo;  // lo <= value <= hi holds for no value (block-uniform)
        if (none) {
        } else if (sg.enc == FLS_ENC_BITPACK) {
            const uint32_t lo_adj = lo - sg.base;
            switch (sg.bw) {
#define FLS_BK_CASE(B) \
    case B: lane_cnt = count_fixed_full<B, VPB>(sg.words, tile.v0, nfull, lo_adj, span); break;
                FLS_WIDTH_CASES(FLS_BK_CASE)
#undef FLS_BK_CASE
                default: break;
            }
            if (sg.bw == 0 && (0u - lo_adj) <= span && threadIdx.x == 0) lane_cnt = 1024u * nfull;
        } else if (sg.enc == FLS_ENC_DELTA) {
            // one width for the whole tile range takes the batched walk
            uint32_t mn = 0xffffffffu, mx = 0u;
            for (uint32_t i = threadIdx.x; i < nfull; i += BLOCK) {
                const uint32_t b = FLS_VMETA_BW(ldg(sg.vmeta + tile.v0 + i));
                mn = min(mn, b);
                mx = max(mx, b);
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                mn = min(mn, (uint32_t)__shfl_xor(mn, off, 64));
                mx = max(mx, (uint32_t)__shfl_xor(mx, off, 64));
            }
            if (dev::is_wave_leader()) {
                s_mm[2 * wave] = mn;
                s_mm[2 * wave + 1] = mx;
            }
            __syncthreads();
            mn = s_mm[0];
            mx = s_mm[1];
#pragma unroll
            for (int w = 1; w < WAVES; ++w) {
                mn = min(mn, s_mm[2 * w]);
                mx = max(mx, s_mm[2 * w + 1]);
            }
            if (nfull && mn == mx) {
                switch (mn) {
#define FLS_DK_CASE(B)                                                                                       \
    case B: lane_cnt = count_delta_full<B, VPB>(sg.words, sg.vmeta, sg.bases, tile.v0, nfull, lo, span); \
        break;
                    FLS_WIDTH_CASES(FLS_DK_CASE)
#undef FLS_DK_CASE
                    default: break;
                }
            } else {
                // widths differ inside the tile: one vector per half-wave at its own width
                for (uint32_t vb = hw; vb < nfull; vb += VPB) {
                    const uint64_t v = tile.v0 + vb;
                    const fls_vmeta m = ldg(sg.vmeta + v);
                    uint32_t c = 0;
                    auto f = [&](int, uint32_t x) { c += (uint32_t)(x - lo <= span); };
                    walk_delta(sg.words + FLS_VMETA_OFFSET(m) + dev::lane32(), FLS_VMETA_BW(m),
                               ldg(sg.bases + v * 32u + dev::lane32()), f);
                    lane_cnt += c;
                }
            }
        } else if (sg.enc == FLS_ENC_RLE) {
            lane_cnt = count_rle_full<VPB, kRleK>(sg, tile.v0, nfull, lo, span, flags_lds);
        } else {  // MIXED: every vector through the slow path
            for (uint32_t vb = hw; vb < nfull; vb += VPB)
                lane_cnt += count_vector_slow(sg, tile.v0 + vb, 0xffffffffu, lo, span);
        }
        if (has_partial && hw == 0 && !none) {
            const uint64_t v = seg_vecs - 1u;
            lane_cnt += count_vector_slow(sg, v, live_row_mask((uint32_t)(sg.n_values - v * 1024u), sg.transposed != 0),
                                          lo, span);
        }
        const uint32_t total_wave = dev::wave_sum(lane_cnt);
        if (dev::is_wave_leader()) s_part[wave] = total_wave;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t total = 0;
#pragma unroll
            for (int w = 0; w < WAVES; ++w) total += s_part[w];
            if (total) atomicAdd((unsigned long long*)&counts[tile.seg], (unsigned long long)total);
        }
        __syncthreads();  // s_part / s_mm / s_tile / flags are rewritten next iteration
    }
}

__global__ void zero_u32_kernel(uint32_t* __restrict p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// ------------------------------------------------------------ generation
struct SegParams {
    uint32_t enc, bw;
    uint64_t len;
};

SegParams c4_params(uint64_t seed, uint32_t s, uint32_t lmin, uint32_t lmax) {
    const uint64_t key = rng::stream_key(seed, rng::kSegParams + s);
    const uint32_t e = lmin + rng::uniform(key, 0, lmax - lmin);
    SegParams p;
    p.len = (1ull << e) + rng::uniform(key, 1, 1ull << e);
    p.bw = 1u + rng::uniform(key, 2, 32);
    const uint32_t r = rng::uniform(key, 3, 10);
    p.enc = r < 4 ? FLS_ENC_BITPACK : r < 7 ? FLS_ENC_DELTA : FLS_ENC_RLE;
    return p;
}

__host__ __device__ inline uint32_t bit_width64(uint64_t x) {
    uint32_t w = 0;
    while (x) {
        ++w;
        x >>= 1;
    }
    return w;
}

// Values of segment s (see fls.h: fls_batch_create_c4).
__global__ void gen_segment_kernel(uint64_t seed, uint32_t s, uint32_t enc, uint32_t bw, uint64_t len,
                                   uint32_t* __restrict out) {
    const uint64_t kv = rng::stream_key(seed, rng::kSegValues + s);
    const uint64_t kr = rng::stream_key(seed, rng::kSegRuns + s);
    uint32_t b = 32u - bit_width64(len);
    b = b < 1u ? 1u : b;
    b = bw < b ? bw : b;
    const uint64_t m = 1ull << (b - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x;
        if (enc == FLS_ENC_BITPACK) {
            x = rng::uniform(kv, i, 1ull << bw);
        } else if (enc == FLS_ENC_DELTA) {
            x = (uint32_t)(i * m) + rng::uniform(kv, i, m);
        } else {
            uint64_t j = i;
            while (j > 0 && rng::uniform(kr, j, 8) != 0) --j;  // start of row i's run
            x = rng::uniform(kv, j, 1ull << bw);
        }
        out[i] = x;
    }
}

__global__ void minmax_kernel(const uint32_t* __restrict in, uint64_t n, uint32_t* __restrict mm) {
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = in[i];
        lo = min(lo, x);
        hi = max(hi, x);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor(lo, off, 64));
        hi = max(hi, (uint32_t)__shfl_xor(hi, off, 64));
    }
    // one atomic pair per block (blockDim.x == 256)
    __shared__ uint32_t s_lo[4], s_hi[4];
    if (dev::is_wave_leader()) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void mm_init_kernel(uint32_t* __restrict mm, uint32_t n_seg) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_seg; i += gridDim.x * blockDim.x) {
        mm[2 * i] = 0xffffffffu;
        mm[2 * i + 1] = 0u;
    }
}

uint32_t grid_for(uint64_t n) {
    const uint64_t g = ceil_div(n, 256);
    return (uint32_t)(g < 4096 ? (g ? g : 1) : 4096);
}

}  // namespace
}  // namespace fls

// ------------------------------------------------------------------- batch
struct fls_batch {
    int device = 0;
    std::vector<fls::SegDev> segs;
    std::vector<fls_segment_info> info;
    fls::SegDev* d_segs = nullptr;
    // owned pools (C4 builder); null for caller-described segments
    uint32_t* d_words = nullptr;
    fls_vmeta* d_vmeta = nullptr;
    uint32_t* d_bases = nullptr;
    fls_rle_meta* d_rmeta = nullptr;
    // tiles for the current tiling (tile_vectors, class_tiles, class filter)
    uint32_t tile_vectors = 0, class_tiles = 0;
    int tile_filter = -2;
    fls::TileDev* d_tiles = nullptr;
    uint32_t n_tiles = 0;
    std::vector<uint32_t> seg_tile_off, seg_tile_cnt;
    uint32_t* d_queue = nullptr;
    // graph
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    fls_tune graph_tune{};
    uint64_t* graph_counts = nullptr;
    uint64_t n_values = 0, n_vectors = 0, encoded_bytes = 0;
    uint32_t count_by_enc[3] = {0, 0, 0};
};

namespace fls {
namespace {

// Tiling class of segment s: 2 * encoding + (more than 8 encoded bits per
// value), so narrow and wide segments of each encoding can get their own tile
// size (fls_tune.class_tiles).
uint32_t seg_class(const fls_batch* b, uint32_t s) {
    const fls_segment_info& i = b->info[s];
    const uint64_t bits = i.n_values ? 32ull * i.n_words / i.n_values : 0u;
    return 2u * (i.encoding < 3u ? i.encoding : 3u) + (bits > 8u ? 1u : 0u);
}

uint32_t class_tile(const fls_tune& t, uint32_t cls) {
    const uint32_t nib = (t.class_tiles >> (4u * cls)) & 15u;
    return nib ? (1u << nib) : t.tile_vectors;
}

// Cut the segments into tiles (only those of class `only`, when >= 0) and
// upload the tile list, unless the current one already is that tiling.
int set_tiles(fls_batch* b, const fls_tune& t, int only = -1) {
    if (t.tile_vectors == 0) return FLS_ERR_ARG;
    if (b->d_tiles && b->tile_vectors == t.tile_vectors && b->class_tiles == t.class_tiles && b->tile_filter == only)
        return FLS_OK;
    std::vector<TileDev> tiles;
    b->seg_tile_off.assign(b->segs.size(), 0);
    b->seg_tile_cnt.assign(b->segs.size(), 0);
    for (uint32_t s = 0; s < b->segs.size(); ++s) {
        const uint32_t cls = seg_class(b, s);
        b->seg_tile_off[s] = (uint32_t)tiles.size();
        if (only >= 0 && (int)cls != only) continue;
        const uint32_t tv = class_tile(t, cls);
        const uint64_t nv = ceil_div(b->segs[s].n_values, 1024);
        for (uint64_t v0 = 0; v0 < nv; v0 += tv)
            tiles.push_back(TileDev{s, (uint32_t)v0, (uint32_t)(nv - v0 < tv ? nv - v0 : tv), 0});
        b->seg_tile_cnt[s] = (uint32_t)tiles.size() - b->seg_tile_off[s];
    }
    if (b->d_tiles) FLS_HIP_CHECK(hipFree(b->d_tiles));
    b->d_tiles = nullptr;
    FLS_HIP_CHECK(hipMalloc((void**)&b->d_tiles, sizeof(TileDev) * (tiles.size() ? tiles.size() : 1)));
    if (!tiles.empty())
        FLS_HIP_CHECK(hipMemcpy(b->d_tiles, tiles.data(), sizeof(TileDev) * tiles.size(), hipMemcpyHostToDevice));
    b->n_tiles = (uint32_t)tiles.size();
    b->tile_vectors = t.tile_vectors;
    b->class_tiles = t.class_tiles;
    b->tile_filter = only;
    return FLS_OK;
}

int upload_segs(fls_batch* b) {
    if (b->d_segs) FLS_HIP_CHECK(hipFree(b->d_segs));
    b->d_segs = nullptr;
    FLS_HIP_CHECK(hipMalloc((void**)&b->d_segs, sizeof(SegDev) * (b->segs.size() ? b->segs.size() : 1)));
    if (!b->segs.empty())
        FLS_HIP_CHECK(hipMemcpy(b->d_segs, b->segs.data(), sizeof(SegDev) * b->segs.size(), hipMemcpyHostToDevice));
    if (!b->d_queue) FLS_HIP_CHECK(hipMalloc((void**)&b->d_queue, 256));
    b->n_values = 0;
    b->n_vectors = 0;
    for (const auto& s : b->segs) {
        b->n_values += s.n_values;
        b->n_vectors += ceil_div(s.n_values, 1024);
    }
    return FLS_OK;
}

template <int BLOCK>
int launch_batch(fls_batch* b, uint64_t* d_counts, bool dynamic, uint32_t tile_off, uint32_t n_tiles, uint32_t grid,
                 hipStream_t s) {
    if (n_tiles == 0) return FLS_OK;
    hipLaunchKernelGGL(batch_kernel<BLOCK>, dim3(grid), dim3(BLOCK), 0, s, b->d_segs, b->d_tiles + tile_off, n_tiles,
                       d_counts, b->d_queue, (uint32_t)dynamic);
    return check_launch();
}

int check_tune(const fls_tune* t) {
    if (!t) return FLS_ERR_NULL;
    if (t->tile_vectors == 0 || t->tile_vectors > (1u << 20)) return FLS_ERR_ARG;
    if (t->block_threads != 256 && t->block_threads != 512) return FLS_ERR_ARG;
    if (t->schedule > FLS_SCHED_DYNAMIC) return FLS_ERR_ARG;
    return FLS_OK;  // every class_tiles nibble is valid (tiles of 2 .. 32768 vectors)
}

// zero counters + queue, then the persistent launch over the tiles of class
// `only` (all classes when < 0)
int run_batch(fls_batch* b, const fls_tune* t, uint64_t* d_counts, hipStream_t s, int only = -1) {
    int st;
    if ((st = check_tune(t)) != FLS_OK) return st;
    if ((st = set_tiles(b, *t, only)) != FLS_OK) return st;
    if ((st = zero_u64(d_counts, b->segs.size(), s)) != FLS_OK) return st;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, b->d_queue, 1u);
    const bool dynamic = t->schedule == FLS_SCHED_DYNAMIC;
    uint32_t grid = 0;
    if (t->block_threads == 512) {
        if ((st = resident_grid_of((const void*)batch_kernel<512>, 512, 0, b->n_tiles, &grid)) != FLS_OK) return st;
        return launch_batch<512>(b, d_counts, dynamic, 0, b->n_tiles, grid, s);
    }
    if ((st = resident_grid_of((const void*)batch_kernel<256>, 256, 0, b->n_tiles, &grid)) != FLS_OK) return st;
    return launch_batch<256>(b, d_counts, dynamic, 0, b->n_tiles, grid, s);
}

// ---------------------------------------------------------------- autotune
std::string default_cache_path() {
    const char* env = getenv("FLS_AUTOTUNE_CACHE");
    if (env && *env) return env;
    const char* home = getenv("HOME");
    return std::string(home && *home ? home : "/tmp") + "/.cache/fls_mi355x/autotune.txt";
}

void mkdirs_for(const std::string& path) {
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
}

std::string batch_key(const fls_batch* b) {
    hipDeviceProp_t prop;
    char name[64] = "gfx";
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) snprintf(name, sizeof name, "%s", prop.gcnArchName);
    for (char* c = name; *c; ++c)
        if (*c == ':' || *c == ' ') *c = '_';
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the segment shape
    for (const auto& s : b->segs) {
        const uint64_t parts[3] = {s.enc, s.bw, s.n_values};
        for (uint64_t p : parts) {
            h ^= p;
            h *= 1099511628211ull;
        }
    }
    uint32_t cus = 0;
    device_cus(&cus);
    char buf[160];
    snprintf(buf, sizeof buf, "%s-cu%u-seg%zu-vec%llu-%016llx", name, cus, b->segs.size(),
             (unsigned long long)b->n_vectors, (unsigned long long)h);
    return buf;
}

bool cache_lookup(const std::string& path, const std::string& key, fls_tune* t) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    char k[200];
    unsigned tile, block, sched;
    double ms;
    bool found = false;
    unsigned classes;
    while (fscanf(f, "%199s %u %u %u %x %lf", k, &tile, &block, &sched, &classes, &ms) == 6) {
        if (key == k) {
            *t = fls_tune{tile, block, sched, classes};
            found = true;  // keep the last entry for the key
        }
    }
    fclose(f);
    return found && check_tune(t) == FLS_OK;
}

void cache_store(const std::string& path, const std::string& key, const fls_tune& t, double ms) {
    mkdirs_for(path);
    FILE* f = fopen(path.c_str(), "a");
    if (!f) return;
    fprintf(f, "%s %u %u %u %x %.6f\n", key.c_str(), t.tile_vectors, t.block_threads, t.schedule, t.class_tiles, ms);
    fclose(f);
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_batch_create(const fls_segment* segs, uint32_t n_segments, fls_batch** out) {
    using namespace fls;
    FLS_RETURN_IF(!out || (n_segments && !segs), FLS_ERR_NULL);
    fls_batch* b = new fls_batch();
    FLS_HIP_CHECK(hipGetDevice(&b->device));
    for (uint32_t s = 0; s < n_segments; ++s) {
        const fls_segment& g = segs[s];
        const fls_column& c = g.col;
        ColDev tmp;
        // RLE metadata comes from the segment's rmeta or, when that is NULL, the column's aux
        const fls_rle_meta* rmeta = g.rmeta ? g.rmeta : reinterpret_cast<const fls_rle_meta*>(c.aux);
        fls_column cc = c;
        if (c.encoding == FLS_ENC_RLE) cc.aux = reinterpret_cast<const uint64_t*>(rmeta);
        const int st = to_dev_column(cc, g.n_values, &tmp, true);
        if (st) {
            delete b;
            return st;
        }
        const uint32_t tr = c.encoding == FLS_ENC_BITPACK ? (c.transposed ? 1u : 0u) : 1u;
        b->segs.push_back(
            SegDev{c.encoding, c.bw, tr, c.encoding == FLS_ENC_BITPACK ? c.base : 0u, g.n_values, c.words, c.vmeta,
                   c.bases, rmeta, c.aux, g.lo, g.hi});
        b->info.push_back(fls_segment_info{c.encoding, c.bw, g.n_values, g.lo, g.hi, c.n_words});
        if (c.encoding <= FLS_ENC_RLE) b->count_by_enc[c.encoding]++;
        b->encoded_bytes += 4ull * c.n_words;
    }
    int st = upload_segs(b);
    if (st) {
        fls_batch_destroy(b);
        return st;
    }
    *out = b;
    return FLS_OK;
}

void fls_c4_segment_params(uint64_t seed, uint32_t s, uint32_t log2_min_len, uint32_t log2_max_len,
                           uint32_t* encoding, uint32_t* bw, uint64_t* n_values) {
    const fls::SegParams p = fls::c4_params(seed, s, log2_min_len, log2_max_len);
    if (encoding) *encoding = p.enc;
    if (bw) *bw = p.bw;
    if (n_values) *n_values = p.len;
}

int fls_gen_c4_segment(uint64_t seed, uint32_t s, uint32_t log2_min_len, uint32_t log2_max_len, uint32_t* d_out,
                       void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!d_out, FLS_ERR_NULL);
    FLS_RETURN_IF(log2_min_len >= log2_max_len || log2_max_len > 30, FLS_ERR_ARG);
    const SegParams p = c4_params(seed, s, log2_min_len, log2_max_len);
    hipLaunchKernelGGL(gen_segment_kernel, dim3(grid_for(p.len)), dim3(256), 0, as_stream(stream), seed, s, p.enc,
                       p.bw, p.len, d_out);
    return check_launch();
}

int fls_batch_create_c4(uint64_t seed, uint32_t n_segments, uint32_t log2_min_len, uint32_t log2_max_len,
                        fls_batch** out, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!out, FLS_ERR_NULL);
    FLS_RETURN_IF(n_segments == 0 || log2_min_len >= log2_max_len || log2_max_len > 30, FLS_ERR_ARG);
    hipStream_t s = as_stream(stream);
    std::vector<SegParams> P(n_segments);
    uint64_t max_len = 0, meta_vec = 0;
    std::vector<uint64_t> meta_off(n_segments, 0);
    for (uint32_t i = 0; i < n_segments; ++i) {
        P[i] = c4_params(seed, i, log2_min_len, log2_max_len);
        max_len = P[i].len > max_len ? P[i].len : max_len;
        meta_off[i] = meta_vec;
        if (P[i].enc != FLS_ENC_BITPACK) meta_vec += ceil_div(P[i].len, 1024);
    }
    fls_batch* b = new fls_batch();
    int st = FLS_OK;
    uint32_t* d_raw = nullptr;
    void* d_scratch = nullptr;
    uint64_t* d_totals = nullptr;
    uint32_t* d_mm = nullptr;
    std::vector<uint64_t> totals(n_segments, 0);
    std::vector<uint32_t> mm(2ull * n_segments, 0);
    std::vector<uint64_t> word_off(n_segments, 0);
    uint64_t words_total = 0;
#define FLS_B(expr)                                  \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) {                      \
            st = record_hip_error(_e);               \
            goto fail;                               \
        }                                            \
    } while (0)
#define FLS_S(expr)          \
    do {                     \
        st = (expr);         \
        if (st) goto fail;   \
    } while (0)
    FLS_B(hipGetDevice(&b->device));
    FLS_B(hipMalloc((void**)&d_raw, 4ull * max_len));
    FLS_B(hipMalloc(&d_scratch, fls_encode_scratch_bytes(max_len)));
    FLS_B(hipMalloc((void**)&d_totals, 8ull * n_segments));
    FLS_B(hipMalloc((void**)&d_mm, 8ull * n_segments));
    FLS_B(hipMalloc((void**)&b->d_vmeta, 8ull * (meta_vec ? meta_vec : 1)));
    FLS_B(hipMalloc((void**)&b->d_bases, 128ull * (meta_vec ? meta_vec : 1)));
    FLS_B(hipMalloc((void**)&b->d_rmeta, 8ull * (meta_vec ? meta_vec : 1)));
    FLS_S(zero_u64(d_totals, n_segments, s));
    hipLaunchKernelGGL(mm_init_kernel, dim3(64), dim3(256), 0, s, d_mm, n_segments);
    // pass 1: generate, min/max, plan variable-width encodings
    for (uint32_t i = 0; i < n_segments; ++i) {
        hipLaunchKernelGGL(gen_segment_kernel, dim3(grid_for(P[i].len)), dim3(256), 0, s, seed, i, P[i].enc, P[i].bw,
                           P[i].len, d_raw);
        hipLaunchKernelGGL(minmax_kernel, dim3(grid_for(P[i].len)), dim3(256), 0, s, d_raw, P[i].len, d_mm + 2 * i);
        if (P[i].enc == FLS_ENC_DELTA)
            FLS_S(fls_delta_plan_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], b->d_bases + 32 * meta_off[i],
                                        d_totals + i, d_scratch, s));
        else if (P[i].enc == FLS_ENC_RLE)
            FLS_S(fls_rle_plan_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], b->d_rmeta + meta_off[i],
                                      b->d_bases + 32 * meta_off[i], d_totals + i, d_scratch, s));
        FLS_S(check_launch());
    }
    FLS_B(hipMemcpyAsync(totals.data(), d_totals, 8ull * n_segments, hipMemcpyDeviceToHost, s));
    FLS_B(hipMemcpyAsync(mm.data(), d_mm, 8ull * n_segments, hipMemcpyDeviceToHost, s));
    FLS_B(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n_segments; ++i) {
        word_off[i] = words_total;
        words_total += P[i].enc == FLS_ENC_BITPACK ? ceil_div(P[i].len, 1024) * 32u * P[i].bw : totals[i];
    }
    FLS_B(hipMalloc((void**)&b->d_words, 4ull * (words_total ? words_total : 1)));
    // pass 2: generate again, pack
    for (uint32_t i = 0; i < n_segments; ++i) {
        hipLaunchKernelGGL(gen_segment_kernel, dim3(grid_for(P[i].len)), dim3(256), 0, s, seed, i, P[i].enc, P[i].bw,
                           P[i].len, d_raw);
        uint32_t* w = b->d_words + word_off[i];
        if (P[i].enc == FLS_ENC_BITPACK) FLS_S(fls_pack_column_device(d_raw, P[i].len, P[i].bw, 0, 0, w, s));
        else if (P[i].enc == FLS_ENC_DELTA) FLS_S(fls_delta_pack_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], w, s));
        else FLS_S(fls_rle_pack_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], b->d_rmeta + meta_off[i], w, s));
        const uint32_t mn = mm[2 * i], mx = mm[2 * i + 1];
        const uint64_t range = (uint64_t)(mx - mn);
        const uint32_t lo = mn + (uint32_t)(range / 4), hi = mn + (uint32_t)(range * 3 / 4);
        const bool meta = P[i].enc != FLS_ENC_BITPACK;
        b->segs.push_back(SegDev{P[i].enc, P[i].bw, P[i].enc == FLS_ENC_BITPACK ? 0u : 1u, 0u, P[i].len, w,
                                 meta ? b->d_vmeta + meta_off[i] : nullptr,
                                 meta ? b->d_bases + 32 * meta_off[i] : nullptr,
                                 P[i].enc == FLS_ENC_RLE ? b->d_rmeta + meta_off[i] : nullptr, nullptr, lo, hi});
        const uint64_t seg_words = P[i].enc == FLS_ENC_BITPACK ? ceil_div(P[i].len, 1024) * 32u * P[i].bw : totals[i];
        b->info.push_back(fls_segment_info{P[i].enc, P[i].bw, P[i].len, lo, hi, seg_words});
        const uint64_t nv = ceil_div(P[i].len, 1024);
        b->encoded_bytes += 4ull * seg_words + (meta ? nv * (8u + 128u) : 0u) + (P[i].enc == FLS_ENC_RLE ? nv * 8u : 0u);
        b->count_by_enc[P[i].enc]++;
    }
    FLS_B(hipStreamSynchronize(s));
    FLS_S(upload_segs(b));
    (void)hipFree(d_raw);
    (void)hipFree(d_scratch);
    (void)hipFree(d_totals);
    (void)hipFree(d_mm);
    *out = b;
    return FLS_OK;
fail:
    if (d_raw) (void)hipFree(d_raw);
    if (d_scratch) (void)hipFree(d_scratch);
    if (d_totals) (void)hipFree(d_totals);
    if (d_mm) (void)hipFree(d_mm);
    fls_batch_destroy(b);
    return st;
#undef FLS_B
#undef FLS_S
}

int fls_batch_get_info(const fls_batch* b, fls_batch_info* info) {
    if (!b || !info) return FLS_ERR_NULL;
    info->n_segments = (uint32_t)b->segs.size();
    info->n_values = b->n_values;
    info->n_vectors = b->n_vectors;
    info->encoded_bytes = b->encoded_bytes;
    info->n_bitpack = b->count_by_enc[0];
    info->n_delta = b->count_by_enc[1];
    info->n_rle = b->count_by_enc[2];
    return FLS_OK;
}

int fls_batch_segment(const fls_batch* b, uint32_t s, fls_segment_info* info) {
    if (!b || !info) return FLS_ERR_NULL;
    if (s >= b->info.size()) return FLS_ERR_ARG;
    *info = b->info[s];
    return FLS_OK;
}

int fls_batch_segment_column(const fls_batch* b, uint32_t s, fls_segment* seg) {
    if (!b || !seg) return FLS_ERR_NULL;
    if (s >= b->segs.size()) return FLS_ERR_ARG;
    const fls::SegDev& d = b->segs[s];
    fls_segment g{};
    g.col.encoding = d.enc;
    g.col.bw = d.bw;
    g.col.base = d.base;
    g.col.transposed = d.transposed;
    g.col.n_words = b->info[s].n_words;
    g.col.words = d.words;
    g.col.vmeta = d.vmeta;
    g.col.bases = d.bases;
    g.col.aux = d.enc == FLS_ENC_RLE ? reinterpret_cast<const uint64_t*>(d.rmeta) : d.aux;
    g.rmeta = d.rmeta;
    g.n_values = d.n_values;
    g.lo = d.lo;
    g.hi = d.hi;
    *seg = g;
    return FLS_OK;
}

int fls_batch_run(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream) {
    if (!b || !d_counts) return FLS_ERR_NULL;
    return fls::run_batch(b, tune, d_counts, fls::as_stream(stream));
}

int fls_batch_run_per_segment(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream) {
    using namespace fls;
    if (!b || !d_counts) return FLS_ERR_NULL;
    int st;
    if ((st = check_tune(tune)) != FLS_OK) return st;
    if ((st = set_tiles(b, *tune)) != FLS_OK) return st;
    hipStream_t s = as_stream(stream);
    if ((st = zero_u64(d_counts, b->segs.size(), s)) != FLS_OK) return st;
    for (uint32_t i = 0; i < b->segs.size(); ++i) {
        // one launch per segment: one block per tile of the segment
        const uint32_t nt = b->seg_tile_cnt[i];
        st = tune->block_threads == 512 ? launch_batch<512>(b, d_counts, false, b->seg_tile_off[i], nt, nt, s)
                                        : launch_batch<256>(b, d_counts, false, b->seg_tile_off[i], nt, nt, s);
        if (st) return st;
    }
    return FLS_OK;
}

int fls_batch_graph_launch(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream) {
    using namespace fls;
    if (!b || !d_counts) return FLS_ERR_NULL;
    int st;
    if ((st = check_tune(tune)) != FLS_OK) return st;
    const bool same = b->exec && b->graph_counts == d_counts && b->graph_tune.tile_vectors == tune->tile_vectors &&
                      b->graph_tune.block_threads == tune->block_threads && b->graph_tune.schedule == tune->schedule &&
                      b->graph_tune.class_tiles == tune->class_tiles;
    if (!same) {
        if (b->exec) FLS_HIP_CHECK(hipGraphExecDestroy(b->exec));
        if (b->graph) FLS_HIP_CHECK(hipGraphDestroy(b->graph));
        b->exec = nullptr;
        b->graph = nullptr;
        if ((st = set_tiles(b, *tune)) != FLS_OK) return st;  // uploads before capture
        if (!b->cap_stream) FLS_HIP_CHECK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        FLS_HIP_CHECK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        st = run_batch(b, tune, d_counts, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return record_hip_error(e);
        b->graph = g;
        FLS_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_counts = d_counts;
        b->graph_tune = *tune;
    }
    FLS_HIP_CHECK(hipGraphLaunch(b->exec, as_stream(stream)));
    return FLS_OK;
}

int fls_batch_autotune(fls_batch* b, const char* cache_path, int force, uint64_t* d_counts, fls_tune* best,
                       double* best_ms, void* stream) {
    using namespace fls;
    if (!b || !best || !d_counts) return FLS_ERR_NULL;
    const std::string path = cache_path && *cache_path ? cache_path : default_cache_path();
    const std::string key = batch_key(b);
    if (!force && cache_lookup(path, key, best)) {
        if (best_ms) *best_ms = -1.0;
        return FLS_OK;
    }
    hipStream_t s = as_stream(stream);
    hipEvent_t e0, e1;
    FLS_HIP_CHECK(hipEventCreate(&e0));
    FLS_HIP_CHECK(hipEventCreate(&e1));
    // best of three timed runs after a warm-up (which also uploads the tiles)
    auto timed = [&](const fls_tune& t, int only, double* out) -> int {
        int rc = run_batch(b, &t, d_counts, s, only);
        double tmin = 1e30;
        for (int rep = 0; rep < 3 && rc == FLS_OK; ++rep) {
            (void)hipEventRecord(e0, s);
            rc = run_batch(b, &t, d_counts, s, only);
            (void)hipEventRecord(e1, s);
            (void)hipEventSynchronize(e1);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            tmin = ms < tmin ? ms : tmin;
        }
        *out = tmin;
        return rc;
    };
    const uint32_t tiles[] = {8, 16, 32, 64, 128, 256, 512};
    const uint32_t blocks[] = {256, 512};
    double best_t = 1e30;
    int st = FLS_OK;
    // 1) one tile size, block size and schedule for every segment
    for (uint32_t tv : tiles)
        for (uint32_t bt : blocks)
            for (uint32_t sched = 0; sched <= FLS_SCHED_DYNAMIC && st == FLS_OK; ++sched) {
                const fls_tune t{tv, bt, sched, 0};
                double ms = 0;
                if ((st = timed(t, -1, &ms)) == FLS_OK && ms < best_t) {
                    best_t = ms;
                    *best = t;
                }
            }
    // 2) at that block size and schedule, a tile size per (encoding, width)
    //    class, each timed on the class's segments alone; kept if the whole
    //    batch then runs faster than with one tile size
    bool present[8] = {false, false, false, false, false, false, false, false};
    for (uint32_t i = 0; i < b->segs.size(); ++i) present[seg_class(b, i)] = true;
    uint32_t classes = 0;
    for (uint32_t c = 0; c < 8 && st == FLS_OK; ++c) {
        if (!present[c]) continue;
        double cbest = 1e30;
        uint32_t cl = 0;
        for (uint32_t tv : tiles) {
            const fls_tune t{tv, best->block_threads, best->schedule, 0};
            double ms = 0;
            if ((st = timed(t, (int)c, &ms)) != FLS_OK) break;
            if (ms < cbest) {
                cbest = ms;
                cl = (uint32_t)__builtin_ctz(tv);
            }
        }
        classes |= cl << (4u * c);
    }
    if (st == FLS_OK && classes) {
        const fls_tune t{best->tile_vectors, best->block_threads, best->schedule, classes};
        double ms = 0;
        if ((st = timed(t, -1, &ms)) == FLS_OK && ms < best_t) {
            best_t = ms;
            *best = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    cache_store(path, key, *best, best_t);
    if (best_ms) *best_ms = best_t;
    return FLS_OK;
}

void fls_batch_destroy(fls_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_segs) (void)hipFree(b->d_segs);
    if (b->d_queue) (void)hipFree(b->d_queue);
    if (b->d_words) (void)hipFree(b->d_words);
    if (b->d_vmeta) (void)hipFree(b->d_vmeta);
    if (b->d_bases) (void)hipFree(b->d_bases);
    if (b->d_rmeta) (void)hipFree(b->d_rmeta);
    delete b;
}

}  // extern "C"
