// Per-vector encoding selection and the mixed-encoding column.
//
//  * stats / exhaustive search: for each 1024-value vector, the exact encoded
//    size of each of the four encodings (bit-pack, FOR, FastLanes delta,
//    FastLanes RLE) at its smallest width, and the best of the 40 classes
//    (encoding x width bucket).  The size is computed from the statistics the
//    encoder itself would derive (widths, run count), so this IS the size the
//    pack step produces.
//  * features: 16 cheap statistics of a 1/8 sample of the vector (four
//    128-byte lines, the first and the last among them) -- the learned
//    selector's input.
//  * mixed column: each vector stored with its own encoding, all in FastLanes
//    transposed row order so mixed columns of one table stay row-aligned.
//    plan (chosen encoding's statistics -> size) -> scan -> pack.
//
// Work-item t of a 256-work-item block serves lane t%32 of vector
// blockIdx.x*8 + t/32.  Loads are coalesced: lane l reads rows 32j + l
// (j = 0..31), i.e. element l of every 32-row chain j -- FastLanes delta's
// chains are 32 consecutive rows, chain c owned in the packed format by the
// lane l with 2*(l%16) + l/16 = c, its element j in register row rsum_row(j)
// (rsum.cuh:8-106).  Chain deltas are therefore neighbour-lane differences
// (one shuffle), and the pack step moves chains to their owning lanes
// through a padded LDS tile.  Every exchange stays inside one wave.
#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_encode.hpp"
#include "fls_rng.hpp"
#include "fls_runtime.hpp"
#include "fls_select.hpp"

namespace fls {
namespace {

constexpr int kBlock = 256;
constexpr int kVecPerBlock = kBlock / 32;

__device__ __forceinline__ uint32_t wid(uint32_t x) { return x ? 32u - __clz(x) : 0u; }
// Half-wave (32-lane) reductions.  Within a 16-lane row DPP moves data at
// VALU speed (quad_perm swaps at distance 1 and 2, then row_ror:4 and
// row_ror:8 -- after the quad steps each lane holds its quad, and each
// rotation doubles the span, so every row lane ends with the row total, for
// sums too); one lane exchange then joins the two rows of the half-wave.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowRor4 = 0x124, kRowRor8 = 0x128;
constexpr int kWaveShr1 = 0x138;  // lane i <- lane i-1

template <typename Op>
__device__ __forceinline__ uint32_t half_reduce(uint32_t x, Op op) {
    x = op(x, dpp<kQuadXor1>(x));
    x = op(x, dpp<kQuadXor2>(x));
    x = op(x, dpp<kRowRor4>(x));
    x = op(x, dpp<kRowRor8>(x));
    return op(x, (uint32_t)__shfl_xor(x, 16, 64));
}
__device__ __forceinline__ uint32_t h_or(uint32_t x) {
    return half_reduce(x, [](uint32_t a, uint32_t b) { return a | b; });
}
__device__ __forceinline__ uint32_t h_min(uint32_t x) {
    return half_reduce(x, [](uint32_t a, uint32_t b) { return min(a, b); });
}
__device__ __forceinline__ uint32_t h_max(uint32_t x) {
    return half_reduce(x, [](uint32_t a, uint32_t b) { return max(a, b); });
}
__device__ __forceinline__ uint32_t h_sum(uint32_t x) {
    return half_reduce(x, [](uint32_t a, uint32_t b) { return a + b; });
}

__device__ __forceinline__ uint32_t chain_slot(uint32_t l) { return 2u * (l & 15u) + (l >> 4); }

__device__ __forceinline__ uint32_t padded(const uint32_t* __restrict in, uint64_t n, uint64_t k) {
    return in[k < n ? k : n - 1];
}

// Coalesced load of vector v: x[j] = row 32j + l (rows past n repeat row n-1).
__device__ __forceinline__ void load_rows(const uint32_t* __restrict in, uint64_t n, uint64_t v, uint32_t (&x)[32]) {
    const uint64_t base = v * 1024u + dev::lane32();
    if (v * 1024u + 1024u <= n) {
#pragma unroll
        for (int j = 0; j < 32; ++j) x[j] = __builtin_nontemporal_load(in + base + 32u * j);
    } else {
#pragma unroll
        for (int j = 0; j < 32; ++j) x[j] = padded(in, n, base + 32u * j);
    }
}

// Statistics of the vector from the coalesced layout (half-wave reductions).
// ALL: every encoding's; otherwise only those of encoding `enc`.
template <bool ALL>
__device__ __forceinline__ VecStats analyze_rows(const uint32_t (&x)[32], uint32_t enc) {
    const uint32_t l = dev::lane32();
    const bool need_or = ALL || enc == MIX_BITPACK;
    const bool need_range = ALL || enc == MIX_FOR || enc == MIX_RLE;
    const bool need_delta = ALL || enc == MIX_DELTA;
    const bool need_runs = ALL || enc == MIX_RLE;
    uint32_t orv = 0, mn = 0xffffffffu, mx = 0, dor = 0, starts = 0, inner = 0, last = 0;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        orv |= x[j];
        mn = min(mn, x[j]);
        mx = max(mx, x[j]);
        if (need_delta || need_runs) {
            const uint32_t up = dpp<kWaveShr1>(x[j]);  // element l-1 of chain j (l >= 1)
            if (l) {
                dor |= x[j] - up;
                const uint32_t st = x[j] != up;
                starts += st;
                inner |= st;
            } else {
                starts += j ? (uint32_t)(x[j] != last) : 1u;  // across the chain boundary
            }
            if (need_runs) {  // element 31 of chain j, of this half-wave's vector
                const uint32_t e0 = (uint32_t)__builtin_amdgcn_readlane((int)x[j], 31);
                const uint32_t e1 = (uint32_t)__builtin_amdgcn_readlane((int)x[j], 63);
                last = (threadIdx.x & 32u) ? e1 : e0;
            }
        }
    }
    VecStats st{};
    if (need_or) st.w_bp = wid(h_or(orv));
    if (need_range) {
        st.vmin = h_min(mn);
        st.w_for = wid(h_max(mx) - st.vmin);
    }
    if (need_delta) st.w_delta = wid(h_or(dor));
    if (need_runs) {
        st.n_runs = h_sum(starts);
        st.ibw = h_or(inner);
    }
    return st;
}

__device__ __forceinline__ uint32_t enc_width(uint32_t enc, const VecStats& st) {
    return enc == MIX_BITPACK ? st.w_bp : enc == MIX_FOR ? st.w_for : enc == MIX_DELTA ? st.w_delta : st.w_for;
}

__device__ __forceinline__ uint64_t enc_aux(uint32_t enc, const VecStats& st) {
    if (enc == MIX_FOR) return st.vmin;
    if (enc == MIX_RLE) return (uint64_t)st.vmin | ((uint64_t)st.n_runs << 32) | ((uint64_t)st.ibw << 48);
    return 0;
}

// smallest of the four encodings (ties to the lower encoding number)
__device__ __forceinline__ uint32_t best_encoding(const VecStats& st) {
    uint32_t best = 0, best_size = mix_bytes(0, st);
    for (uint32_t e = 1; e < 4; ++e) {
        const uint32_t b = mix_bytes(e, st);
        if (b < best_size) {
            best_size = b;
            best = e;
        }
    }
    return best;
}

// ----------------------------------------------------------------- kernels
// exhaustive search over the 40 classes + the four sizes
__global__ __launch_bounds__(kBlock) void exhaustive_kernel(const uint32_t* __restrict in, uint64_t n,
                                                            uint64_t n_vec, uint8_t* __restrict cls,
                                                            uint32_t* __restrict sizes4) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    if (v >= n_vec) return;
    uint32_t x[32];
    load_rows(in, n, v, x);
    const VecStats st = analyze_rows<true>(x, 0);
    if (dev::lane32() == 0) {
        uint32_t best = 0, best_size = 0xffffffffu;
        for (uint32_t e = 0; e < 4; ++e) {
            const uint32_t b = mix_bytes(e, st);
            if (sizes4) sizes4[4 * v + e] = b;
            if (b < best_size) {
                best_size = b;
                best = e;
            }
        }
        cls[v] = (uint8_t)mix_class(best, enc_width(best, st));
    }
}

// 16 features of a 1/8 sample: four 32-row (128-byte, one memory line)
// segments, lines 0, 10, 20 and 31 of the vector (line k*31/3), so the sample
// reads 1/8 of the vector's lines and includes its first and last rows (a
// sorted vector's sample spans its whole range); lane l takes rows
// 32*((l/8)*31/3) + 4*(l%8) + 0..3
// the 16 features of one vector from its sample (4 rows per lane)
__device__ __forceinline__ void vector_features(const uint32_t (&s)[4], float* __restrict f) {
    uint32_t orv = 0, mn = 0xffffffffu, mx = 0, dor = 0, zero = 0, nondec = 0, absmax = 0, dsum = 0, big = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        orv |= s[j];
        mn = min(mn, s[j]);
        mx = max(mx, s[j]);
        if (j) {
            const uint32_t d = s[j] - s[j - 1];
            const uint32_t a = (int32_t)d < 0 ? 0u - d : d;
            dor |= d;
            zero += d == 0;
            nondec += s[j] >= s[j - 1];
            absmax = max(absmax, a);
            dsum += min(a, 0xffffu);
            big += a >= 256u;
        }
    }
    const uint32_t all_eq = (s[0] == s[1]) & (s[1] == s[2]) & (s[2] == s[3]);
    const uint32_t g_or = h_or(orv), g_min = h_min(mn), g_max = h_max(mx), g_dor = h_or(dor);
    const uint32_t g_zero = h_sum(zero), g_nondec = h_sum(nondec), g_abs = h_max(absmax), g_dsum = h_sum(dsum);
    const uint32_t g_big = h_sum(big), g_eq = h_sum(all_eq);
    const uint32_t first_min = h_min(s[0]), first_max = h_max(s[0]);
    const uint32_t at_min = h_sum((uint32_t)(s[0] == g_min) + (s[1] == g_min) + (s[2] == g_min) + (s[3] == g_min));
    const uint32_t wsum = h_sum(wid(s[0]) + wid(s[1]) + wid(s[2]) + wid(s[3]));
    if (dev::lane32() == 0) {
        f[0] = wid(g_or) / 32.f;
        f[1] = wid(g_max - g_min) / 32.f;
        f[2] = wid(g_dor) / 32.f;
        f[3] = g_zero / 96.f;
        f[4] = g_nondec / 96.f;
        f[5] = wid(g_abs) / 32.f;
        f[6] = at_min / 128.f;
        f[7] = wid(g_dsum / 96u) / 32.f;
        f[8] = wid(first_max - first_min) / 32.f;
        f[9] = g_eq / 32.f;
        f[10] = wid(g_min) / 32.f;
        f[11] = g_big / 96.f;
        f[12] = wsum / (128.f * 32.f);
        f[13] = wid(g_max) / 32.f;
        f[14] = (float)(wid(g_max - g_min) > wid(g_dor));
        f[15] = 1.0f;
    }
}

// A half-wave takes kFeatVec consecutive vectors and issues all their sample
// loads (16 independent 16-byte reads per lane) before reducing any of them:
// one vector's 512 sampled bytes alone are too few to cover HBM latency.
constexpr int kFeatVec = 4;

__global__ __launch_bounds__(kBlock) void features_kernel(const uint32_t* __restrict in, uint64_t n, uint64_t n_vec,
                                                          float* __restrict feat) {
    const uint64_t v0 = ((uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5)) * kFeatVec;
    if (v0 >= n_vec) return;  // uniform over the half-wave, which is all the reductions use
    const uint32_t l = dev::lane32();
    const uint32_t off = 32u * (((l >> 3) * 31u) / 3u) + 4u * (l & 7u);
    uint32_t s[kFeatVec][4];
#pragma unroll
    for (int k = 0; k < kFeatVec; ++k) {
        const uint64_t v = v0 + k < n_vec ? v0 + k : n_vec - 1;
#pragma unroll
        for (int j = 0; j < 4; ++j) s[k][j] = padded(in, n, v * 1024u + off + j);
    }
#pragma unroll
    for (int k = 0; k < kFeatVec; ++k)
        if (v0 + k < n_vec) vector_features(s[k], feat + 16 * (v0 + k));
}

// plan the chosen encoding of every vector: words and metadata.  With a
// confidence array, a vector whose confidence is below conf_min is planned by
// the exhaustive search instead (all four encodings' statistics, from the
// same registers: the fallback costs no extra memory traffic) and its class
// is rewritten.
__global__ __launch_bounds__(kBlock) void mixed_plan_kernel(const uint32_t* __restrict in, uint64_t n, uint64_t n_vec,
                                                            uint8_t* __restrict cls, const float* __restrict conf,
                                                            float conf_min, uint32_t* __restrict words,
                                                            uint32_t* __restrict encw, uint64_t* __restrict aux,
                                                            uint32_t* __restrict fallbacks) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    if (v >= n_vec) return;
    uint32_t x[32];
    load_rows(in, n, v, x);
    const bool fallback = conf && conf[v] < conf_min;
    uint32_t enc;
    VecStats st;
    if (fallback) {
        st = analyze_rows<true>(x, 0);
        enc = best_encoding(st);
    } else {
        enc = mix_class_enc(cls[v]);
        st = analyze_rows<false>(x, enc);
    }
    if (dev::lane32() == 0) {
        const uint32_t w = enc_width(enc, st);
        words[v] = (mix_bytes(enc, st) - 16u) / 4u;
        encw[v] = (enc << 6) | w;
        aux[v] = enc_aux(enc, st);
        if (fallback) {
            cls[v] = (uint8_t)mix_class(enc, w);
            if (fallbacks) atomicAdd(fallbacks, 1u);
        }
    }
}

// pack every vector with its planned encoding.  The vector is loaded
// coalesced, moved to chain-per-lane order through a padded LDS tile
// (32 x 33 words: row j of chain c at 33c + j, conflict-free both ways), and
// packed from registers.  RLE additionally needs the prefix of run starts
// over the chains before the lane's and the run values, both through LDS.
// A vector lives in one half-wave, so every exchange is wave-local: the
// kernel has no block barrier.
constexpr int kTile = 32 * 33;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct PackLds {
    uint32_t tile[kVecPerBlock][kTile];  // transpose tile, then the RLE run values
    uint32_t cnt[kVecPerBlock][32];
};

// Pack one vector (this half-wave's) with encoding enc, width w and aux ax at
// word offset off of out, from its rows in the coalesced layout (x[j] = row
// 32j + l).  tile / cnt are the half-wave's LDS slots; every half-wave of the
// wave calls this together (wave-level exchanges), live or not.
__device__ __forceinline__ void pack_vector(const uint32_t (&r)[32], bool live, uint32_t enc, uint32_t w,
                                            uint64_t off, uint64_t ax, uint32_t* __restrict tile,
                                            uint32_t* __restrict cnt, uint32_t* __restrict out) {
    const uint32_t l = dev::lane32();
    const uint32_t c = chain_slot(l);  // the chain this lane packs
    uint32_t x[32];
    if (live) {
#pragma unroll
        for (int j = 0; j < 32; ++j) tile[33 * j + l] = r[j];  // element l of chain j
    }
    wave_sync();
    uint32_t before = 0;
    if (live) {
#pragma unroll
        for (int j = 0; j < 32; ++j) x[j] = tile[33 * c + j];
        before = c ? tile[33 * (c - 1) + 31] : 0u;
    }
    // RLE: run starts of this chain, then the index of its first run
    uint32_t rows[32];
    uint32_t first_idx = 0;
    const bool rle = live && enc == MIX_RLE;
    if (rle) {
        uint32_t starts = c == 0 ? 1u : (uint32_t)(x[0] != before);
#pragma unroll
        for (int j = 1; j < 32; ++j) starts += x[j] != x[j - 1];
        cnt[c] = starts;
    }
    wave_sync();  // cnt written; every lane has read its chain from the tile
    if (rle) {
        uint32_t before_cnt = 0;
        for (uint32_t k = 0; k < c; ++k) before_cnt += cnt[k];
        const uint32_t s0 = c == 0 ? 1u : (uint32_t)(x[0] != before);
        first_idx = before_cnt + s0 - 1u;
        uint32_t idx = first_idx;
        tile[idx] = x[0];  // run values reuse the tile (idx < 1024)
#pragma unroll
        for (int j = 1; j < 32; ++j) {
            const uint32_t st = x[j] != x[j - 1];
            idx += st;
            rows[dev::rsum_row(j)] = st;
            if (st) tile[idx] = x[j];
        }
        rows[0] = 0u;
    }
    wave_sync();  // run values in place
    if (!live) return;
    uint32_t* o = out + off;
    if (enc == MIX_BITPACK || enc == MIX_FOR) {
        const uint32_t base = enc == MIX_FOR ? (uint32_t)ax : 0u;
#pragma unroll
        for (int j = 0; j < 32; ++j) rows[dev::rsum_row(j)] = x[j] - base;
        pack_rows(rows, o, w);
    } else if (enc == MIX_DELTA) {
        o[l] = x[0];
        rows[0] = 0u;
#pragma unroll
        for (int j = 1; j < 32; ++j) rows[dev::rsum_row(j)] = x[j] - x[j - 1];
        pack_rows(rows, o + 32, w);
    } else {
        const uint32_t ibw = (uint32_t)(ax >> 48) & 1u, nr = (uint32_t)(ax >> 32) & 0xffffu;
        const uint32_t vbase = (uint32_t)ax;
        o[l] = first_idx;
        pack_rows(rows, o + 32, ibw);
        uint32_t* vo = o + 32 + 32 * ibw;
        const uint32_t n_words = (nr * w + 31u) / 32u;
        for (uint32_t wi = l; wi < n_words; wi += 32u) {
            uint32_t word = 0;
            const uint32_t bit0 = 32u * wi;
            const uint32_t k0 = bit0 / w, k1 = min(nr - 1u, (bit0 + 31u) / w);
            for (uint32_t k = k0; k <= k1; ++k) {
                const uint32_t kb = k * w;
                const uint32_t val = tile[k] - vbase;
                if (kb >= bit0) {
                    if (kb - bit0 < 32u) word |= val << (kb - bit0);
                } else {
                    word |= val >> (bit0 - kb);
                }
            }
            vo[wi] = word;
        }
    }
}

__global__ __launch_bounds__(kBlock) void mixed_pack_kernel(const uint32_t* __restrict in, uint64_t n, uint64_t n_vec,
                                                            const fls_vmeta* __restrict vmeta,
                                                            const uint64_t* __restrict aux,
                                                            uint32_t* __restrict out) {
    __shared__ PackLds lds;
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const int slot = threadIdx.x >> 5;
    const bool live = v < n_vec;
    uint32_t r[32];
    uint32_t enc = 0, w = 0;
    uint64_t off = 0, ax = 0;
    if (live) {
        load_rows(in, n, v, r);
        const fls_vmeta m = vmeta[v];
        enc = (uint32_t)(m >> 6) & 3u;
        w = (uint32_t)m & 63u;
        off = FLS_VMETA_OFFSET(m);
        ax = aux[v];
    }
    pack_vector(r, live, enc, w, off, ax, lds.tile[slot], lds.cnt[slot], out);
}

// Plan of one vector (this half-wave's): its encoding (the class's, or the
// exhaustive search's when its confidence is below conf_min), statistics and
// width.
struct VecPlan {
    uint32_t enc, w;
    VecStats st;
    bool fallback;
};
__device__ __forceinline__ VecPlan plan_vector(const uint32_t (&x)[32], uint8_t c, const float* __restrict conf,
                                               float conf_min, uint64_t v) {
    VecPlan p;
    p.fallback = conf && conf[v] < conf_min;
    if (p.fallback) {
        p.st = analyze_rows<true>(x, 0);
        p.enc = best_encoding(p.st);
    } else {
        p.enc = mix_class_enc(c);
        p.st = analyze_rows<false>(x, p.enc);
    }
    p.w = enc_width(p.enc, p.st);
    return p;
}

// Single-pass mixed encode: plan, offsets and pack in one launch (after a
// fill of its look-back state).  At 120M rows it measured slower than plan +
// scan + pack (0.20-0.24 ms against 0.15-0.20 ms per column): the plan's
// work joins the pack's in blocks held to the pack's LDS occupancy, and
// without the look-back at all (offsets at fixed slots, a timing variant) it
// was still slower, so the encoders default to plan + pack.  Blocks take
// their place in the output in the order they start (a ticket drawn from
// state[0]), publish their 8 vectors' word count in state[2 + b], and find
// their offset by looking back over the blocks before them: a predecessor's
// aggregate (kAgg) adds to the running prefix and the walk goes on, its
// inclusive prefix (kInc) ends it (decoupled look-back).  The first wave reads
// 64 predecessors per round trip (lane i the i-th nearest), so a walk is
// usually one round trip.  A block waits only on blocks that drew their
// ticket before it, and every block publishes its aggregate before it waits
// on anything, so each walk ends; a walk that polls kSpinMax rounds without
// an answer stops and sets state[1] instead of spinning on.  state holds
// n_blocks + 2 words, zeroed before the launch.
constexpr uint64_t kAgg = 1ull << 62, kInc = 2ull << 62, kPrefixMask = (1ull << 62) - 1u;
constexpr uint32_t kSpinMax = 1u << 26;

struct EncodeLds {
    PackLds pack;
    uint32_t words[kVecPerBlock];
    uint64_t base;
    uint32_t ticket;
};

__global__ __launch_bounds__(kBlock) void mixed_encode_kernel(const uint32_t* __restrict in, uint64_t n,
                                                              uint64_t n_vec, uint8_t* __restrict cls,
                                                              const float* __restrict conf, float conf_min,
                                                              fls_vmeta* __restrict vmeta, uint64_t* __restrict aux,
                                                              uint32_t* __restrict out, uint64_t* __restrict total,
                                                              uint32_t* __restrict fallbacks,
                                                              unsigned long long* __restrict state) {
    __shared__ EncodeLds lds;
    if (threadIdx.x == 0) lds.ticket = atomicAdd(reinterpret_cast<unsigned int*>(state), 1u);
    __syncthreads();
    const uint64_t b = lds.ticket;
    const int slot = threadIdx.x >> 5;
    const uint64_t v = b * kVecPerBlock + slot;
    const bool live = v < n_vec;
    uint32_t r[32];
    VecPlan p{};
    if (live) {
        load_rows(in, n, v, r);
        p = plan_vector(r, cls[v], conf, conf_min, v);
    }
    const uint32_t words = live ? (mix_bytes(p.enc, p.st) - 16u) / 4u : 0u;
    if (dev::lane32() == 0) lds.words[slot] = words;
    __syncthreads();
    if (threadIdx.x < 64) {  // the first wave: publish, look back, publish
        const uint32_t lane = threadIdx.x;
        uint32_t local = 0;
#pragma unroll
        for (int k = 0; k < kVecPerBlock; ++k) local += lds.words[k];
        unsigned long long* status = state + 2;
        uint64_t prefix = 0;
        if (b == 0) {
            if (lane == 0) atomicExch(&status[0], kInc | local);
        } else {
            if (lane == 0) atomicExch(&status[b], kAgg | local);
            int64_t top = (int64_t)b - 1;  // nearest predecessor of this round
            uint32_t rounds = 0;
            while (true) {
                const int64_t q = top - (int64_t)lane;
                // read at the coherence point; before block 0: an inclusive 0
                const uint64_t st = q >= 0 ? atomicAdd(&status[q], 0ull) : kInc;
                const uint64_t unpub = __ballot(st == 0), inc = __ballot((st & kInc) != 0);
                const uint32_t first_inc = inc ? (uint32_t)__ffsll((long long)inc) - 1u : 64u;
                const uint32_t first_unpub = unpub ? (uint32_t)__ffsll((long long)unpub) - 1u : 64u;
                if (first_unpub > first_inc || (first_unpub == 64u && first_inc == 64u)) {
                    // lanes 0..first_inc (or all 64 aggregates) are published
                    const uint64_t mine = lane <= first_inc ? (st & kPrefixMask) : 0ull;
                    prefix += dev::wave_sum(mine);
                    if (first_inc < 64u) break;
                    top -= 64;
                    continue;
                }
                if (++rounds == kSpinMax) {  // not expected: give up rather than spin on
                    if (lane == 0) atomicExch(state + 1, 1ull);
                    break;
                }
            }
            if (lane == 0) atomicExch(&status[b], kInc | (prefix + local));
        }
        if (lane == 0) {
            lds.base = prefix;
            if (b == ceil_div(n_vec, kVecPerBlock) - 1) *total = prefix + local;
        }
    }
    __syncthreads();
    uint64_t off = lds.base;
    for (int k = 0; k < slot; ++k) off += lds.words[k];
    if (live && dev::lane32() == 0) {
        vmeta[v] = FLS_VMETA(off, (p.enc << 6) | p.w);
        aux[v] = enc_aux(p.enc, p.st);
        if (p.fallback) {
            cls[v] = (uint8_t)mix_class(p.enc, p.w);
            if (fallbacks) atomicAdd(fallbacks, 1u);
        }
    }
    pack_vector(r, live, p.enc, p.w, off, enc_aux(p.enc, p.st), lds.pack.tile[slot], lds.pack.cnt[slot], out);
}

__global__ __launch_bounds__(kBlock) void mixed_decode_kernel(const uint32_t* __restrict words,
                                                              const fls_vmeta* __restrict vmeta,
                                                              const uint64_t* __restrict aux, uint64_t n_vec,
                                                              uint32_t* __restrict out) {
    __shared__ uint32_t stage[kVecPerBlock][1024];
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    const bool live = v < n_vec;
    ColDev c{FLS_ENC_MIXED, 0, 0, 1, words, vmeta, nullptr, aux};
    uint32_t x[32];
    decode_mixed(c, v, stage[threadIdx.x >> 5], live, x);
    if (!live) return;
    uint32_t* o = out + v * 1024u + dev::lane32();
#pragma unroll
    for (int r = 0; r < 32; ++r) o[32 * r] = x[r];
}

// Try-all-encodings search, last two steps: every vector has been encoded
// four times (candidate columns cand[e]); keep the smallest (ties to the
// lower encoding number) ...
struct Candidates {
    const fls_vmeta* vmeta[4];
    const uint64_t* aux[4];
    const uint64_t* total[4];
    const uint32_t* words[4];
};

__global__ void trial_choose_kernel(Candidates c, uint64_t n_vec, uint8_t* __restrict cls, uint32_t* __restrict words,
                                    uint32_t* __restrict encw, uint64_t* __restrict aux, uint8_t* __restrict pick) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n_vec; v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t best = 0;
        uint64_t best_size = ~0ull;
        for (uint32_t e = 0; e < 4; ++e) {
            const uint64_t end = v + 1 < n_vec ? FLS_VMETA_OFFSET(c.vmeta[e][v + 1]) : *c.total[e];
            const uint64_t size = end - FLS_VMETA_OFFSET(c.vmeta[e][v]);
            if (size < best_size) {
                best_size = size;
                best = e;
            }
        }
        const uint32_t ew = (uint32_t)c.vmeta[best][v] & 0xffu;
        words[v] = (uint32_t)best_size;
        encw[v] = ew;
        aux[v] = c.aux[best][v];
        pick[v] = (uint8_t)best;
        cls[v] = (uint8_t)mix_class(best, ew & 63u);
    }
}

// ... and copy each vector's chosen encoding to its place in the output.
__global__ __launch_bounds__(kBlock) void trial_gather_kernel(Candidates c, uint64_t n_vec,
                                                              const uint8_t* __restrict pick,
                                                              const fls_vmeta* __restrict vmeta,
                                                              uint32_t* __restrict out) {
    const uint64_t v = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
    if (v >= n_vec) return;
    const uint32_t e = pick[v];
    const uint64_t src = FLS_VMETA_OFFSET(c.vmeta[e][v]);
    const uint64_t end = v + 1 < n_vec ? FLS_VMETA_OFFSET(c.vmeta[e][v + 1]) : *c.total[e];
    const uint64_t dst = FLS_VMETA_OFFSET(vmeta[v]);
    for (uint64_t k = dev::lane32(); k < end - src; k += 32u) out[dst + k] = c.words[e][src + k];
}

__global__ void fill_u8_kernel(uint8_t* __restrict p, uint64_t n, uint8_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// Training vectors for the selector: vector v belongs to one of four
// families (chosen from the seed), all integer arithmetic:
//  0 uniform: base + U[0, 2^bw), bw U{1..32}, base 0 or U[0,2^31) (half each)
//  1 Zipf s=1.1 over ranks 1..4096 (inverse-CDF table): base + rank * stride,
//    stride 2^U{0..11}, base U[0,2^20)
//  2 sorted + Gaussian noise sigma 16: start + i*step + noise, step U[0,256),
//    start U[0,2^24); noise = 16 * (sum of 12 U[0,2^16) - 6*2^16) / 2^16
//    (Irwin-Hall), clamped at 0
//  3 runs of geometric length (mean 32): a run starts at row i with
//    probability 1/32, run value U[0, 2^bw), bw U{1..32}
__global__ void train_vectors_kernel(uint64_t seed, uint64_t v0, uint64_t n_vec, const uint32_t* __restrict zipf_cdf,
                                     uint32_t* __restrict out) {
    const uint64_t kf = rng::stream_key(seed, rng::kTrainFamily), kv = rng::stream_key(seed, rng::kTrainValues),
                   kn = rng::stream_key(seed, rng::kTrainNoise), kr = rng::stream_key(seed, rng::kTrainRuns);
    const uint64_t total = n_vec * 1024u;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = v0 + e / 1024u, i = e % 1024u, g = v * 1024u + i;
        const uint32_t fam = rng::uniform(kf, 8 * v, 4);
        uint32_t x;
        if (fam == 0) {
            const uint32_t bw = 1u + rng::uniform(kf, 8 * v + 1, 32);
            const uint32_t base = rng::uniform(kf, 8 * v + 2, 2) ? rng::uniform(kf, 8 * v + 3, 1ull << 31) : 0u;
            x = base + rng::uniform(kv, g, 1ull << bw);
        } else if (fam == 1) {
            const uint32_t u = rng::uniform(kv, g, 1ull << 32);
            uint32_t lo = 0, hi = 4095;  // first rank whose CDF exceeds u
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (zipf_cdf[mid] > u) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t stride = 1u << rng::uniform(kf, 8 * v + 1, 12);
            x = rng::uniform(kf, 8 * v + 2, 1u << 20) + (lo + 1u) * stride;
        } else if (fam == 2) {
            const uint32_t step = rng::uniform(kf, 8 * v + 1, 256);
            const int64_t start = rng::uniform(kf, 8 * v + 2, 1u << 24);
            int64_t sum = 0;
            for (int k = 0; k < 12; ++k) sum += rng::uniform(kn, 12 * g + k, 1u << 16);
            const int64_t noise = (16 * (sum - 6 * 65536)) / 65536;
            const int64_t val = start + (int64_t)i * step + noise;
            x = (uint32_t)(val < 0 ? 0 : val);
        } else {
            const uint32_t bw = 1u + rng::uniform(kf, 8 * v + 1, 32);
            uint64_t j = i;
            while (j > 0 && rng::uniform(kr, v * 1024u + j, 32) != 0) --j;
            x = rng::uniform(kv, v * 1024u + j, 1ull << bw);
        }
        out[e] = x;
    }
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_mixed_decode_u32(const uint32_t* d_words, const fls_vmeta* d_vmeta, const uint64_t* d_aux, uint64_t n_vec,
                         uint32_t* d_out, void* stream) {
    using namespace fls;
    if (n_vec == 0) return FLS_OK;
    FLS_RETURN_IF(!d_words || !d_vmeta || !d_aux || !d_out, FLS_ERR_NULL);
    hipLaunchKernelGGL(mixed_decode_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0,
                       as_stream(stream), d_words, d_vmeta, d_aux, n_vec, d_out);
    return check_launch();
}

int fls_gen_train_vectors(uint64_t seed, uint64_t first_vector, uint64_t n_vec, const uint32_t* d_zipf_cdf,
                          uint32_t* d_out, void* stream) {
    using namespace fls;
    if (n_vec == 0) return FLS_OK;
    FLS_RETURN_IF(!d_zipf_cdf || !d_out, FLS_ERR_NULL);
    const uint64_t blocks = ceil_div(n_vec * 1024u, 256);
    hipLaunchKernelGGL(train_vectors_kernel, dim3((uint32_t)(blocks < 16384 ? blocks : 16384)), dim3(256), 0,
                       as_stream(stream), seed, first_vector, n_vec, d_zipf_cdf, d_out);
    return check_launch();
}

int fls_exhaustive_select(const uint32_t* d_in, uint64_t n, uint8_t* d_class, uint32_t* d_sizes4, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_class, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    hipLaunchKernelGGL(exhaustive_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0,
                       as_stream(stream), d_in, n, n_vec, d_class, d_sizes4);
    return check_launch();
}

int fls_vector_features(const uint32_t* d_in, uint64_t n, float* d_feat, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_feat, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    hipLaunchKernelGGL(features_kernel, dim3((uint32_t)ceil_div(n_vec, (uint64_t)kVecPerBlock * kFeatVec)),
                       dim3(kBlock), 0, as_stream(stream), d_in, n, n_vec, d_feat);
    return check_launch();
}

uint64_t fls_mixed_max_words(uint64_t n) { return fls::ceil_div(n, 1024) * fls::kMixMaxWords; }

int fls_mixed_plan_device(const uint32_t* d_in, uint64_t n, uint8_t* d_class, const float* d_conf, float conf_min,
                          fls_vmeta* d_vmeta, uint64_t* d_aux, uint64_t* d_total_words, uint32_t* d_fallbacks,
                          void* d_scratch, void* stream) {
    using namespace fls;
    if (n == 0) return zero_u64(d_total_words, 1, as_stream(stream));
    FLS_RETURN_IF(!d_in || !d_class || !d_vmeta || !d_aux || !d_total_words || !d_scratch, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    uint32_t* encw = static_cast<uint32_t*>(d_scratch);
    uint32_t* words = encw + n_vec;
    uint64_t* bsum = reinterpret_cast<uint64_t*>(static_cast<char*>(d_scratch) + (2 * n_vec * 4 + 255) / 256 * 256);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(mixed_plan_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0, s, d_in, n,
                       n_vec, d_class, d_conf, conf_min, words, encw, d_aux, d_fallbacks);
    if (const int rc = check_launch()) return rc;
    return scan_offsets(words, encw, n_vec, bsum, d_vmeta, d_total_words, s);
}

int fls_mixed_encode_device(const uint32_t* d_in, uint64_t n, uint8_t* d_class, const float* d_conf, float conf_min,
                            fls_vmeta* d_vmeta, uint64_t* d_aux, uint32_t* d_out, uint64_t* d_total_words,
                            uint32_t* d_fallbacks, void* d_scratch, void* stream) {
    using namespace fls;
    if (n == 0) return zero_u64(d_total_words, 1, as_stream(stream));
    FLS_RETURN_IF(!d_in || !d_class || !d_vmeta || !d_aux || !d_out || !d_total_words || !d_scratch, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    const uint64_t blocks = ceil_div(n_vec, kVecPerBlock);
    FLS_RETURN_IF(blocks > 0x7fffffffu, FLS_ERR_SIZE);
    // look-back state (ticket, error flag, one status word per block) in the
    // scratch, which holds at least 8 bytes per vector
    unsigned long long* state = static_cast<unsigned long long*>(d_scratch);
    hipStream_t s = as_stream(stream);
    if (const int rc = zero_u64(reinterpret_cast<uint64_t*>(state), blocks + 2, s)) return rc;
    hipLaunchKernelGGL(mixed_encode_kernel, dim3((uint32_t)blocks), dim3(kBlock), 0, s, d_in, n, n_vec, d_class,
                       d_conf, conf_min, d_vmeta, d_aux, d_out, d_total_words, d_fallbacks, state);
    return check_launch();
}

// workspace of the try-all-encodings search: four candidate columns
// (words, vmeta, aux, total) + the plan/scan scratch + the per-vector pick
static uint64_t trial_layout(uint64_t n, uint64_t off[8]) {
    using namespace fls;
    const uint64_t n_vec = ceil_div(n, 1024);
    auto al = [](uint64_t b) { return (b + 255) / 256 * 256; };
    uint64_t o = 0;
    off[0] = o;  // words[4]
    o += 4 * al(fls_mixed_max_words(n) * 4);
    off[1] = o;  // vmeta[4]
    o += 4 * al(n_vec * 8);
    off[2] = o;  // aux[4]
    o += 4 * al(n_vec * 8);
    off[3] = o;  // total[4]
    o += al(4 * 8);
    off[4] = o;  // class per candidate
    o += al(n_vec);
    off[5] = o;  // pick
    o += al(n_vec);
    off[6] = o;  // plan/scan scratch
    o += al(fls_encode_scratch_bytes(n));
    off[7] = o;
    return o;
}

uint64_t fls_trial_workspace_bytes(uint64_t n) {
    uint64_t off[8];
    return trial_layout(n, off);
}

int fls_mixed_encode_trial(const uint32_t* d_in, uint64_t n, uint8_t* d_class, fls_vmeta* d_vmeta, uint64_t* d_aux,
                           uint32_t* d_out, uint64_t* d_total_words, void* d_work, void* stream) {
    using namespace fls;
    if (n == 0) return zero_u64(d_total_words, 1, as_stream(stream));
    FLS_RETURN_IF(!d_in || !d_class || !d_vmeta || !d_aux || !d_out || !d_total_words || !d_work, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    uint64_t off[8];
    trial_layout(n, off);
    char* w = static_cast<char*>(d_work);
    const uint64_t wstride = (off[1] - off[0]) / 4, mstride = (off[2] - off[1]) / 4;
    uint8_t* ccls = reinterpret_cast<uint8_t*>(w + off[4]);
    uint8_t* pick = reinterpret_cast<uint8_t*>(w + off[5]);
    void* scratch = w + off[6];
    hipStream_t s = as_stream(stream);
    const uint32_t vblocks = (uint32_t)ceil_div(n_vec, kVecPerBlock);
    Candidates c;
    for (int e = 0; e < 4; ++e) {
        uint32_t* cw = reinterpret_cast<uint32_t*>(w + off[0] + e * wstride);
        fls_vmeta* cm = reinterpret_cast<fls_vmeta*>(w + off[1] + e * mstride);
        uint64_t* ca = reinterpret_cast<uint64_t*>(w + off[2] + e * mstride);
        uint64_t* ct = reinterpret_cast<uint64_t*>(w + off[3]) + e;
        hipLaunchKernelGGL(fill_u8_kernel, dim3(256), dim3(256), 0, s, ccls, n_vec, (uint8_t)mix_class(e, 0));
        if (const int rc = check_launch()) return rc;
        if (const int rc = fls_mixed_plan_device(d_in, n, ccls, nullptr, 0.f, cm, ca, ct, nullptr, scratch, s)) return rc;
        hipLaunchKernelGGL(mixed_pack_kernel, dim3(vblocks), dim3(kBlock), 0, s, d_in, n, n_vec, cm, ca, cw);
        if (const int rc = check_launch()) return rc;
        c.vmeta[e] = cm;
        c.aux[e] = ca;
        c.total[e] = ct;
        c.words[e] = cw;
    }
    uint32_t* encw = static_cast<uint32_t*>(scratch);
    uint32_t* words = encw + n_vec;
    uint64_t* bsum = reinterpret_cast<uint64_t*>(static_cast<char*>(scratch) + (2 * n_vec * 4 + 255) / 256 * 256);
    hipLaunchKernelGGL(trial_choose_kernel, dim3((uint32_t)ceil_div(n_vec, 256)), dim3(256), 0, s, c, n_vec, d_class,
                       words, encw, d_aux, pick);
    if (const int rc = check_launch()) return rc;
    if (const int rc = scan_offsets(words, encw, n_vec, bsum, d_vmeta, d_total_words, s)) return rc;
    hipLaunchKernelGGL(trial_gather_kernel, dim3(vblocks), dim3(kBlock), 0, s, c, n_vec, pick, d_vmeta, d_out);
    return check_launch();
}

int fls_mixed_pack_device(const uint32_t* d_in, uint64_t n, const fls_vmeta* d_vmeta, const uint64_t* d_aux,
                          uint32_t* d_out, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_in || !d_vmeta || !d_aux || !d_out, FLS_ERR_NULL);
    const uint64_t n_vec = ceil_div(n, 1024);
    hipLaunchKernelGGL(mixed_pack_kernel, dim3((uint32_t)ceil_div(n_vec, kVecPerBlock)), dim3(kBlock), 0,
                       as_stream(stream), d_in, n, n_vec, d_vmeta, d_aux, d_out);
    return check_launch();
}

}  // extern "C"
