// Device view of an fls_column and the per-work-item column decoders shared by
// the fused query kernels.  Internal to libfls.so.
#pragma once

#include "fls/device/unpack.hpp"
#include "fls_runtime.hpp"

namespace fls {

// Loads of column words and metadata, which always live in device memory,
// through global address-space pointers.  A pointer the compiler cannot place
// (one read from memory, e.g. a segment descriptor) is generic, and loads
// through it compile to flat loads: those count on both the vector-memory and
// the LDS counters, so every wait for one also waits for every LDS access in
// flight and vice versa.  (A cast to the global space and back is folded
// away; the load itself has to go through the global-space pointer.)
template <typename T>
__device__ __forceinline__ T ldg(const T* p) {
    if constexpr (__is_scalar(T)) {
        return *(const __attribute__((address_space(1))) T*)p;
    } else {  // a plain struct of whole words, read word by word
        static_assert(sizeof(T) % 4 == 0, "struct of whole 32-bit words");
        T t;
        uint32_t* d = reinterpret_cast<uint32_t*>(&t);
        const __attribute__((address_space(1))) uint32_t* g = (const __attribute__((address_space(1))) uint32_t*)p;
#pragma unroll
        for (unsigned i = 0; i < sizeof(T) / 4; ++i) d[i] = g[i];
        return t;
    }
}
__device__ __forceinline__ uint32_t ldg_nt(const uint32_t* p) {
    return __builtin_nontemporal_load((const __attribute__((address_space(1))) uint32_t*)p);
}
template <int BW>
__device__ __forceinline__ void load_lane_g(const uint32_t* lane_in, uint32_t (&w)[BW > 0 ? BW : 1]) {
#pragma unroll
    for (int k = 0; k < BW; ++k) w[k] = ldg_nt(lane_in + 32 * k);
}

struct ColDev {
    uint32_t enc, bw, base, transposed;
    const uint32_t* words;
    const fls_vmeta* vmeta;
    const uint32_t* bases;
    const uint64_t* aux;  // RLE: fls_rle_meta per vector; MIXED: aux per vector
};

// Validate a host-side column descriptor for n_rows rows and build its device
// view.  Fixed-width columns are bounds-checked against n_words.
int to_dev_column(const fls_column& c, uint64_t n_rows, ColDev* out, bool allow_var);

// Grid of a grid-stride kernel over `work` equal units with at most
// max_blocks resident: the fewest blocks that keep the same number of rounds
// (ceil(work / max_blocks)), so every block gets the same number of units and
// the last round has no idle blocks (14,649 units on 1,536 slots: 1,465
// blocks x 10 units instead of a 10th round with 54 % of the blocks busy).
uint32_t balanced_grid(uint64_t max_blocks, uint64_t work);

// Grid size for a persistent kernel: balanced_grid(CUs * blocks_per_cu, work).
int persistent_grid(uint64_t work, uint32_t blocks_per_cu, uint32_t* grid);

// Number of compute units of the current device (cached per device).
int device_cus(uint32_t* cus);

// Resident blocks per CU of `kernel` at `block` work-items and `lds` bytes of
// dynamic LDS, queried once per (kernel, block, lds) and cached.
int blocks_per_cu(const void* kernel, int block, size_t lds, int* per_cu);

// Grid of a persistent kernel: every block resident at once (CUs x the
// kernel's occupancy), capped by the work.  A larger grid would leave the
// excess blocks to run as a tail after the first ones retire.
int resident_grid_of(const void* kernel, int block, size_t lds, uint64_t work, uint32_t* grid);

template <typename Kernel>
int resident_grid(Kernel kernel, int block, uint64_t work, uint32_t* grid) {
    return resident_grid_of(reinterpret_cast<const void*>(kernel), block, 0, work, grid);
}

// Fixed-width (BITPACK + FOR) column: BW >= 0 is a compile-time width that
// must equal c.bw; BW < 0 dispatches on c.bw at run time.
template <int BW>
__device__ __forceinline__ void decode_fixed(const ColDev& c, uint64_t v, uint32_t (&x)[32]) {
    if constexpr (BW >= 0) {
        dev::unpack_lane<BW>(c.words + v * (32u * BW), x);
    } else {
        dev::unpack_device(c.words + v * (32u * c.bw), x, c.bw);
    }
    if (c.base) {
#pragma unroll
        for (int r = 0; r < 32; ++r) x[r] += c.base;
    }
}

// FastLanes delta column: unpack at the vector's own width, then the
// in-register rsum seeded with the lane's base.
__device__ __forceinline__ void decode_delta(const ColDev& c, uint64_t v, uint32_t (&x)[32]) {
    const fls_vmeta m = c.vmeta[v];
    dev::unpack_device(c.words + FLS_VMETA_OFFSET(m), x, FLS_VMETA_BW(m));
    dev::rsum_lane(x, c.bases[v * 32u + dev::lane32()]);
}

// Value k of a linear LSB-first bit-packing at width bw (bw <= 32).
__device__ __forceinline__ uint32_t linear_extract(const uint32_t* __restrict w, uint32_t k, uint32_t bw) {
    if (bw == 0) return 0u;
    const uint32_t bit = k * bw, i = bit >> 5, s = bit & 31u;
    const uint32_t lo = ldg(w + i);
    const uint32_t hi = (s + bw > 32u) ? ldg(w + i + 1) : 0u;
    const uint32_t mask = bw >= 32 ? 0xffffffffu : ((1u << bw) - 1u);
    return __builtin_amdgcn_alignbit(hi, lo, s) & mask;
}

// Any column encoding (BITPACK, DELTA, RLE, MIXED) decoded into registers with
// no LDS and no wave barrier, for kernels whose LDS holds dimension tables.
// Every encoding reduces to one unpack at a run-time width, then a FOR add, or
// the in-register rsum (delta, and the run index of RLE), and for RLE a gather
// of the run values straight from the column (a vector's run values are at
// most 1024 packed values, so the gathers hit the cache after the first).
// Row order: normal for BITPACK unless c.transposed, transposed otherwise.
__device__ __forceinline__ void decode_regs(const ColDev& c, uint64_t v, uint32_t (&x)[32]) {
    const uint32_t l = dev::lane32();
    const uint32_t* p;
    uint32_t w, add = 0, seed = 0, nr = 0, gbase = 0, gbw = 0;
    bool rsum = false;
    const uint32_t* vals = nullptr;
    if (c.enc == FLS_ENC_BITPACK) {
        p = c.words + v * (32u * c.bw);
        w = c.bw;
        add = c.base;
    } else if (c.enc == FLS_ENC_MIXED) {
        const fls_vmeta m = c.vmeta[v];
        const uint32_t enc = (uint32_t)(m >> 6) & 3u, wm = (uint32_t)m & 63u;
        const uint64_t ax = c.aux[v];
        const uint32_t* q = c.words + FLS_VMETA_OFFSET(m);
        p = q;
        w = wm;
        if (enc == FLS_MIX_FOR) {
            add = (uint32_t)ax;
        } else if (enc == FLS_MIX_DELTA || enc == FLS_MIX_RLE) {
            rsum = true;
            seed = q[l];
            p = q + 32u;
            if (enc == FLS_MIX_RLE) {
                const uint32_t ibw = (uint32_t)(ax >> 48) & 1u, n_field = (uint32_t)(ax >> 32) & 0xffffu;
                w = ibw;
                vals = q + 32u + 32u * ibw;
                gbase = (uint32_t)ax;
                gbw = wm;
                nr = n_field < 1024u ? n_field : 1024u;
            }
        }
    } else {  // DELTA, RLE: (offset, width) per vector, 32 bases per vector
        const fls_vmeta m = c.vmeta[v];
        p = c.words + FLS_VMETA_OFFSET(m);
        w = FLS_VMETA_BW(m);
        rsum = true;
        seed = c.bases[v * 32u + l];
        if (c.enc == FLS_ENC_RLE) {
            const fls_rle_meta rm = reinterpret_cast<const fls_rle_meta*>(c.aux)[v];
            vals = p + 32u * w;
            gbase = rm.vbase;
            gbw = rm.vbw;
            nr = rm.n_runs < 1024u ? rm.n_runs : 1024u;
        }
    }
    dev::unpack_device(p, x, w);
    if (rsum) dev::rsum_lane(x, seed);
    if (vals) {
        const uint32_t last = nr ? nr - 1u : 0u;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            x[r] = gbase + linear_extract(vals, min(x[r], last), gbw);
            // one row's gather at a time: left free, the scheduler hoists all
            // 32 rows' loads (64 addresses of 64 bits) and the kernels that
            // hold other columns' codes beside this one spill
            __builtin_amdgcn_sched_barrier(0);
        }
    } else if (add) {
#pragma unroll
        for (int r = 0; r < 32; ++r) x[r] += add;
    }
}

// Keep the asm-visible value of each word live here: the loads that produced
// them are then issued (and complete) before any data-dependent branch that
// follows, so a full scan reads every byte even when a wave skips its ALU work.
template <int N>
__device__ __forceinline__ void pin_words(uint32_t (&w)[N]) {
#pragma unroll
    for (int k = 0; k < N; ++k) asm volatile("" : "+v"(w[k]));
}

// One lane of one vector of a column, held for row-by-row use.
// BW >= 0: a BITPACK column of that width: the packed words (BW registers),
// each row extracted on demand.
// BW < 0: any encoding and width, known only at run time: the 32 rows are
// decoded up front by decode_regs.
template <int BW>
struct LaneCol {
    uint32_t w[BW > 0 ? BW : (BW == 0 ? 1 : 32)];
    uint32_t base;
    __device__ __forceinline__ void load(const ColDev& c, uint64_t v) {
        if constexpr (BW >= 0) {
            base = c.base;
            dev::load_lane<BW>(c.words + v * (32u * BW) + dev::lane32(), w);
        } else {
            base = 0;
            decode_regs(c, v, w);
        }
    }
    // BW > 0 only: vector `first + slot` of the column as a wave-uniform base
    // (vector `first`, in scalar registers) and the lane's 32-bit word offset
    // from it, read through the global address space.  Each load is then one
    // scalar base, one offset register and an immediate, and its result is
    // waited for by the in-order memory counter alone: generic (flat) loads
    // also count on the LDS counter, so the first LDS read after them waits
    // for every one still in flight.
    __device__ __forceinline__ void load_at(const ColDev& c, uint64_t first, uint32_t slot) {
        static_assert(BW > 0, "fixed-width bit-packed columns");
        typedef const __attribute__((address_space(1))) uint32_t* gptr;
        base = c.base;
        const gptr p = (gptr)(c.words + first * (32u * BW));
        const uint32_t off = slot * (32u * BW) + dev::lane32();
#pragma unroll
        for (int k = 0; k < BW; ++k) w[k] = __builtin_nontemporal_load(p + off + 32u * k);
    }
    __device__ __forceinline__ uint32_t get(int r) const {
        if constexpr (BW >= 0) return dev::extract_at<BW>(w, r) + base;
        else return w[r] + base;
    }
    __device__ __forceinline__ void pin() { pin_words(w); }
};

// The column with its words pointer hidden from loop-invariant code motion.
// Loaded inside a loop, a kernel near its register limit otherwise keeps
// words + lane of every column in a VGPR pair across the loop and spills
// them; recomputing the address costs one instruction per column and step.
__device__ __forceinline__ ColDev opaque_words(ColDev c) {
    asm volatile("" : "+s"(c.words));
    return c;
}

// Row r of lane `lane` of vector v of a fixed-width bit-packed column, read
// from memory: the word that holds its first bit and, when the value crosses
// a word boundary, the next one (a late read for the few rows that survive a
// query's filters; the rest of the vector's words are never fetched).
template <int BW>
__device__ __forceinline__ uint32_t read_row(const ColDev& c, uint64_t v, uint32_t lane, uint32_t r) {
    static_assert(BW >= 0 && BW <= 32, "fixed widths only");
    if constexpr (BW == 0) {
        return c.base;
    } else {
        const uint32_t* p = c.words + v * (32u * BW) + lane;
        const uint32_t bit = r * (uint32_t)BW, k = bit >> 5, s = bit & 31u;
        const uint32_t k1 = s + BW > 32u ? k + 1u : k;  // k + 1 < BW whenever it is taken
        const uint32_t x = __builtin_amdgcn_alignbit(p[32u * k1], p[32u * k], s);
        return (BW == 32 ? x : x & dev::Mask<BW>::value) + c.base;
    }
}

// Bit key - kmin of a bitmap of n >= 1 bits, 0 for keys out of range.  The
// load reads word 0 for those and masks the result, so it is unconditional and
// a lane's 32 rows' loads issue back to back (a load under a range test
// compiles to a branch and a wait per row).
__device__ __forceinline__ uint32_t bit_at(const uint32_t* __restrict bits, uint32_t key, uint32_t kmin, uint32_t n) {
    const uint32_t k = key - kmin;
    const uint32_t kk = k < n ? k : 0u;
    return (bits[kk >> 5] >> (kk & 31u)) & (k < n ? 1u : 0u);
}

// m &= {rows r of the lane whose key's bit is set in `bits`} (bit_at per row),
// B bitmap words read before any is used: written as one bit_at per row, the
// compiler waits for each LDS read before issuing the next.
template <int B, typename Col>
__device__ __forceinline__ void bitmap_filter(const Col& col, const uint32_t* __restrict bits, uint32_t kmin,
                                              uint32_t n, uint32_t& m) {
#pragma unroll
    for (int rb = 0; rb < 32; rb += B) {
        uint32_t k[B], wv[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            k[j] = col.get(rb + j) - kmin;
            wv[j] = bits[(k[j] < n ? k[j] : 0u) >> 5];
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const uint32_t hit = k[j] < n ? (wv[j] >> (k[j] & 31u)) & 1u : 0u;
            m &= ~((hit ^ 1u) << (rb + j));
        }
    }
}

// m &= {rows r of the lane whose key's bit is set}, for a fixed-width lane of
// packed words w and a bitmap that holds a zero bit at index n: a key tests bit
// min(raw + d, n), d = the column's base - kmin (keys out of range -- below
// kmin they wrap -- read the zero bit), so a test is a min, a shift-add and a
// bit extract.  B words are read before any is used.
template <int B, int BW>
__device__ __forceinline__ void bitmap_filter_min(const uint32_t (&w)[BW > 0 ? BW : 1],
                                                  const uint32_t* __restrict bits, uint32_t d, uint32_t n,
                                                  uint32_t& m) {
    uint32_t hit = 0;
#pragma unroll
    for (int rb = 0; rb < 32; rb += B) {
        uint32_t k[B], wv[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            k[j] = min(dev::extract_at<BW>(w, rb + j) + d, n);
            uint32_t word;  // k >> 5 as an opaque shift: the compiler's (k >> 3) & ~3 costs one more
            asm("v_lshrrev_b32 %0, 5, %1" : "=v"(word) : "v"(k[j]));
            wv[j] = bits[word];
        }
#pragma unroll
        for (int j = 0; j < B; ++j) hit |= __builtin_amdgcn_ubfe(wv[j], k[j], 1u) << (rb + j);
    }
    m &= hit;
}

// bitmap_filter with the bit index of a key computed by `idx` (e.g. a
// reference hash table's slot, (key - min) % len), which must be below n
// or read as a miss.
template <int B, typename Col, typename Idx>
__device__ __forceinline__ void bitmap_filter_by(const Col& col, const uint32_t* __restrict bits, Idx idx,
                                                 uint32_t n, uint32_t& m) {
#pragma unroll
    for (int rb = 0; rb < 32; rb += B) {
        uint32_t k[B], wv[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            k[j] = idx(col.get(rb + j));
            wv[j] = bits[(k[j] < n ? k[j] : 0u) >> 5];
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const uint32_t hit = k[j] < n ? (wv[j] >> (k[j] & 31u)) & 1u : 0u;
            m &= ~((hit ^ 1u) << (rb + j));
        }
    }
}

// Row mask (bit r) of the rows of this lane that exist in a vector holding
// `live` <= 1024 rows, in normal or transposed row order.  Closed forms (no
// per-row index table, which the compiler would hoist into 32 registers):
//  * normal order: row r is original row l + 32r, so the mask is a prefix;
//  * transposed order: register row rsum_row(j) holds original row c0 + j of
//    the lane's chain (c0 = 32*(2*(l%16) + l/16)), so the live rows are the
//    first J = clamp(live - c0, 0, 32) chain elements.  Chain element j sits
//    in row 4*(j%8) + {0,2,1,3}[j/8]: four groups of eight rows spaced 4 apart.
__device__ __forceinline__ uint32_t live_row_mask(uint32_t live, bool transposed) {
    const uint32_t l = dev::lane32();
    if (!transposed) {
        const uint32_t k = live > l ? min((live - l + 31u) >> 5, 32u) : 0u;
        return k >= 32u ? 0xffffffffu : ((1u << k) - 1u);
    }
    const uint32_t c0 = 32u * (2u * (l & 15u) + (l >> 4));
    const uint32_t J = live > c0 ? min(live - c0, 32u) : 0u;
    const uint32_t q = J >> 3, s = J & 7u;
    uint32_t m = (q > 0 ? 0x11111111u : 0u) | (q > 1 ? 0x44444444u : 0u) | (q > 2 ? 0x22222222u : 0u) |
                 (q > 3 ? 0x88888888u : 0u);
    const uint32_t part = 0x11111111u & ((1u << (4u * s)) - 1u);  // s <= 7
    const uint32_t sh = q < 4 ? (0x3120u >> (4u * q)) & 3u : 0u;
    return m | (part << sh);
}

// FastLanes RLE decode of vector v, one lane, into 32 rows (transposed order).
// `stage` is this vector's 1024-entry LDS slice; the 32 lanes of the vector
// first unpack the run values into it (about 4 per lane at 128 runs), then
// each row reads its value by run index.  All 64 work-items of the wave must
// call it (it waits on the wave's own LDS writes only).
__device__ __forceinline__ void decode_rle(const ColDev& c, const fls_rle_meta* __restrict rmeta, uint64_t v,
                                           uint32_t* __restrict stage, bool live, uint32_t (&x)[32]) {
    fls_vmeta m = 0;
    fls_rle_meta rm{};
    if (live) {
        m = c.vmeta[v];
        rm = rmeta[v];
        const uint32_t* vals = c.words + FLS_VMETA_OFFSET(m) + 32u * FLS_VMETA_BW(m);
        const uint32_t nr = rm.n_runs < 1024u ? rm.n_runs : 1024u;  // the stage holds one vector's runs
        for (uint32_t k = dev::lane32(); k < nr; k += 32u) stage[k] = rm.vbase + linear_extract(vals, k, rm.vbw);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!live) return;
    dev::unpack_device(c.words + FLS_VMETA_OFFSET(m), x, FLS_VMETA_BW(m));
    dev::rsum_lane(x, c.bases[v * 32u + dev::lane32()]);
    const uint32_t last = rm.n_runs ? min((uint32_t)rm.n_runs, 1024u) - 1u : 0u;  // inside this vector's stage
#pragma unroll
    for (int r = 0; r < 32; ++r) x[r] = stage[min(x[r], last)];
}

// MIXED column: vector v decoded with its own encoding (fls.h FLS_ENC_MIXED),
// transposed order.  Like decode_rle, every work-item of the wave calls it.
__device__ __forceinline__ void decode_mixed(const ColDev& c, uint64_t v, uint32_t* __restrict stage, bool live,
                                             uint32_t (&x)[32]) {
    uint32_t enc = 0, w = 0;
    uint64_t ax = 0;
    const uint32_t* p = c.words;
    if (live) {
        const fls_vmeta m = c.vmeta[v];
        enc = (uint32_t)(m >> 6) & 3u;
        w = (uint32_t)m & 63u;
        ax = c.aux[v];
        p = c.words + FLS_VMETA_OFFSET(m);
    }
    const uint32_t ibw = (uint32_t)(ax >> 48) & 1u, nr_field = (uint32_t)(ax >> 32) & 0xffffu;
    const uint32_t nr = nr_field < 1024u ? nr_field : 1024u;  // the stage holds one vector's runs
    if (live && enc == FLS_MIX_RLE) {
        const uint32_t* vals = p + 32u + 32u * ibw;
        for (uint32_t k = dev::lane32(); k < nr; k += 32u) stage[k] = (uint32_t)ax + linear_extract(vals, k, w);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (!live) return;
    const uint32_t l = dev::lane32();
    switch (enc) {
        case FLS_MIX_BITPACK: dev::unpack_device(p, x, w); break;
        case FLS_MIX_FOR:
            dev::unpack_device(p, x, w);
#pragma unroll
            for (int r = 0; r < 32; ++r) x[r] += (uint32_t)ax;
            break;
        case FLS_MIX_DELTA:
            dev::unpack_device(p + 32, x, w);
            dev::rsum_lane(x, p[l]);
            break;
        default: {
            dev::unpack_device(p + 32, x, ibw);
            dev::rsum_lane(x, p[l]);
            const uint32_t last = nr ? nr - 1u : 0u;
#pragma unroll
            for (int r = 0; r < 32; ++r) x[r] = stage[min(x[r], last)];
        }
    }
}

// Any column encoding, width known at run time.  `stage` is this vector's
// 1024-entry LDS slice (RLE / MIXED); every work-item of the wave calls it.
__device__ __forceinline__ void decode_col(const ColDev& c, uint64_t v, uint32_t* __restrict stage, bool live,
                                           uint32_t (&x)[32]) {
    switch (c.enc) {
        case FLS_ENC_BITPACK:
            if (live) decode_fixed<-1>(c, v, x);
            break;
        case FLS_ENC_DELTA:
            if (live) decode_delta(c, v, x);
            break;
        case FLS_ENC_RLE: decode_rle(c, reinterpret_cast<const fls_rle_meta*>(c.aux), v, stage, live, x); break;
        default: decode_mixed(c, v, stage, live, x); break;
    }
}

// Predicate-fused FastLanes delta walk of one lane at compile-time width BW:
// the running sum visits the rows in chain order (rsum.cuh:8-106) and each
// decoded value is reduced to one bit of the returned row mask at once, so
// only the BW packed words stay live.
template <int BW, typename Pred>
__device__ __forceinline__ uint32_t delta_mask_lane(const uint32_t* __restrict lane_in, uint32_t base, Pred pred) {
    uint32_t w[BW > 0 ? BW : 1];
    dev::load_lane<BW>(lane_in, w);
    if constexpr (BW == 0) {
        return pred(base) ? 0xffffffffu : 0u;
    } else {
        uint32_t acc = base, m = 0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int r = dev::rsum_row(j);
            acc += dev::extract_at<BW>(w, r);
            m |= (uint32_t)pred(acc) << r;
        }
        return m;
    }
}

// Run-time-width delta walk of one lane: each row reads the (at most two)
// words it straddles, which the first touch brings into the cache, and
// extracts with one funnel shift.  Partially unrolled so that only a few row
// loads are in flight and few VGPRs are held, whatever the width.
template <typename Pred>
__device__ __forceinline__ uint32_t delta_mask_walk(const uint32_t* __restrict in, uint32_t bw, uint32_t base,
                                                    Pred pred) {
    const uint32_t mask = bw >= 32 ? 0xffffffffu : ((1u << bw) - 1u);
    uint32_t acc = base, m = 0;
#pragma unroll 4
    for (int j = 0; j < 32; ++j) {
        const int r = dev::rsum_row(j);
        const uint32_t bit = (uint32_t)r * bw, k = bit >> 5, s = bit & 31u;
        const uint32_t lo = in[32u * k];
        const uint32_t hi = (k + 1u < bw) ? in[32u * (k + 1u)] : 0u;
        acc += __builtin_amdgcn_alignbit(hi, lo, s) & mask;
        m |= (uint32_t)pred(acc) << r;
    }
    return m;
}

// Row mask {r : pred(value of row r)} of vector v of a DELTA column.  Widths
// up to 16 take a fully unrolled per-width walk (all words in registers, one
// extract per row); wider vectors take the run-time walk, which bounds the
// VGPRs a kernel must reserve for this column at about 16 + a few.
// m = the vector's vmeta, base = this lane's delta base (already loaded, so a
// kernel can fetch them one iteration ahead).
template <typename Pred>
__device__ __forceinline__ uint32_t delta_mask_pre(const ColDev& c, fls_vmeta m, uint32_t base, Pred pred) {
    const uint32_t* in = c.words + FLS_VMETA_OFFSET(m) + dev::lane32();
    const uint32_t bw = FLS_VMETA_BW(m);
    switch (bw) {
#define FLS_DM_CASE(B) \
    case B: return delta_mask_lane<B>(in, base, pred);
        FLS_DM_CASE(0) FLS_DM_CASE(1) FLS_DM_CASE(2) FLS_DM_CASE(3) FLS_DM_CASE(4) FLS_DM_CASE(5) FLS_DM_CASE(6)
        FLS_DM_CASE(7) FLS_DM_CASE(8) FLS_DM_CASE(9) FLS_DM_CASE(10) FLS_DM_CASE(11) FLS_DM_CASE(12)
        FLS_DM_CASE(13) FLS_DM_CASE(14) FLS_DM_CASE(15) FLS_DM_CASE(16)
#undef FLS_DM_CASE
        default: return delta_mask_walk(in, bw, base, pred);
    }
}

template <typename Pred>
__device__ __forceinline__ uint32_t delta_mask(const ColDev& c, uint64_t v, Pred pred) {
    return delta_mask_pre(c, c.vmeta[v], c.bases[v * 32u + dev::lane32()], pred);
}

// ------------------------------------------------------------------ row walks
// One lane of one vector visited row by row: f(r, value) is called once per
// register row r (a compile-time constant after unrolling, so a caller's
// per-row arrays stay in registers), with only the packed words of the
// vector's width held -- no 32-row decoded array is needed by the walk
// itself.  Row masks are the walk with f = "set bit r if pred(value)"; a
// kernel that handles every width and encoding this way stays near 40-80
// VGPRs.

// fixed width (bit-pack, + FOR base); lane_in = first word of this lane
template <int BW, typename F>
__device__ __forceinline__ void walk_fixed_lane(const uint32_t* __restrict lane_in, uint32_t base, F& f) {
    if constexpr (BW == 0) {
#pragma unroll
        for (int r = 0; r < 32; ++r) f(r, base);
    } else {
        uint32_t w[BW];
        load_lane_g<BW>(lane_in, w);
#pragma unroll
        for (int r = 0; r < 32; ++r) f(r, dev::extract_at<BW>(w, r) + base);
    }
}

// FastLanes delta: the running sum visits the rows in chain order
template <int BW, typename F>
__device__ __forceinline__ void walk_delta_lane(const uint32_t* __restrict lane_in, uint32_t first, F& f) {
    if constexpr (BW == 0) {
#pragma unroll
        for (int r = 0; r < 32; ++r) f(r, first);
    } else {
        uint32_t w[BW];
        load_lane_g<BW>(lane_in, w);
        uint32_t acc = first;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int r = dev::rsum_row(j);
            acc += dev::extract_at<BW>(w, r);
            f(r, acc);
        }
    }
}

#define FLS_WIDTH_CASES(X)                                                                                           \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19)  \
        X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

template <typename F>
__device__ __forceinline__ void walk_fixed(const uint32_t* __restrict lane_in, uint32_t bw, uint32_t base, F& f) {
    switch (bw) {
#define FLS_WF_CASE(B)                               \
    case B: walk_fixed_lane<B>(lane_in, base, f); \
        break;
        FLS_WIDTH_CASES(FLS_WF_CASE)
#undef FLS_WF_CASE
        default: break;
    }
}

template <typename F>
__device__ __forceinline__ void walk_delta(const uint32_t* __restrict lane_in, uint32_t bw, uint32_t first, F& f) {
    switch (bw) {
#define FLS_WD_CASE(B)                                \
    case B: walk_delta_lane<B>(lane_in, first, f); \
        break;
        FLS_WIDTH_CASES(FLS_WD_CASE)
#undef FLS_WD_CASE
        default: break;
    }
}

// FastLanes RLE: the run index of the lane's rows is a delta chain of width
// ibw (0: the whole lane is one run; 1: one word, bit r = 1 where row r starts
// a run); `stage` holds the vector's run values, `last` = n_runs - 1.
template <typename F>
__device__ __forceinline__ void walk_rle(const uint32_t* __restrict idx_lane, uint32_t ibw, uint32_t first_idx,
                                         const uint32_t* __restrict stage, uint32_t last, F& f) {
    if (ibw == 0) {
        const uint32_t x = stage[min(first_idx, last)];
#pragma unroll
        for (int r = 0; r < 32; ++r) f(r, x);
        return;
    }
    const uint32_t starts = ldg_nt(idx_lane);
    uint32_t acc = first_idx;
#pragma unroll
    for (int j = 0; j < 32; ++j) {
        const int r = dev::rsum_row(j);
        acc += (starts >> r) & 1u;
        f(r, stage[min(acc, last)]);
    }
}

// Stage the n_runs (<= 1024) run values of a vector into LDS; every
// work-item of the wave calls it (`active` = this half-wave has a vector).
__device__ __forceinline__ void stage_runs(const uint32_t* __restrict vals, uint32_t n_runs, uint32_t vbw,
                                           uint32_t vbase, uint32_t* __restrict stage, bool active) {
    if (active) {
        const uint32_t nr = n_runs < 1024u ? n_runs : 1024u;
        for (uint32_t k = dev::lane32(); k < nr; k += 32u) stage[k] = vbase + linear_extract(vals, k, vbw);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Walk vector v of any column encoding (run-time width and encoding).  Every
// work-item of the wave calls it (RLE run values are staged in `stage`, this
// vector's 1024-entry LDS slice); f is called only when `live`.  `rmeta`:
// the RLE metadata (c.aux of an RLE column).
template <typename F>
__device__ __forceinline__ void column_walk(const ColDev& c, const fls_rle_meta* __restrict rmeta, uint64_t v,
                                            uint32_t* __restrict stage, bool live, F& f) {
    const uint32_t l = dev::lane32();
    if (c.enc == FLS_ENC_BITPACK) {
        if (live) walk_fixed(c.words + v * (32u * c.bw) + l, c.bw, c.base, f);
        return;
    }
    if (c.enc == FLS_ENC_DELTA) {
        if (!live) return;
        const fls_vmeta m = ldg(c.vmeta + v);
        walk_delta(c.words + FLS_VMETA_OFFSET(m) + l, FLS_VMETA_BW(m), ldg(c.bases + v * 32u + l), f);
        return;
    }
    if (c.enc == FLS_ENC_RLE) {
        fls_vmeta m = 0;
        fls_rle_meta rm{};
        if (live) {
            m = ldg(c.vmeta + v);
            rm = ldg(rmeta + v);
        }
        const uint32_t* p = c.words + FLS_VMETA_OFFSET(m);
        const uint32_t ibw = FLS_VMETA_BW(m) ? 1u : 0u;
        stage_runs(p + 32u * ibw, rm.n_runs, rm.vbw, rm.vbase, stage, live);
        if (!live) return;
        const uint32_t last = rm.n_runs ? min((uint32_t)rm.n_runs, 1024u) - 1u : 0u;
        walk_rle(p + l, ibw, ldg(c.bases + v * 32u + l), stage, last, f);
        return;
    }
    // MIXED
    uint32_t enc = 0, w = 0;
    uint64_t ax = 0;
    const uint32_t* p = c.words;
    if (live) {
        const fls_vmeta m = ldg(c.vmeta + v);
        enc = (uint32_t)(m >> 6) & 3u;
        w = (uint32_t)m & 63u;
        ax = ldg(c.aux + v);
        p = c.words + FLS_VMETA_OFFSET(m);
    }
    const uint32_t ibw = (uint32_t)(ax >> 48) & 1u, nr = (uint32_t)(ax >> 32) & 0xffffu;
    stage_runs(p + 32u + 32u * ibw, nr, w, (uint32_t)ax, stage, live && enc == FLS_MIX_RLE);
    if (!live) return;
    switch (enc) {
        case FLS_MIX_BITPACK: walk_fixed(p + l, w, 0u, f); break;
        case FLS_MIX_FOR: walk_fixed(p + l, w, (uint32_t)ax, f); break;
        case FLS_MIX_DELTA: walk_delta(p + 32u + l, w, ldg(p + l), f); break;
        default: {
            const uint32_t last = nr ? min(nr, 1024u) - 1u : 0u;
            walk_rle(p + 32u + l, ibw, ldg(p + l), stage, last, f);
        }
    }
}

// Row mask {r : pred(value of row r)} of vector v, any encoding.
template <typename Pred>
__device__ __forceinline__ uint32_t column_mask(const ColDev& c, const fls_rle_meta* __restrict rmeta, uint64_t v,
                                                uint32_t* __restrict stage, bool live, Pred pred) {
    uint32_t m = 0;
    auto f = [&](int r, uint32_t x) { m |= (uint32_t)pred(x) << r; };
    column_walk(c, rmeta, v, stage, live, f);
    return m;
}

}  // namespace fls
