// Fused decode + star joins + group-by for the Q3.1 and Q4.1 shapes
// (reference: probe_v1 and build_hashtable_* of fastlanes/src/ssb/fls_q31.cu
// and fls_q41.cu).  MI355X-native: the dimension tables are reduced once into
// a compact workspace (one byte per supplier / customer / date: its nation or
// year offset, 0xFF when the row is filtered out; one bit per part and per
// qualifying customer), the probe kernel stages the small ones in LDS and
// keeps the customer codes (600 KB) in L2, and the group-by runs in a
// per-block LDS table (4,375 or 175 slots, or a hash table) flushed once per
// block: the ~4 M (Q3.1) / ~2 M (Q4.1) qualifying rows of 120 M land in about
// 150 / 35 live slots, which global per-row atomics would serialise on.
#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_fastdiv.hpp"
#include "fls_group_hash.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

constexpr uint8_t kNone = 0xFF;
constexpr uint32_t kMaxSlots = 8192;

// workspace layout (byte offsets, 256-aligned)
struct StarLayout {
    uint64_t s_code, c_code, p_bits, y_code, c_bits, s_bits, masks, total;
    uint32_t p_words, c_words, s_words;
};

StarLayout star_layout(const fls_star_dims& d) {
    auto up = [](uint64_t x) { return (x + 255u) & ~(uint64_t)255u; };
    StarLayout L{};
    L.p_words = (d.n_part + 31u) / 32u;
    L.c_words = (d.n_cust + 31u) / 32u;
    L.s_words = (d.n_supp + 31u) / 32u;
    L.s_code = 0;
    L.c_code = up(L.s_code + d.n_supp);
    L.p_bits = up(L.c_code + d.n_cust);
    L.y_code = up(L.p_bits + 4ull * L.p_words);
    L.c_bits = up(L.y_code + d.n_date);
    L.s_bits = up(L.c_bits + 4ull * L.c_words);
    L.masks = up(L.s_bits + 4ull * L.s_words);  // Q3.1: qualifying supplier and customer nations, 8 words each
    L.total = up(L.masks + 64u);
    return L;
}

__global__ void star_zero_words(uint32_t* __restrict w, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
}

// Q3.1 builds (build_hashtable_s/_c/_d, fls_q31.cu:300-372): nation of the
// qualifying suppliers and customers (and a bit per qualifying customer;
// c_bits zeroed before), year offset of the dates in range
// (and the set of nations that qualify on each side, 256 bits each in
// `masks`, zeroed before: supplier nations in words 0-7, customer in 8-15)
__global__ __launch_bounds__(256) void q31_build_kernel(fls_star_dims d, fls_q31_params p, uint8_t* __restrict s_code,
                                                        uint8_t* __restrict c_code, uint8_t* __restrict y_code,
                                                        uint32_t* __restrict c_bits, uint32_t* __restrict masks) {
    __shared__ uint32_t nm[16];
    if (threadIdx.x < 16) nm[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_supp; i += stride) {
        const uint32_t n = d.s_nation[i];
        const bool hit = d.s_region[i] == p.s_region && n < kNone;
        s_code[i] = hit ? (uint8_t)n : kNone;
        if (hit) atomicOr(&nm[n >> 5], 1u << (n & 31u));
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_cust; i += stride) {
        const uint32_t n = d.c_nation[i];
        const bool hit = d.c_region[i] == p.c_region && n < kNone;
        c_code[i] = hit ? (uint8_t)n : kNone;
        if (hit) {
            atomicOr(&c_bits[i >> 5], 1u << (i & 31u));
            atomicOr(&nm[8u + (n >> 5)], 1u << (n & 31u));
        }
    }
    __syncthreads();
    if (threadIdx.x < 16 && nm[threadIdx.x]) atomicOr(&masks[threadIdx.x], nm[threadIdx.x]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_date; i += stride) {
        const uint32_t y = d.d_year[i];
        y_code[i] = (y >= p.year_lo && y <= p.year_hi) ? (uint8_t)(y - p.year_base) : kNone;
    }
}

// Q4.1 builds (fls_q41.cu:186-271): qualifying suppliers (code 0 and a bit),
// parts (a bit), customers (their nation), every date (its year offset);
// p_bits and s_bits zeroed before
__global__ void q41_build_kernel(fls_star_dims d, fls_q41_params p, uint8_t* __restrict s_code,
                                 uint8_t* __restrict c_code, uint32_t* __restrict p_bits,
                                 uint8_t* __restrict y_code, uint32_t* __restrict s_bits) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_supp; i += stride) {
        const bool hit = d.s_region[i] == p.s_region;
        s_code[i] = hit ? (uint8_t)0 : kNone;
        if (hit) atomicOr(&s_bits[i >> 5], 1u << (i & 31u));
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_cust; i += stride)
        c_code[i] = (d.c_region[i] == p.c_region && d.c_nation[i] < kNone) ? (uint8_t)d.c_nation[i] : kNone;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_part; i += stride)
        if (d.p_mfgr[i] >= p.mfgr_lo && d.p_mfgr[i] <= p.mfgr_hi) atomicOr(&p_bits[i >> 5], 1u << (i & 31u));
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < d.n_date; i += stride) {
        const uint32_t y = d.d_year[i] - p.year_base;  // wraps below year_base: no match
        y_code[i] = y < kNone ? (uint8_t)y : kNone;
    }
}

struct StarDev {
    ColDev date, cust, supp, part, rev, cost;
    uint64_t n_rows, n_vec, n_groups;
    const uint8_t *s_code, *c_code, *y_code;
    const uint32_t *p_bits, *c_bits, *s_bits;
    const uint32_t* masks;  // Q3.1: qualifying supplier nations (words 0-7) and customer nations (8-15)
    uint32_t n_supp, n_cust, n_part, n_date, keys_min, p_words, c_words, s_words;
    uint32_t y_span;  // Q3.1: year codes run 0 .. y_span - 1
    uint32_t slot_mul_s, slot_mul_c, n_slots;  // slot = (s*slot_mul_s + c*slot_mul_c + y) % n_slots
    FastDiv slot_div;                          // n_slots
    int64_t* sums;
    uint64_t* counts;
};

constexpr int kStarBlock = 1024;
constexpr int kB31 = 16;  // rows per batch of dimension lookups
constexpr int kStarVecs = kStarBlock / 32;

// per-block LDS: group table, then (USE_LDS) the staged dimension codes
struct StarLds {
    unsigned long long* sum;
    uint32_t* cnt;
    const uint8_t *s_code, *y_code;
    const uint32_t* p_bits;
};

template <bool USE_LDS, bool PARTS>
__device__ __forceinline__ StarLds star_stage(const StarDev& a, uint8_t* lds) {
    // 16-byte aligned (the caller's reserve, if any, is a multiple of 16)
    StarLds t;
    t.sum = reinterpret_cast<unsigned long long*>(lds);
    t.cnt = reinterpret_cast<uint32_t*>(lds + 8ull * a.n_slots);
    for (uint32_t i = threadIdx.x; i < a.n_slots; i += blockDim.x) {
        t.sum[i] = 0ull;
        t.cnt[i] = 0u;
    }
    t.s_code = a.s_code;
    t.y_code = a.y_code;
    t.p_bits = a.p_bits;
    if constexpr (USE_LDS) {
        uint8_t* q = lds + 12ull * a.n_slots;
        uint32_t* pb = reinterpret_cast<uint32_t*>(q);  // part bits first: 4-byte aligned
        if constexpr (PARTS) {
            for (uint32_t i = threadIdx.x; i < a.p_words; i += blockDim.x) pb[i] = a.p_bits[i];
            q += 4ull * a.p_words;
            t.p_bits = pb;
        }
        uint8_t* s = q;
        uint8_t* y = q + a.n_supp;
        for (uint32_t i = threadIdx.x; i < a.n_supp; i += blockDim.x) s[i] = a.s_code[i];
        for (uint32_t i = threadIdx.x; i < a.n_date; i += blockDim.x) y[i] = a.y_code[i];
        t.s_code = s;
        t.y_code = y;
    }
    __syncthreads();
    return t;
}

__device__ __forceinline__ void star_flush(const StarDev& a, const StarLds& t) {
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.n_slots; i += blockDim.x)
        if (t.cnt[i]) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.sums[i]), t.sum[i]);
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[i]), (unsigned long long)t.cnt[i]);
        }
}

// Code of a key (kNone out of range), branch-free like bit_at.  Every
// dimension holds at least one row.
__device__ __forceinline__ uint32_t code_at(const uint8_t* __restrict code, uint32_t key, uint32_t kmin, uint32_t n) {
    const uint32_t k = key - kmin;
    const uint32_t x = code[k < n ? k : 0u];
    return k < n ? x : kNone;
}

// Table entries for rows [r0, r0 + B) of a lane: table[key - kmin] for the
// rows live in m whose key is in range, `none` for the others.  Every load is
// issued (a dead row reads entry 0) before any result is used, so the B
// lookups wait for one memory round trip instead of B dependent ones.
template <int B, typename T, typename Col>
__device__ __forceinline__ void lookup_batch(const Col& col, const T* __restrict table, uint32_t kmin, uint32_t n,
                                             uint32_t m, int r0, uint32_t none, uint32_t (&out)[B]) {
    uint32_t idx[B];
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const uint32_t k = col.get(r0 + j) - kmin;
        idx[j] = (((m >> (r0 + j)) & 1u) && k < n) ? k : 0xFFFFFFFFu;
    }
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const uint32_t x = table[idx[j] != 0xFFFFFFFFu ? idx[j] : 0u];
        out[j] = idx[j] != 0xFFFFFFFFu ? x : none;
    }
}

template <int B>
constexpr uint32_t batch_mask(int r0) {
    return (B >= 32 ? 0xFFFFFFFFu : ((1u << B) - 1u)) << r0;
}

// One byte per row r of a lane, four rows per register (r compile-time).
__device__ __forceinline__ void put8(uint32_t (&p)[8], int r, uint32_t v) { p[r >> 2] |= v << (8 * (r & 3)); }
__device__ __forceinline__ uint32_t get8(const uint32_t (&p)[8], int r) { return (p[r >> 2] >> (8 * (r & 3))) & 0xFFu; }

// A lane's 32 decoded rows (any encoding, decode_regs) seen as a column.
struct RegCol {
    uint32_t x[32];
    __device__ __forceinline__ void load(const ColDev& c, uint64_t v) { decode_regs(c, v, x); }
    __device__ __forceinline__ uint32_t get(int r) const { return x[r]; }
};

// Batched code lookups of the rows live in m: each row's code is kept in
// `codes`, and rows whose code is kNone (filtered, or key out of range) leave m.
template <int B, typename Col>
__device__ __forceinline__ void filter_codes(const Col& col, const uint8_t* __restrict table, uint32_t kmin,
                                             uint32_t n, uint32_t& m, uint32_t (&codes)[8]) {
#pragma unroll
    for (int rb = 0; rb < 32; rb += B) {
        if (!__any((m & batch_mask<B>(rb)) != 0)) continue;
        uint32_t cv[B];
        lookup_batch<B>(col, table, kmin, n, m, rb, kNone, cv);
#pragma unroll
        for (int j = 0; j < B; ++j) {
            put8(codes, rb + j, cv[j]);
            if (cv[j] == kNone) m &= ~(1u << (rb + j));
        }
    }
}

__device__ __forceinline__ void group_add_lds(const StarLds& t, uint32_t slot, uint64_t value) {
    atomicAdd(&t.sum[slot], (unsigned long long)value);
    atomicAdd(&t.cnt[slot], 1u);
}

// Q3.1 probe at fixed widths, three ways of holding the dimensions:
//  kQ31Codes  group table in LDS, supplier / date codes in global memory;
//  kQ31Staged group table, supplier and date codes in LDS;
//  kQ31Bitmap supplier and date codes, a customer filter bitmap and a group
//             hash table in LDS (SSB sizes: 40 KB + 2.5 KB + 75 KB + 10 KB).
// All four columns' loads are issued before the first probe (bytes in
// flight); the supplier nation (one byte per supplier) is looked up on every
// row, and kQ31Bitmap also tests every row's customer bit in LDS.  The
// survivors (about one in five, or one in 25 with the bitmap) go to a
// per-wave LDS queue: customer key, date key | supplier nation << 24,
// revenue.  The flush takes one queued row per lane, two rows per lane in
// flight: year offset, customer nation, group.  The customer lookups are L2
// gathers of a 600 KB table, each a distinct cache line: at 24 M of them the
// L2 request rate bounds the kernel, which the bitmap cuts to the 4.8 M rows
// that pass both filters (and the queue's flushes five-fold).
enum Q31Mode { kQ31Codes = 0, kQ31Staged = 1, kQ31Bitmap = 2 };
constexpr uint32_t kQ31Queue = 128;
constexpr size_t kQ31QueueBytes = (size_t)kStarVecs / 2u * 3u * kQ31Queue * 4u;  // 16 waves, 24 KB
using Q31Hash = GroupHash<9>;  // 512 entries, 10 KB

// kQ31Codes and kQ31Staged (the kQ31Bitmap probe is q31_bitmap_kernel)
template <int MODE>
__device__ __forceinline__ void q31_flush(const StarDev& a, const StarLds& t, const uint32_t* q, uint32_t qn) {
    constexpr int kPer = kQ31Queue / 64;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t l = threadIdx.x & 63u;
    uint32_t ck[kPer], dk[kPer], cv[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        const uint32_t x = q[i];
        ck[j] = i < qn ? x : 0xFFFFFFFFu;  // rows past the queue: out of range, dropped
        dk[j] = q[kQ31Queue + i];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) cv[j] = code_at(a.c_code, ck[j], 0u, a.n_cust);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        const uint32_t y = code_at(t.y_code, dk[j] & 0xFFFFFFu, 0u, a.n_date);
        if (cv[j] != kNone && y != kNone) {
            const uint32_t slot = fast_mod((dk[j] >> 24) * a.slot_mul_s + cv[j] * a.slot_mul_c + y, a.slot_div);
            group_add_lds(t, slot, q[2u * kQ31Queue + i]);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Date keys are clamped to 24 bits in the queue (n_date < 2^24).
template <int DBW, int CBW, int SBW, int RBW, int MODE>
__global__ __launch_bounds__(kStarBlock) void q31_kernel(const StarDev a) {
    extern __shared__ __align__(16) uint8_t lds[];
    uint32_t* q = reinterpret_cast<uint32_t*>(lds) + (threadIdx.x >> 6) * 3u * kQ31Queue;
    static_assert(MODE == kQ31Codes || MODE == kQ31Staged, "kQ31Bitmap runs q31_bitmap_kernel");
    const StarLds t = star_stage<MODE == kQ31Staged, false>(a, lds + kQ31QueueBytes);
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    uint32_t qn = 0;  // wave-uniform; the queue carries over from step to step
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kStarVecs + (threadIdx.x >> 5);
        // the whole wave stays in step (the queue flush needs every lane): a
        // half-wave past the last vector reads the last vector again (in
        // bounds) and keeps no rows
        const bool vlive = v < a.n_vec;
        if (!__any(vlive)) continue;
        const uint64_t vl = vlive ? v : a.n_vec - 1;
        LaneCol<SBW> sc;
        LaneCol<CBW> cc;
        LaneCol<DBW> dc;
        LaneCol<RBW> rc;
        sc.load(a.supp, vl);
        cc.load(a.cust, vl);
        dc.load(a.date, vl);
        rc.load(a.rev, vl);
        uint32_t m = 0, s8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const uint32_t sn = code_at(t.s_code, sc.get(r), km, a.n_supp);
            put8(s8, r, sn);
            m |= (uint32_t)(sn != kNone) << r;
        }
        if (!vlive) m = 0;
        if (v == a.n_vec - 1) m &= tail;
        cc.pin();
        dc.pin();
        rc.pin();
        if (!__any(m != 0)) continue;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            // flush only when this row slot's rows would not fit: the flushes
            // run nearly full (a chain of round trips each, whatever their size)
            const bool live = (m >> r) & 1u;
            const uint64_t bal = __ballot(live);
            if (qn + (uint32_t)__popcll(bal) > kQ31Queue) {
                q31_flush<MODE>(a, t, q, qn);
                qn = 0;
            }
            if (live) {
                const uint32_t pos =
                    qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                const uint32_t dk = dc.get(r) - km;
                q[pos] = cc.get(r) - km;
                q[kQ31Queue + pos] = min(dk, 0xFFFFFFu) | (get8(s8, r) << 24);
                q[2u * kQ31Queue + pos] = rc.get(r);
            }
            qn += (uint32_t)__popcll(bal);
        }
    }
    if (qn) q31_flush<MODE>(a, t, q, qn);
    star_flush(a, t);
}

// kQ31Bitmap probe.  LDS: the row queues, the group hash table, then the
// supplier codes at a fixed offset (one byte per supplier and a kNone entry
// at index n_supp), the year codes, and the customer filter bitmap (with a
// zero bit at index n_cust).  A key k of a row tests entry min(k, n), so an
// out-of-range key (below keys_min it wraps) reads the sentinel: one min per
// lookup instead of a compare and two selects.  Rows are queued as raw
// packed fields (customer; supplier | date << SBW; revenue) and the flush
// adds the frame bases and keys_min, looks the supplier nation up again and
// the year, and gathers the customer nation from L2.
// per wave: three fields of kQ31Queue entries and a spare entry each
constexpr uint32_t kQ31Stride = kQ31Queue + 1u;
constexpr size_t kQ31BitmapQueueBytes = (size_t)kStarVecs / 2u * 3u * kQ31Stride * 4u;
constexpr size_t kQ31SupplierLds = kQ31BitmapQueueBytes + Q31Hash::kBytes;  // byte offset of the supplier codes

// Dense group table (the hash table's LDS when it fits): one entry per
// (qualifying supplier nation, qualifying customer nation, year code), ranked
// among the nations that qualify -- SSB: 5 x 5 x 6 = 150 groups, added to
// with no probing.  sum[kQ31Dense] (8 bytes), cnt[kQ31Dense], then per side
// a nation's rank (0xFF: not qualifying) and the nation of a rank.
constexpr uint32_t kQ31Dense = 760;
struct Q31Groups {
    uint32_t n_s, n_c, n_groups;
    bool dense;
    // the tables, at fixed offsets from the dynamic LDS base
    static __device__ __forceinline__ unsigned long long* sum(uint8_t* lds) {
        return reinterpret_cast<unsigned long long*>(lds + kQ31BitmapQueueBytes);
    }
    static __device__ __forceinline__ uint32_t* cnt(uint8_t* lds) {
        return reinterpret_cast<uint32_t*>(lds + kQ31BitmapQueueBytes + 8u * kQ31Dense);
    }
    // rank of a nation: side 0 supplier, 1 customer; nation of a rank: 2, 3
    static __device__ __forceinline__ uint8_t* byte_table(uint8_t* lds, uint32_t which) {
        return lds + kQ31BitmapQueueBytes + 12u * kQ31Dense + 256u * which;
    }
};
static_assert(kQ31Dense * 12u + 4u * 256u <= Q31Hash::kBytes, "dense groups share the hash table's LDS");

template <int DBW, int CBW, int SBW, int RBW>
__device__ __forceinline__ void q31_bitmap_flush(const StarDev& a, const uint8_t* s_tab, const uint8_t* y_tab,
                                                 const Q31Hash& h, const Q31Groups& gt, uint8_t* lds,
                                                 const uint32_t* q, uint32_t qn) {
    constexpr int kPer = kQ31Queue / 64;
    constexpr uint32_t kSMask = (1u << SBW) - 1u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t l = threadIdx.x & 63u;
    const uint32_t km = a.keys_min;
    uint32_t ck[kPer], sd[kPer], cv[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        const uint32_t x = q[i];
        ck[j] = i < qn ? x + (a.cust.base - km) : 0xFFFFFFFFu;  // rows past the queue: out of range, dropped
        sd[j] = q[kQ31Stride + i];
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) cv[j] = code_at(a.c_code, ck[j], 0u, a.n_cust);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        const uint32_t sk = min((sd[j] & kSMask) + (a.supp.base - km), a.n_supp);
        const uint32_t y = code_at(y_tab, (sd[j] >> SBW) + (a.date.base - km), 0u, a.n_date);
        if (cv[j] != kNone && y != kNone) {
            const uint32_t sn = s_tab[sk] - 1u;  // staged as nation + 1
            const uint32_t rev = q[2u * kQ31Stride + i] + a.rev.base;
            if (gt.dense) {
                const uint32_t sd = Q31Groups::byte_table(lds, 0)[sn], cd = Q31Groups::byte_table(lds, 1)[cv[j]];
                if (sd < gt.n_s && cd < gt.n_c && y < a.y_span) {
                    const uint32_t g = (sd * gt.n_c + cd) * a.y_span + y;
                    atomicAdd(&Q31Groups::sum(lds)[g], (unsigned long long)rev);
                    atomicAdd(&Q31Groups::cnt(lds)[g], 1u);
                    continue;
                }
            }
            const uint32_t slot = fast_mod(sn * a.slot_mul_s + cv[j] * a.slot_mul_c + y, a.slot_div);
            if (gt.dense) {
                atomicAdd(reinterpret_cast<unsigned long long*>(&a.sums[slot]), (unsigned long long)rev);
                atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[slot]), 1ull);
            } else {
                h.add(slot, rev, a.sums, a.counts);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}


size_t q31_bitmap_lds(uint32_t n_supp, uint32_t n_date, uint32_t c_words) {
    return ((kQ31SupplierLds + n_supp + 1u + n_date + 3u) & ~(size_t)3u) + 4ull * (c_words + 1u);
}

template <int DBW, int CBW, int SBW, int RBW>
__global__ __launch_bounds__(kStarBlock) void q31_bitmap_kernel(const StarDev a) {
    static_assert(DBW >= 0 && CBW >= 0 && SBW >= 0 && RBW >= 0 && DBW + SBW <= 32,
                  "fixed widths; supplier and date share a queue word");
    extern __shared__ __align__(16) uint8_t lds[];
    uint32_t* q = reinterpret_cast<uint32_t*>(lds) + (threadIdx.x >> 6) * 3u * kQ31Stride;
    // the qualifying nations (every block reads the same 16 words and takes
    // the same branch)
    Q31Groups gt{};
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        gt.n_s += (uint32_t)__popc(a.masks[w]);
        gt.n_c += (uint32_t)__popc(a.masks[8 + w]);
    }
    gt.n_groups = gt.n_s * gt.n_c * a.y_span;
    gt.dense = a.y_span <= 255u && gt.n_groups <= kQ31Dense;
    Q31Hash h;
    if (gt.dense) {
        for (uint32_t i = threadIdx.x; i < gt.n_groups; i += blockDim.x) {
            Q31Groups::sum(lds)[i] = 0ull;
            Q31Groups::cnt(lds)[i] = 0u;
        }
        if (threadIdx.x < 512u) {
            const uint32_t side = threadIdx.x >> 8, n = threadIdx.x & 255u;  // side 0: supplier, 1: customer
            const uint32_t* mk = a.masks + 8u * side;
            uint32_t rank = 0;
            for (uint32_t w = 0; w < (n >> 5); ++w) rank += (uint32_t)__popc(mk[w]);
            const uint32_t mw = mk[n >> 5];
            rank += (uint32_t)__popc(mw & ((1u << (n & 31u)) - 1u));
            const bool set = (mw >> (n & 31u)) & 1u;
            Q31Groups::byte_table(lds, side)[n] = set ? (uint8_t)rank : kNone;
            if (set) Q31Groups::byte_table(lds, 2u + side)[rank] = (uint8_t)n;
        }
    } else {
        h.init(lds + kQ31BitmapQueueBytes);
    }
    uint8_t* s_tab = lds + kQ31SupplierLds;
    uint8_t* y_tab = s_tab + a.n_supp + 1u;
    uint32_t* c_tab =
        reinterpret_cast<uint32_t*>(lds + ((kQ31SupplierLds + a.n_supp + 1u + a.n_date + 3u) & ~(size_t)3u));
    // supplier codes staged as nation + 1 (0: not qualifying, and the
    // sentinel), so a lookup's hit bit is min(code, 1)
    for (uint32_t i = threadIdx.x; i <= a.n_supp; i += blockDim.x)
        s_tab[i] = i < a.n_supp ? (uint8_t)(a.s_code[i] + 1u) : (uint8_t)0;
    for (uint32_t i = threadIdx.x; i < a.n_date; i += blockDim.x) y_tab[i] = a.y_code[i];
    for (uint32_t i = threadIdx.x; i <= a.c_words; i += blockDim.x) c_tab[i] = i < a.c_words ? a.c_bits[i] : 0u;
    __syncthreads();
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    const uint32_t d_s = a.supp.base - km, d_c = a.cust.base - km;
    uint32_t qn = 0;  // wave-uniform; the queue carries over from step to step
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kStarVecs + (threadIdx.x >> 5);
        // the whole wave stays in step (the queue flush needs every lane): a
        // half-wave past the last vector reads the group's first vector (in
        // bounds: the group has a live vector) and keeps no rows
        const bool vlive = v < a.n_vec;
        if (!__any(vlive)) continue;
        const uint64_t g0 = g * kStarVecs;
        const uint32_t slot = vlive ? (threadIdx.x >> 5) : 0u;
        LaneCol<SBW> sc;
        LaneCol<CBW> cc;
        LaneCol<DBW> dc;
        LaneCol<RBW> rc;
        sc.load_at(a.supp, g0, slot);
        cc.load_at(a.cust, g0, slot);
        dc.load_at(a.date, g0, slot);
        rc.load_at(a.rev, g0, slot);
        // every load in flight before the first row is tested (the scheduler
        // otherwise interleaves address arithmetic that waits for early loads)
        __builtin_amdgcn_sched_barrier(0);
        // supplier semi-join on every row
        uint32_t m = 0;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const uint32_t k = min(dev::extract_at<SBW>(sc.w, r) + d_s, a.n_supp);
            // the hit bit as one opaque min (the compiler turns min(code, 1)
            // into a compare and a select)
            uint32_t bit;
            asm("v_min_u32 %0, 1, %1" : "=v"(bit) : "v"((uint32_t)s_tab[k]));
            m |= bit << r;
        }
        if (!vlive) m = 0;
        if (v == a.n_vec - 1) m &= tail;
        cc.pin();
        dc.pin();
        rc.pin();
        if (!__any(m != 0)) continue;
        // customer semi-join on every row from the LDS bitmap: only the rows
        // that pass both filters (about 4 %) reach the queue
        uint32_t hit = 0;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const uint32_t k = min(dev::extract_at<CBW>(cc.w, r) + d_c, a.n_cust);
            // the word index k >> 5 from an opaque shift, so the address is
            // one shift-add (the compiler's (k >> 3) & ~3 form costs two
            // more); v_bfe_u32 reads the bit offset from k's low five bits
            uint32_t word;
            asm("v_lshrrev_b32 %0, 5, %1" : "=v"(word) : "v"(k));
            hit |= __builtin_amdgcn_ubfe(c_tab[word], k, 1u) << r;
        }
        m &= hit;
        if (!__any(m != 0)) continue;
        // queue slots: the wave's survivors in (lane, row) order, from a wave
        // prefix sum of the per-lane counts; a dead row slot writes the spare
        // entry kQ31Queue, so the writes need no branch.  When more rows
        // survive than a queue holds (never at SSB selectivities) they go two
        // row slots at a time (at most 128 rows).
        uint32_t part = 0xFFFFFFFFu;  // row slots of this round (wave-uniform)
        while (true) {
            const uint32_t mm = m & part;
            const uint32_t cnt = (uint32_t)__popc(mm);
            const uint32_t incl = dev::wave_inclusive_sum(cnt);
            const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
            if (total > kQ31Queue) {
                part = 3u;
                continue;
            }
            if (qn + total > kQ31Queue) {
                q31_bitmap_flush<DBW, CBW, SBW, RBW>(a, s_tab, y_tab, h, gt, lds, q, qn);
                qn = 0;
            }
            const uint32_t first = qn + incl - cnt;
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const uint32_t below = (uint32_t)__popc(mm & ((1u << r) - 1u));
                const uint32_t pos = (mm >> r) & 1u ? first + below : kQ31Queue;
                q[pos] = dev::extract_at<CBW>(cc.w, r);
                q[kQ31Stride + pos] = dev::extract_at<SBW>(sc.w, r) | dev::extract_at<DBW>(dc.w, r) << SBW;
                q[2u * kQ31Stride + pos] = dev::extract_at<RBW>(rc.w, r);
            }
            qn += total;
            m &= ~mm;
            if (part == 0xFFFFFFFFu || part == 0xC0000000u) break;
            part <<= 2;
        }
    }
    if (qn) q31_bitmap_flush<DBW, CBW, SBW, RBW>(a, s_tab, y_tab, h, gt, lds, q, qn);
    if (gt.dense) {
        __syncthreads();
        for (uint32_t g = threadIdx.x; g < gt.n_groups; g += blockDim.x) {
            const uint32_t c = Q31Groups::cnt(lds)[g];
            if (c) {
                const uint32_t y = g % a.y_span, sc = g / a.y_span;
                const uint32_t slot = fast_mod(Q31Groups::byte_table(lds, 2)[sc / gt.n_c] * a.slot_mul_s +
                                                   Q31Groups::byte_table(lds, 3)[sc % gt.n_c] * a.slot_mul_c + y,
                                               a.slot_div);
                atomicAdd(reinterpret_cast<unsigned long long*>(&a.sums[slot]), Q31Groups::sum(lds)[g]);
                atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[slot]), (unsigned long long)c);
            }
        }
    } else {
        h.flush(a.sums, a.counts);
    }
}

// Q3.1 probe over columns of any encoding and width: one column decoded at a
// time (32 registers), only the per-row codes carried between them.
template <bool USE_LDS>
__global__ __launch_bounds__(kStarBlock) void q31_any_kernel(const StarDev a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const StarLds t = star_stage<USE_LDS, false>(a, lds);
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kStarVecs + (threadIdx.x >> 5);
        if (v >= a.n_vec) continue;
        RegCol col;
        uint32_t m = (v == a.n_vec - 1) ? tail : 0xFFFFFFFFu;
        uint32_t s8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, c8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, y8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        col.load(a.supp, v);
        filter_codes<kB31>(col, t.s_code, km, a.n_supp, m, s8);
        if (!__any(m != 0)) continue;
        col.load(a.cust, v);
        filter_codes<kB31>(col, a.c_code, km, a.n_cust, m, c8);
        if (!__any(m != 0)) continue;
        col.load(a.date, v);
        filter_codes<kB31>(col, t.y_code, km, a.n_date, m, y8);
        if (!__any(m != 0)) continue;
        col.load(a.rev, v);
#pragma unroll
        for (int r = 0; r < 32; ++r)
            if ((m >> r) & 1u)
                group_add_lds(
                    t, fast_mod(get8(s8, r) * a.slot_mul_s + get8(c8, r) * a.slot_mul_c + get8(y8, r), a.slot_div),
                    col.get(r));
    }
    star_flush(a, t);
}

// Q4.1 probe at fixed widths: supplier and part pass (LDS), then the last four
// columns' loads; the rows that pass (about 8 %) go to a per-wave LDS queue
// (customer key, date key, revenue, supplycost) that carries over from step
// to step and is flushed when a row slot's rows would not fit: one queued row
// per lane, customer nation (L2), year offset (LDS), revenue - supplycost
// into the LDS group table.  Issuing all six columns' loads before the filter
// measured no faster (the 16 waves of a CU already cover the second round
// trip) and, kept there by a scheduling barrier, spilled 4 VGPRs and ran 5 %
// slower.
constexpr uint32_t kQ41Queue = 128;
constexpr size_t kQ41QueueBytes = (size_t)kStarVecs / 2u * 4u * kQ41Queue * 4u;  // 16 waves, 32 KB

__device__ __forceinline__ void q41_flush(const StarDev& a, const StarLds& t, const uint32_t* q, uint32_t qn) {
    constexpr int kPer = kQ41Queue / 64;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t l = threadIdx.x & 63u;
    uint32_t ck[kPer], cv[kPer], y[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        const uint32_t x = q[i];
        ck[j] = i < qn ? x : 0xFFFFFFFFu;  // rows past the queue: out of range, dropped
        y[j] = code_at(t.y_code, q[kQ41Queue + i], 0u, a.n_date);
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) cv[j] = code_at(a.c_code, ck[j], 0u, a.n_cust);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        if (cv[j] != kNone && y[j] != kNone)
            group_add_lds(t, fast_mod(cv[j] * a.slot_mul_c + y[j], a.slot_div),
                          (uint64_t)((int64_t)q[2u * kQ41Queue + i] - (int64_t)q[3u * kQ41Queue + i]));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// USE_LDS: group table, part bits, supplier bits and year codes in LDS (SSB
// sizes: 2.1 KB + 100 KB + 5 KB + 2.5 KB); otherwise the group table alone.
template <int DBW, int CBW, int SBW, int PBW, int RBW, int KBW, bool USE_LDS>
__global__ __launch_bounds__(kStarBlock) void q41_kernel(const StarDev a) {
    extern __shared__ __align__(16) uint8_t lds[];
    uint32_t* q = reinterpret_cast<uint32_t*>(lds) + (threadIdx.x >> 6) * 4u * kQ41Queue;
    StarLds t;
    uint8_t* base = lds + kQ41QueueBytes;
    t.sum = reinterpret_cast<unsigned long long*>(base);
    t.cnt = reinterpret_cast<uint32_t*>(base + 8ull * a.n_slots);
    for (uint32_t i = threadIdx.x; i < a.n_slots; i += blockDim.x) {
        t.sum[i] = 0ull;
        t.cnt[i] = 0u;
    }
    t.s_code = a.s_code;
    t.y_code = a.y_code;
    t.p_bits = a.p_bits;
    const uint32_t* s_bits = a.s_bits;
    if constexpr (USE_LDS) {
        uint32_t* pb = reinterpret_cast<uint32_t*>(base + 12ull * a.n_slots);  // 4-byte aligned
        uint32_t* sb = pb + a.p_words;
        uint8_t* y = reinterpret_cast<uint8_t*>(sb + a.s_words);
        for (uint32_t i = threadIdx.x; i < a.p_words; i += blockDim.x) pb[i] = a.p_bits[i];
        for (uint32_t i = threadIdx.x; i < a.s_words; i += blockDim.x) sb[i] = a.s_bits[i];
        for (uint32_t i = threadIdx.x; i < a.n_date; i += blockDim.x) y[i] = a.y_code[i];
        t.p_bits = pb;
        t.y_code = y;
        s_bits = sb;
    }
    __syncthreads();
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    uint32_t qn = 0;  // wave-uniform; the queue carries over from step to step
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kStarVecs + (threadIdx.x >> 5);
        // the whole wave stays in step (the queue flush needs every lane): a
        // half-wave past the last vector reads the last vector again (in
        // bounds) and keeps no rows
        const bool vlive = v < a.n_vec;
        if (!__any(vlive)) continue;
        const uint64_t vl = vlive ? v : a.n_vec - 1;
        uint32_t m = 0;
        {
            LaneCol<SBW> sc;
            LaneCol<PBW> pc;
            sc.load(a.supp, vl);
            pc.load(a.part, vl);
#pragma unroll
            for (int r = 0; r < 32; ++r)
                m |= (bit_at(s_bits, sc.get(r), km, a.n_supp) & bit_at(t.p_bits, pc.get(r), km, a.n_part)) << r;
        }
        if (!vlive) m = 0;
        if (v == a.n_vec - 1) m &= tail;
        LaneCol<CBW> cc;
        LaneCol<DBW> dc;
        LaneCol<RBW> rc;
        LaneCol<KBW> kc;
        cc.load(a.cust, vl);
        dc.load(a.date, vl);
        rc.load(a.rev, vl);
        kc.load(a.cost, vl);
        cc.pin();
        dc.pin();
        rc.pin();
        kc.pin();
        if (!__any(m != 0)) continue;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const bool live = (m >> r) & 1u;
            const uint64_t bal = __ballot(live);
            if (qn + (uint32_t)__popcll(bal) > kQ41Queue) {
                q41_flush(a, t, q, qn);
                qn = 0;
            }
            if (live) {
                const uint32_t pos =
                    qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                q[pos] = cc.get(r) - km;
                q[kQ41Queue + pos] = dc.get(r) - km;
                q[2u * kQ41Queue + pos] = rc.get(r);
                q[3u * kQ41Queue + pos] = kc.get(r);
            }
            qn += (uint32_t)__popcll(bal);
        }
    }
    if (qn) q41_flush(a, t, q, qn);
    star_flush(a, t);
}

// kQ41 bitmap probe (fixed widths, SSB-sized dimensions).  LDS: the row
// queues (3 fields of kQ41BQueue entries and a spare entry per wave), then at
// fixed offsets the supplier and part filter bitmaps (a zero sentinel bit at
// index n, so a key k tests bit min(k, n): out-of-range keys, below keys_min
// they wrap, read the sentinel), the year codes and the 175-slot group table.
// The survivors (8 % of the rows) are queued at slots from a wave prefix sum
// of the lanes' counts (dead row slots write the spare entry) as raw packed
// fields -- customer | date << CBW, revenue, supplycost -- and the flush adds
// the frame bases and keys_min, gathers the customer nation from L2 and
// looks the year up, four queued rows per lane.
constexpr uint32_t kQ41BQueue = 256;
constexpr uint32_t kQ41BStride = kQ41BQueue + 1u;
constexpr size_t kQ41BQueueBytes = (size_t)kStarVecs / 2u * 3u * kQ41BStride * 4u;  // 16 waves, 49,344 B

size_t q41_bitmap_lds(uint32_t s_words, uint32_t p_words, uint32_t n_date, uint32_t n_slots) {
    const size_t y = kQ41BQueueBytes + 4ull * (s_words + 1u) + 4ull * (p_words + 1u);
    return ((y + n_date + 7u) & ~(size_t)7u) + 12ull * n_slots;
}

template <int DBW, int CBW, int RBW, int KBW>
__device__ __forceinline__ void q41_bitmap_flush(const StarDev& a, const uint8_t* __restrict y_tab,
                                                 unsigned long long* __restrict sum, uint32_t* __restrict cnt,
                                                 const uint32_t* q, uint32_t qn) {
    constexpr int kPer = kQ41BQueue / 64;
    constexpr uint32_t kCMask = CBW == 32 ? 0xFFFFFFFFu : (1u << CBW) - 1u;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t l = threadIdx.x & 63u;
    const uint32_t km = a.keys_min;
    uint32_t ck[kPer], cv[kPer], y[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        const uint32_t w = q[i];
        ck[j] = i < qn ? (w & kCMask) + (a.cust.base - km) : 0xFFFFFFFFu;  // rows past the queue: dropped
        y[j] = code_at(y_tab, (w >> CBW) + (a.date.base - km), 0u, a.n_date);
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) cv[j] = code_at(a.c_code, ck[j], 0u, a.n_cust);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const uint32_t i = l + 64u * j;
        if (cv[j] != kNone && y[j] != kNone) {
            const uint32_t rev = q[kQ41BStride + i] + a.rev.base, cost = q[2u * kQ41BStride + i] + a.cost.base;
            const uint32_t slot = fast_mod(cv[j] * a.slot_mul_c + y[j], a.slot_div);
            atomicAdd(&sum[slot], (unsigned long long)((int64_t)rev - (int64_t)cost));
            atomicAdd(&cnt[slot], 1u);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int DBW, int CBW, int SBW, int PBW, int RBW, int KBW>
__global__ __launch_bounds__(kStarBlock) void q41_bitmap_kernel(const StarDev a) {
    static_assert(DBW + CBW <= 32, "customer and date share a queue word");
    extern __shared__ __align__(16) uint8_t lds[];
    uint32_t* q = reinterpret_cast<uint32_t*>(lds) + (threadIdx.x >> 6) * 3u * kQ41BStride;
    uint32_t* sb = reinterpret_cast<uint32_t*>(lds + kQ41BQueueBytes);
    uint32_t* pb = sb + a.s_words + 1u;
    uint8_t* y_tab = reinterpret_cast<uint8_t*>(pb + a.p_words + 1u);
    const size_t g_off = (kQ41BQueueBytes + 4ull * (a.s_words + 1u) + 4ull * (a.p_words + 1u) + a.n_date + 7u) &
                         ~(size_t)7u;
    unsigned long long* sum = reinterpret_cast<unsigned long long*>(lds + g_off);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(sum + a.n_slots);
    for (uint32_t i = threadIdx.x; i <= a.s_words; i += blockDim.x) sb[i] = i < a.s_words ? a.s_bits[i] : 0u;
    for (uint32_t i = threadIdx.x; i <= a.p_words; i += blockDim.x) pb[i] = i < a.p_words ? a.p_bits[i] : 0u;
    for (uint32_t i = threadIdx.x; i < a.n_date; i += blockDim.x) y_tab[i] = a.y_code[i];
    for (uint32_t i = threadIdx.x; i < a.n_slots; i += blockDim.x) {
        sum[i] = 0ull;
        cnt[i] = 0u;
    }
    __syncthreads();
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    const uint32_t d_s = a.supp.base - km, d_p = a.part.base - km;
    uint32_t qn = 0;  // wave-uniform; the queue carries over from step to step
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kStarVecs + (threadIdx.x >> 5);
        // the whole wave stays in step (the queue flush needs every lane): a
        // half-wave past the last vector reads the group's first vector (in
        // bounds: the group has a live vector) and keeps no rows
        const bool vlive = v < a.n_vec;
        if (!__any(vlive)) continue;
        const uint64_t g0 = g * kStarVecs;
        const uint32_t slot = vlive ? (threadIdx.x >> 5) : 0u;
        uint32_t m = 0;
        {
            LaneCol<SBW> sc;
            LaneCol<PBW> pc;
            sc.load_at(a.supp, g0, slot);
            pc.load_at(a.part, g0, slot);
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const uint32_t ks = min(dev::extract_at<SBW>(sc.w, r) + d_s, a.n_supp);
                const uint32_t kp = min(dev::extract_at<PBW>(pc.w, r) + d_p, a.n_part);
                // word indices as opaque shifts (one shift-add per address);
                // v_bfe_u32 reads the bit offset from the key's low five bits
                uint32_t ws, wp;
                asm("v_lshrrev_b32 %0, 5, %1" : "=v"(ws) : "v"(ks));
                asm("v_lshrrev_b32 %0, 5, %1" : "=v"(wp) : "v"(kp));
                m |= (__builtin_amdgcn_ubfe(sb[ws], ks, 1u) & __builtin_amdgcn_ubfe(pb[wp], kp, 1u)) << r;
            }
        }
        if (!vlive) m = 0;
        if (v == a.n_vec - 1) m &= tail;
        LaneCol<CBW> cc;
        LaneCol<DBW> dc;
        LaneCol<RBW> rc;
        LaneCol<KBW> kc;
        cc.load_at(a.cust, g0, slot);
        dc.load_at(a.date, g0, slot);
        rc.load_at(a.rev, g0, slot);
        kc.load_at(a.cost, g0, slot);
        cc.pin();
        dc.pin();
        rc.pin();
        kc.pin();
        if (!__any(m != 0)) continue;
        // queue slots from a wave prefix sum of the lanes' counts; a dead row
        // slot writes the spare entry.  More survivors than a queue holds
        // (not at SSB selectivities) go four row slots at a time.
        uint32_t part = 0xFFFFFFFFu;  // row slots of this round (wave-uniform)
        while (true) {
            const uint32_t mm = m & part;
            const uint32_t n_mine = (uint32_t)__popc(mm);
            const uint32_t incl = dev::wave_inclusive_sum(n_mine);
            const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
            if (total > kQ41BQueue) {
                part = 0xFu;
                continue;
            }
            if (qn + total > kQ41BQueue) {
                q41_bitmap_flush<DBW, CBW, RBW, KBW>(a, y_tab, sum, cnt, q, qn);
                qn = 0;
            }
            const uint32_t first = qn + incl - n_mine;
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                const uint32_t below = (uint32_t)__popc(mm & ((1u << r) - 1u));
                const uint32_t pos = (mm >> r) & 1u ? first + below : kQ41BQueue;
                q[pos] = dev::extract_at<CBW>(cc.w, r) | dev::extract_at<DBW>(dc.w, r) << CBW;
                q[kQ41BStride + pos] = dev::extract_at<RBW>(rc.w, r);
                q[2u * kQ41BStride + pos] = dev::extract_at<KBW>(kc.w, r);
            }
            qn += total;
            m &= ~mm;
            if (part == 0xFFFFFFFFu || part == 0xF0000000u) break;
            part <<= 4;
        }
    }
    if (qn) q41_bitmap_flush<DBW, CBW, RBW, KBW>(a, y_tab, sum, cnt, q, qn);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < a.n_slots; i += blockDim.x)
        if (cnt[i]) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.sums[i]), sum[i]);
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[i]), (unsigned long long)cnt[i]);
        }
}

// Q4.1 probe over columns of any encoding and width, one column at a time.
// Revenue and supplycost enter the group sums in two passes (+revenue with the
// row count, then -supplycost): the 64-bit wrapping sums are the same.
template <bool USE_LDS>
__global__ __launch_bounds__(kStarBlock) void q41_any_kernel(const StarDev a) {
    extern __shared__ __align__(16) uint8_t lds[];
    const StarLds t = star_stage<USE_LDS, true>(a, lds);
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kStarVecs + (threadIdx.x >> 5);
        if (v >= a.n_vec) continue;
        RegCol col;
        uint32_t m = (v == a.n_vec - 1) ? tail : 0xFFFFFFFFu;
        uint32_t s8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, c8[8] = {0, 0, 0, 0, 0, 0, 0, 0}, y8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        col.load(a.supp, v);
        filter_codes<kB31>(col, t.s_code, km, a.n_supp, m, s8);
        if (!__any(m != 0)) continue;
        col.load(a.part, v);
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            m &= ~((bit_at(t.p_bits, col.get(r), km, a.n_part) ^ 1u) << r);
        }
        if (!__any(m != 0)) continue;
        col.load(a.cust, v);
        filter_codes<kB31>(col, a.c_code, km, a.n_cust, m, c8);
        if (!__any(m != 0)) continue;
        col.load(a.date, v);
        filter_codes<kB31>(col, t.y_code, km, a.n_date, m, y8);
        if (!__any(m != 0)) continue;
        uint32_t slot16[16];  // slot of each row (< kMaxSlots), two per register
#pragma unroll
        for (int k = 0; k < 16; ++k)
            slot16[k] = fast_mod(get8(c8, 2 * k) * a.slot_mul_c + get8(y8, 2 * k), a.slot_div) |
                        (fast_mod(get8(c8, 2 * k + 1) * a.slot_mul_c + get8(y8, 2 * k + 1), a.slot_div) << 16);
        col.load(a.rev, v);
#pragma unroll
        for (int r = 0; r < 32; ++r)
            if ((m >> r) & 1u) group_add_lds(t, (slot16[r >> 1] >> (16 * (r & 1))) & 0xFFFFu, col.get(r));
        col.load(a.cost, v);
#pragma unroll
        for (int r = 0; r < 32; ++r)
            if ((m >> r) & 1u)
                atomicAdd(&t.sum[(slot16[r >> 1] >> (16 * (r & 1))) & 0xFFFFu],
                          (unsigned long long)(0ull - (uint64_t)col.get(r)));
    }
    star_flush(a, t);
}

template <typename Kernel>
int launch_star(Kernel k, size_t lds, const StarDev& a, hipStream_t s) {
    FLS_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)lds));
    uint32_t grid = 0;
    if (const int st = resident_grid_of(reinterpret_cast<const void*>(k), kStarBlock, lds, a.n_groups, &grid))
        return st;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kStarBlock), lds, s, a);
    return check_launch();
}

// Fact columns of any encoding, all in one row order (DELTA, RLE and MIXED
// columns decode in transposed order, as does BITPACK with transposed set).
int star_columns(StarDev& a, const fls_column* const* cols, ColDev* const* outs, int n, uint64_t n_rows) {
    for (int i = 0; i < n; ++i) {
        if (!cols[i]) return FLS_ERR_NULL;
        if (const int st = to_dev_column(*cols[i], n_rows, outs[i], true)) return st;
        if ((outs[i]->transposed != 0) != (outs[0]->transposed != 0)) return FLS_ERR_ARG;
    }
    a.n_rows = n_rows;
    a.n_vec = ceil_div(n_rows, 1024);
    a.n_groups = ceil_div(a.n_vec, kStarVecs);
    return FLS_OK;
}

bool all_bitpack(ColDev* const* outs, int n) {
    for (int i = 0; i < n; ++i)
        if (outs[i]->enc != FLS_ENC_BITPACK) return false;
    return true;
}

int star_workspace(StarDev& a, const fls_star_dims* dims, const void* d_ws) {
    const StarLayout L = star_layout(*dims);
    const char* ws = static_cast<const char*>(d_ws);
    a.s_code = reinterpret_cast<const uint8_t*>(ws + L.s_code);
    a.c_code = reinterpret_cast<const uint8_t*>(ws + L.c_code);
    a.p_bits = reinterpret_cast<const uint32_t*>(ws + L.p_bits);
    a.c_bits = reinterpret_cast<const uint32_t*>(ws + L.c_bits);
    a.s_bits = reinterpret_cast<const uint32_t*>(ws + L.s_bits);
    a.masks = reinterpret_cast<const uint32_t*>(ws + L.masks);
    a.y_code = reinterpret_cast<const uint8_t*>(ws + L.y_code);
    a.n_supp = dims->n_supp;
    a.n_cust = dims->n_cust;
    a.n_part = dims->n_part;
    a.n_date = dims->n_date;
    a.p_words = L.p_words;
    a.c_words = L.c_words;
    a.s_words = L.s_words;
    a.keys_min = dims->keys_min;
    return FLS_OK;
}

int star_outputs(StarDev& a, uint32_t n_slots, int64_t* d_sums, uint64_t* d_counts, uint32_t flags, hipStream_t s) {
    if (!d_sums || !d_counts) return FLS_ERR_NULL;
    if (n_slots == 0 || n_slots > kMaxSlots) return FLS_ERR_ARG;
    a.n_slots = n_slots;
    a.slot_div = make_fastdiv(n_slots);
    a.sums = d_sums;
    a.counts = d_counts;
    if (!(flags & FLS_Q_ACCUMULATE)) {
        if (const int st = zero_u64(reinterpret_cast<uint64_t*>(d_sums), n_slots, s)) return st;
        return zero_u64(d_counts, n_slots, s);
    }
    return FLS_OK;
}

constexpr size_t kLdsBudget = 150u * 1024u;
constexpr size_t kLdsMax = 160u * 1024u;  // gfx950: LDS per CU, and per block

}  // namespace
}  // namespace fls

extern "C" {

uint64_t fls_star_workspace_bytes(const fls_star_dims* dims) {
    return dims ? fls::star_layout(*dims).total : 0;
}

int fls_q31_build(const fls_star_dims* dims, const fls_q31_params* params, void* d_workspace, uint64_t ws_bytes,
                  void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !params || !d_workspace, FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->s_region || !dims->s_nation || !dims->c_region || !dims->c_nation || !dims->d_year,
                  FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->n_supp || !dims->n_cust || !dims->n_date, FLS_ERR_ARG);
    FLS_RETURN_IF(params->year_hi < params->year_lo || params->year_lo < params->year_base ||
                      params->year_hi - params->year_base >= kNone,
                  FLS_ERR_ARG);
    const StarLayout L = star_layout(*dims);
    FLS_RETURN_IF(ws_bytes < L.total, FLS_ERR_CAPACITY);
    uint8_t* ws = static_cast<uint8_t*>(d_workspace);
    hipStream_t s = as_stream(stream);
    uint32_t* c_bits = reinterpret_cast<uint32_t*>(ws + L.c_bits);
    uint32_t* masks = reinterpret_cast<uint32_t*>(ws + L.masks);
    hipLaunchKernelGGL(star_zero_words, dim3(256), dim3(256), 0, s, c_bits, L.c_words);
    hipLaunchKernelGGL(star_zero_words, dim3(1), dim3(64), 0, s, masks, 16u);
    hipLaunchKernelGGL(q31_build_kernel, dim3(1024), dim3(256), 0, s, *dims, *params, ws + L.s_code, ws + L.c_code,
                       ws + L.y_code, c_bits, masks);
    return check_launch();
}

int fls_query_q31(const fls_column* datekey, const fls_column* custkey, const fls_column* suppkey,
                  const fls_column* revenue, uint64_t n_rows, const fls_star_dims* dims, const void* d_workspace,
                  const fls_q31_params* params, int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !d_workspace || !params, FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->n_supp || !dims->n_cust || !dims->n_date, FLS_ERR_ARG);
    StarDev a{};
    const fls_column* cols[4] = {datekey, custkey, suppkey, revenue};
    ColDev* outs[4] = {&a.date, &a.cust, &a.supp, &a.rev};
    if (const int st = star_columns(a, cols, outs, 4, n_rows)) return st;
    hipStream_t s = as_stream(stream);
    if (const int st = star_outputs(a, params->n_slots, d_sums, d_counts, flags, s)) return st;
    if (n_rows == 0) return FLS_OK;
    star_workspace(a, dims, d_workspace);
    a.slot_mul_s = params->n_nations * params->n_years;
    a.slot_mul_c = params->n_years;
    a.y_span = params->year_hi - params->year_base + 1u;
    const size_t table = 12ull * a.n_slots;
    const size_t staged = table + dims->n_supp + dims->n_date;
    const bool fast = all_bitpack(outs, 4) && a.date.bw == 12 && a.cust.bw == 20 && a.supp.bw == 16 &&
                      a.rev.bw == 24 && dims->n_date < (1u << 24);
    if (fast) {
        const size_t bitmap = q31_bitmap_lds(dims->n_supp, dims->n_date, a.c_words);
        if (bitmap <= kLdsMax) return launch_star(q31_bitmap_kernel<12, 20, 16, 24>, bitmap, a, s);
        if (kQ31QueueBytes + staged <= kLdsMax)
            return launch_star(q31_kernel<12, 20, 16, 24, kQ31Staged>, kQ31QueueBytes + staged, a, s);
        return launch_star(q31_kernel<12, 20, 16, 24, kQ31Codes>, kQ31QueueBytes + table, a, s);
    }
    const bool in_lds = staged <= kLdsBudget;
    const size_t lds = in_lds ? staged : table;
    return in_lds ? launch_star(q31_any_kernel<true>, lds, a, s) : launch_star(q31_any_kernel<false>, lds, a, s);
}

int fls_q41_build(const fls_star_dims* dims, const fls_q41_params* params, void* d_workspace, uint64_t ws_bytes,
                  void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !params || !d_workspace, FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->s_region || !dims->c_region || !dims->c_nation || !dims->p_mfgr || !dims->d_year,
                  FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->n_supp || !dims->n_cust || !dims->n_date || !dims->n_part, FLS_ERR_ARG);
    const StarLayout L = star_layout(*dims);
    FLS_RETURN_IF(ws_bytes < L.total, FLS_ERR_CAPACITY);
    uint8_t* ws = static_cast<uint8_t*>(d_workspace);
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(star_zero_words, dim3(256), dim3(256), 0, s, reinterpret_cast<uint32_t*>(ws + L.p_bits),
                       L.p_words);
    hipLaunchKernelGGL(star_zero_words, dim3(256), dim3(256), 0, s, reinterpret_cast<uint32_t*>(ws + L.s_bits),
                       L.s_words);
    hipLaunchKernelGGL(q41_build_kernel, dim3(1024), dim3(256), 0, s, *dims, *params, ws + L.s_code, ws + L.c_code,
                       reinterpret_cast<uint32_t*>(ws + L.p_bits), ws + L.y_code,
                       reinterpret_cast<uint32_t*>(ws + L.s_bits));
    return check_launch();
}

int fls_query_q41(const fls_column* datekey, const fls_column* custkey, const fls_column* suppkey,
                  const fls_column* partkey, const fls_column* revenue, const fls_column* supplycost, uint64_t n_rows,
                  const fls_star_dims* dims, const void* d_workspace, const fls_q41_params* params, int64_t* d_sums,
                  uint64_t* d_counts, uint32_t flags, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !d_workspace || !params, FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->n_supp || !dims->n_cust || !dims->n_date || !dims->n_part, FLS_ERR_ARG);
    StarDev a{};
    const fls_column* cols[6] = {datekey, custkey, suppkey, partkey, revenue, supplycost};
    ColDev* outs[6] = {&a.date, &a.cust, &a.supp, &a.part, &a.rev, &a.cost};
    if (const int st = star_columns(a, cols, outs, 6, n_rows)) return st;
    hipStream_t s = as_stream(stream);
    if (const int st = star_outputs(a, params->n_slots, d_sums, d_counts, flags, s)) return st;
    if (n_rows == 0) return FLS_OK;
    star_workspace(a, dims, d_workspace);
    a.slot_mul_s = 0;
    a.slot_mul_c = params->n_years;
    const size_t table = 12ull * a.n_slots;
    const size_t staged = table + 4ull * a.p_words + dims->n_supp + dims->n_date;
    const bool fast = all_bitpack(outs, 6) && a.date.bw == 12 && a.cust.bw == 20 && a.supp.bw == 16 &&
                      a.part.bw == 20 && a.rev.bw == 24 && a.cost.bw == 16;
    if (fast) {
        const size_t bitmap = q41_bitmap_lds(a.s_words, a.p_words, dims->n_date, a.n_slots);
        if (bitmap <= kLdsMax) return launch_star(q41_bitmap_kernel<12, 20, 16, 20, 24, 16>, bitmap, a, s);
        const size_t bits = table + 4ull * a.p_words + 4ull * a.s_words + dims->n_date;
        const bool fits = kQ41QueueBytes + bits <= kLdsMax;
        const size_t lds = kQ41QueueBytes + (fits ? bits : table);
        return fits ? launch_star(q41_kernel<12, 20, 16, 20, 24, 16, true>, lds, a, s)
                    : launch_star(q41_kernel<12, 20, 16, 20, 24, 16, false>, lds, a, s);
    }
    const bool in_lds = staged <= kLdsBudget;
    const size_t lds = in_lds ? staged : table;
    return in_lds ? launch_star(q41_any_kernel<true>, lds, a, s) : launch_star(q41_any_kernel<false>, lds, a, s);
}

}  // extern "C"
