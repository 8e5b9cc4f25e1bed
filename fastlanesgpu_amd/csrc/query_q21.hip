// Fused FastLanes decode + hash-probe + group-by, Q2.1 shape
// (reference: probe_v1 + build_hashtable_s/p/d, fastlanes/src/ssb/fls_q21.cu:21-78,
// 197-263, driver runQuery<BT,IPT>() fls_q21.cu:265-358).
//
// Two probe paths with identical results:
//  * PHT path: the reference's perfect hash tables (crystal/join.cuh layout:
//    PHT_1 = key per slot, PHT_2 = [key, value] pairs), built and probed with
//    the Crystal primitives; drop-in for callers that already hold such tables.
//  * compact path (MI355X-native): the build reduces each dimension to what
//    the probe needs -- a bitmap of qualifying suppliers, a bitmap of
//    qualifying parts, a 16-bit brand per part, a 16-bit year per date row --
//    and the probe stages the bitmaps and the year table in LDS (about 110 KB
//    at 40k suppliers / 800k parts / 2556 dates), so the two selective
//    semi-joins cost an LDS read per row instead of a random HBM/L2 access.
//    Brand and year are read only for rows that survive both filters.
//
// Both probe kernels are persistent, issue every fact column's loads before
// any data-dependent branch (full scan), and aggregate with 64-bit atomics on
// the n_slots-entry group table: slot = (brand*n_years + year-year_base) % n_slots
// (fls_q21.cu:71) -- the compact probe per block through an LDS hash table,
// then into partial tables that a second kernel sums into the output.
#include <string.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "fls/device/crystal.hpp"
#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_fastdiv.hpp"
#include "fls_group_hash.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

// ------------------------------------------------------------ PHT builds
constexpr int kBuildThreads = 128;
constexpr int kBuildItems = 4;

// build_hashtable_s / _p / _d (fls_q21.cu:197-263) in one kernel: optional
// equality filter, PHT_1 (vals == nullptr) or PHT_2.
__global__ __launch_bounds__(kBuildThreads) void build_pht_kernel(const uint32_t* __restrict filter, uint32_t eq,
                                                                  const uint32_t* __restrict keys,
                                                                  const uint32_t* __restrict vals, uint32_t n,
                                                                  uint32_t* __restrict ht, uint32_t ht_len,
                                                                  uint32_t keys_min) {
    constexpr int tile = kBuildThreads * kBuildItems;
    const uint32_t off = blockIdx.x * tile;
    const int num = (int)min<uint32_t>(tile, n - off);
    uint32_t items[kBuildItems], items2[kBuildItems];
    int flags[kBuildItems];
    InitFlags<kBuildThreads, kBuildItems>(flags);
    if (filter) {
        BlockLoad<uint32_t, kBuildThreads, kBuildItems>(filter + off, items, num);
        BlockPredEQ<uint32_t, kBuildThreads, kBuildItems>(items, eq, flags, num);
    }
    BlockLoad<uint32_t, kBuildThreads, kBuildItems>(keys + off, items, num);
    if (vals) {
        BlockLoad<uint32_t, kBuildThreads, kBuildItems>(vals + off, items2, num);
        BlockBuildSelectivePHT_2<uint32_t, uint32_t, kBuildThreads, kBuildItems>(items, items2, flags, ht,
                                                                                 (int)ht_len, keys_min, num);
    } else {
        BlockBuildSelectivePHT_1<uint32_t, kBuildThreads, kBuildItems>(items, flags, ht, (int)ht_len, keys_min, num);
    }
}

// ----------------------------------------------------------- compact build
struct Q21Layout {
    uint64_t supp_bm, part_bm, brand, year, total;  // byte offsets in the workspace
    uint32_t supp_words, part_words;
};

Q21Layout q21_layout(uint32_t n_supp, uint32_t n_part, uint32_t n_date) {
    auto up = [](uint64_t x) { return (x + 255u) & ~(uint64_t)255u; };
    Q21Layout L{};
    L.supp_words = (n_supp + 31u) / 32u;
    L.part_words = (n_part + 31u) / 32u;
    L.supp_bm = 0;
    L.part_bm = up(L.supp_bm + 4ull * L.supp_words);
    L.brand = up(L.part_bm + 4ull * L.part_words);
    L.year = up(L.brand + 2ull * n_part);
    L.total = up(L.year + 2ull * n_date);
    return L;
}

__global__ void q21_zero_words(uint32_t* __restrict w, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
}

__global__ void q21_build_compact(const uint32_t* __restrict s_region, uint32_t n_supp,
                                  const uint32_t* __restrict p_category, const uint32_t* __restrict p_brand,
                                  uint32_t n_part, const uint32_t* __restrict d_year, uint32_t n_date, uint32_t region,
                                  uint32_t category, uint32_t* __restrict supp_bm, uint32_t* __restrict part_bm,
                                  uint16_t* __restrict brand, uint16_t* __restrict year) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_part; i += stride) {
        const bool q = p_category[i] == category;
        brand[i] = q ? (uint16_t)p_brand[i] : (uint16_t)0;
        if (q) atomicOr(&part_bm[i >> 5], 1u << (i & 31u));
    }
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_supp; i += stride)
        if (s_region[i] == region) atomicOr(&supp_bm[i >> 5], 1u << (i & 31u));
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_date; i += stride) year[i] = (uint16_t)d_year[i];
}

// ------------------------------------------------------------ probe kernels
struct Q21Dev {
    ColDev date, supp, part, rev;
    uint64_t n_rows, n_vec, n_groups;
    // PHT path
    const uint32_t *ht_s, *ht_p, *ht_d;
    uint32_t s_len, p_len, d_len, s_min, p_min, d_min;
    const uint32_t *occ_s, *occ_p;  // occupancy bitmaps of ht_s / ht_p (bit-filtered probe)
    FastDiv s_div, p_div, d_div;    // s_len, p_len, d_len
    // compact path
    const uint32_t *supp_bm, *part_bm;
    const uint16_t *brand, *year;
    uint32_t n_supp, n_part, n_date, keys_min;
    uint32_t supp_words, part_words;
    // group-by
    uint32_t year_base, n_years, n_slots;
    FastDiv slot_div;  // n_slots
    int64_t* sums;
    uint64_t* counts;
    unsigned long long* copies;  // compact probe: this call's set of partial tables, or null (adds to sums / counts)
};

// The compact probe's blocks add their group tables to one of kQ21Copies
// partial tables (block b to copy b % kQ21Copies) and a second kernel sums the
// copies into the output, clearing them.  Adding every block's groups to the
// output directly cost 24-28 us of the 120M-row probe's 164: the ~280 groups
// of SSB Q2.1 sit in 18 cache lines, and the 245 blocks' atomics on them,
// issued as the blocks finish together, serialise.  A copy is a sums and a
// counts array of kQ21MaxSlots; a set (kQ21Copies copies) belongs to one call,
// taken round-robin from a static per-device array: up to kQ21Sets calls may
// run at the same time on one device.  Tables of more slots add to the output
// directly.
constexpr uint32_t kQ21Copies = 16, kQ21MaxSlots = 8192, kQ21Sets = 8;
constexpr uint64_t kQ21SetWords = 2ull * kQ21Copies * kQ21MaxSlots;  // 2 MiB
static __device__ __align__(256) unsigned long long q21_copies[kQ21Sets][kQ21SetWords];

// One work-item per slot: the copies' sums and counts, read and cleared with
// atomics (the probe's atomics are performed at the device's coherence point),
// written to the output or, with FLS_Q_ACCUMULATE, added to it.
__global__ __launch_bounds__(256) void q21_collect_kernel(unsigned long long* __restrict set, uint32_t n_slots,
                                                          int64_t* __restrict sums, uint64_t* __restrict counts,
                                                          uint32_t accumulate) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n_slots) return;
    unsigned long long sm = 0, cn = 0;
#pragma unroll
    for (uint32_t k = 0; k < kQ21Copies; ++k) {
        unsigned long long* cp = set + (uint64_t)k * (2u * kQ21MaxSlots);
        sm += atomicExch(&cp[i], 0ull);
        cn += atomicExch(&cp[kQ21MaxSlots + i], 0ull);
    }
    if (accumulate) {
        sums[i] = (int64_t)((unsigned long long)sums[i] + sm);
        counts[i] += cn;
    } else {
        sums[i] = (int64_t)sm;
        counts[i] = cn;
    }
}

__device__ __forceinline__ void group_add(const Q21Dev& a, uint32_t brand, uint32_t year, uint32_t revenue) {
    const uint32_t slot = fast_mod(brand * a.n_years + (year - a.year_base), a.slot_div);
    atomicAdd((unsigned long long*)&a.sums[slot], (unsigned long long)revenue);
    atomicAdd((unsigned long long*)&a.counts[slot], 1ull);
}

// Reference-layout probe (probe_v1, fls_q21.cu:21-78) through the Crystal
// primitives, one 32-lane tile per FastLanes vector.  The rows of a partial
// last vector are the first `num` striped items in row order; in transposed
// order they are scattered over the lane's registers, so the flags start from
// the live-row mask and the primitives see a whole tile.
template <int DBW, int SBW, int PBW, int RBW>
__global__ __launch_bounds__(256) void q21_pht_kernel(const Q21Dev a) {
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * 8u + (threadIdx.x >> 5);
        if (v >= a.n_vec) continue;
        int num = (int)min<uint64_t>(1024u, a.n_rows - v * 1024u);
        uint32_t items[32], brand[32], year[32];
        int flags[32];
        InitFlags<32, 32>(flags);
        if (num < 1024 && a.rev.transposed) {
            const uint32_t live = live_row_mask((uint32_t)num, true);
#pragma unroll
            for (int i = 0; i < 32; ++i) flags[i] = (int)((live >> i) & 1u);
            num = 1024;
        }
        decode_fixed<SBW>(a.supp, v, items);
        BlockProbeAndPHT_1<uint32_t, 32, 32>(items, flags, a.ht_s, (int)a.s_len, a.s_min, num);
        decode_fixed<PBW>(a.part, v, items);
        BlockProbeAndPHT_2<uint32_t, uint32_t, 32, 32>(items, brand, flags, a.ht_p, (int)a.p_len, a.p_min, num);
        decode_fixed<DBW>(a.date, v, items);
        BlockProbeAndPHT_2<uint32_t, uint32_t, 32, 32>(items, year, flags, a.ht_d, (int)a.d_len, a.d_min, num);
        decode_fixed<RBW>(a.rev, v, items);
        const int tid = dev::lane32();
#pragma unroll
        for (int i = 0; i < 32; ++i)
            if (tid + 32 * i < num && flags[i]) group_add(a, brand[i], year[i], items[i]);
    }
}

// Compact-table probe.  1024 work-items = 32 vectors per step; LDS holds the
// supplier and part bitmaps and the year table when they fit (USE_LDS).
using Q21Hash = GroupHash<10>;  // 1,024 entries, 20 KB

// Per-wave LDS queue of rows that passed every filter, kQueue of each field.
// The flush runs with one lane per row: brand from L2, year from the year
// table, then the block's hash table.  With fixed-width date and revenue
// columns (LATE) a row is queued as its part index and its place (vector;
// lane << 5 | row), and the flush reads its date and revenue words from
// memory beside the brand; otherwise the decoded date and revenue are queued.
constexpr uint32_t kQueue = 64;
constexpr size_t kQueueBytes = 16ull * 3u * kQueue * 4u;  // 16 waves of a 1024-item block

// Compact tables: p = part index, d = date index.  Reference hash tables
// (PHT): p = the part's ht_p slot (its bit was set: the slot holds
// [key, brand]), d = the date key, probed here in ht_d (PHT_2: an empty slot
// drops the row, else year = slot >> 32).
template <bool PHT, bool LATE, int DBW, int RBW>
__device__ __forceinline__ void q21_flush(const Q21Dev& a, uint32_t* q, uint32_t qn, const uint16_t* year_tab,
                                          const Q21Hash& h, int64_t* gs, uint64_t* gc) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint32_t l = threadIdx.x & 63u;
    if (l < qn) {
        const uint32_t p = q[l];
        uint32_t d, rev;
        if constexpr (LATE) {
            const uint64_t v = q[kQueue + l];
            const uint32_t lane = q[2 * kQueue + l] >> 5, r = q[2 * kQueue + l] & 31u;
            d = read_row<DBW>(a.date, v, lane, r);
            rev = read_row<RBW>(a.rev, v, lane, r);
            if constexpr (!PHT) d -= a.keys_min;
        } else {
            d = q[kQueue + l];
            rev = q[2 * kQueue + l];
        }
        if constexpr (PHT) {
            const uint64_t ds =
                *reinterpret_cast<const uint64_t*>(a.ht_d + 2u * fast_mod(d - a.d_min, a.d_div));
            if (ds != 0) {
                const uint32_t brand = a.ht_p[2u * p + 1u];
                const uint32_t slot = fast_mod(brand * a.n_years + ((uint32_t)(ds >> 32) - a.year_base), a.slot_div);
                h.add(slot, rev, gs, gc);
            }
        } else if (d < a.n_date) {
            const uint32_t slot = fast_mod(a.brand[p] * a.n_years + (year_tab[d] - a.year_base), a.slot_div);
            h.add(slot, rev, gs, gc);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// PHT: the probe of the reference's hash tables with their occupancy bitmaps
// (supplier and part bits staged in LDS like the compact bitmaps, indexed by
// the tables' slot (key - min) % len); only the rows that pass both read the
// part table (brand) and the date table.  The same rows qualify as in the
// Crystal-primitive probe, which reads a table slot per row.
// With fixed-width date and revenue columns (LATE) only the supplier and part
// columns are read whole; a surviving row's date and revenue are read at the
// flush (about 17 rows in 2,048 qualify, so most of those two columns' 128-byte
// lines are never fetched).
template <int DBW, int SBW, int PBW, int RBW, bool USE_LDS, bool PHT = false>
__global__ __launch_bounds__(1024) void q21_compact_kernel(const Q21Dev a) {
    constexpr bool LATE = DBW >= 0 && RBW >= 0;
    extern __shared__ __align__(16) uint32_t lds[];
    // the group hash table first (8-byte fields at offset 0), then the
    // queues, then the staged tables
    Q21Hash h;
    h.init(lds);
    uint32_t* q = lds + Q21Hash::kBytes / 4u + (threadIdx.x >> 6) * 3u * kQueue;
    uint32_t* stage = lds + (Q21Hash::kBytes + kQueueBytes) / 4u;
    const uint32_t* supp_bm = PHT ? a.occ_s : a.supp_bm;
    const uint32_t* part_bm = PHT ? a.occ_p : a.part_bm;
    const uint16_t* year_tab = a.year;
    int64_t* gs = a.sums;  // where this block's groups go
    uint64_t* gc = a.counts;
    if (a.copies) {
        unsigned long long* cp = a.copies + (uint64_t)(blockIdx.x % kQ21Copies) * (2u * kQ21MaxSlots);
        gs = reinterpret_cast<int64_t*>(cp);
        gc = reinterpret_cast<uint64_t*>(cp + kQ21MaxSlots);
    }
    if constexpr (USE_LDS) {
        // each bitmap with one more zero word: the compact tables' bit n is zero
        // (the sentinel of bitmap_filter_min)
        uint32_t* s = stage;
        uint32_t* p = stage + a.supp_words + 1u;
        uint16_t* y = reinterpret_cast<uint16_t*>(p + a.part_words + 1u);
        for (uint32_t i = threadIdx.x; i <= a.supp_words; i += blockDim.x) s[i] = i < a.supp_words ? supp_bm[i] : 0u;
        for (uint32_t i = threadIdx.x; i <= a.part_words; i += blockDim.x) p[i] = i < a.part_words ? part_bm[i] : 0u;
        if constexpr (!PHT)
            for (uint32_t i = threadIdx.x; i < a.n_date; i += blockDim.x) y[i] = a.year[i];
        supp_bm = s;
        part_bm = p;
        year_tab = y;
    }
    __syncthreads();
    // rows of the (possibly partial) last vector, computed once
    const uint32_t tail =
        live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    uint32_t qn = 0;  // wave-uniform; the queue carries over from step to step
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * 32u + (threadIdx.x >> 5);
        // the whole wave stays in step (the queue flush needs every lane): a
        // half-wave past the last vector reads the last vector again (in
        // bounds, loads unconditional) and keeps no rows
        const bool vlive = v < a.n_vec;
        if (!__any(vlive)) continue;
        const uint64_t vl = vlive ? v : a.n_vec - 1;
        LaneCol<SBW> sc;
        LaneCol<PBW> pc;
        LaneCol<LATE ? 0 : DBW> dc;
        LaneCol<LATE ? 0 : RBW> rc;
        sc.load(a.supp, vl);
        pc.load(a.part, vl);
        if constexpr (!LATE) {
            dc.load(a.date, vl);
            rc.load(a.rev, vl);
        }
        // supplier semi-join on every row
        const uint32_t km = a.keys_min;
        uint32_t m = 0;
        if constexpr (PHT) {
#pragma unroll
            for (int r = 0; r < 32; ++r) m |= bit_at(supp_bm, fast_mod(sc.get(r) - a.s_min, a.s_div), 0u, a.s_len) << r;
        } else if constexpr (USE_LDS && SBW >= 0) {
            m = 0xFFFFFFFFu;
            bitmap_filter_min<16, SBW>(sc.w, supp_bm, sc.base - km, a.n_supp, m);
        } else {
#pragma unroll
            for (int r = 0; r < 32; ++r) m |= bit_at(supp_bm, sc.get(r), km, a.n_supp) << r;
        }
        if (!vlive) m = 0;
        if (v == a.n_vec - 1) m &= tail;
        pc.pin();
        if constexpr (!LATE) {
            dc.pin();
            rc.pin();
        }
        if (!__any(m != 0)) continue;
        // part semi-join: every row from the LDS bitmap; from L2 only the rows
        // that passed the supplier (rows no lane of the wave holds skipped)
        // part slot (PHT) or index (compact) of a key
        auto part_idx = [&](uint32_t key) { return PHT ? fast_mod(key - a.p_min, a.p_div) : key - km; };
        const uint32_t n_part_bits = PHT ? a.p_len : a.n_part;
        if constexpr (USE_LDS && !PHT && PBW >= 0) {
            bitmap_filter_min<16, PBW>(pc.w, part_bm, pc.base - km, n_part_bits, m);
        } else if constexpr (USE_LDS) {
            bitmap_filter_by<16>(pc, part_bm, part_idx, n_part_bits, m);
        } else {
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                if (__any((m >> r) & 1u)) {
                    const uint32_t p = part_idx(pc.get(r));
                    const bool live = (m >> r) & 1u;
                    const uint32_t hit = (live && p < n_part_bits) ? (part_bm[p >> 5] >> (p & 31u)) & 1u : 0u;
                    if (!hit) m &= ~(1u << r);
                }
            }
        }
        if (!__any(m != 0)) continue;
        // group-by on the rows that passed both (about 17 per 2,048 here):
        // compacted into the wave's LDS queue, then looked up and added with
        // one lane per row -- one brand round trip per queue, not per row slot
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const bool live = (m >> r) & 1u;
            const uint64_t bal = __ballot(live);
            if (bal == 0) continue;
            const uint32_t cnt = (uint32_t)__popcll(bal);
            if (qn + cnt > kQueue) {
                q21_flush<PHT, LATE, DBW, RBW>(a, q, qn, year_tab, h, gs, gc);
                qn = 0;
            }
            if (live) {
                const uint32_t pos =
                    qn + __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
                q[pos] = part_idx(pc.get(r));
                if constexpr (LATE) {
                    q[kQueue + pos] = (uint32_t)vl;  // n_vec < 2^32 (checked at launch)
                    q[2 * kQueue + pos] = dev::lane32() << 5 | (uint32_t)r;
                } else {
                    q[kQueue + pos] = PHT ? dc.get(r) : dc.get(r) - km;
                    q[2 * kQueue + pos] = rc.get(r);
                }
            }
            qn += cnt;
        }
    }
    if (qn) q21_flush<PHT, LATE, DBW, RBW>(a, q, qn, year_tab, h, gs, gc);
    h.flush(gs, gc);
}

// Occupancy bitmap of a reference hash table: lane i of a wave tests slot
// base + i, one ballot per 64 slots gives two bitmap words.
__global__ __launch_bounds__(256) void pht_occupancy_kernel(const uint32_t* __restrict ht, uint32_t n, uint32_t pair,
                                                            uint32_t* __restrict bits) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t waves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint32_t words = (n + 31u) / 32u;
    for (uint64_t base = wave * 64u; base < n; base += waves * 64u) {  // uniform per wave
        const uint64_t slot = base + lane;
        bool occ = false;
        if (slot < n) occ = pair ? (ht[2u * slot] | ht[2u * slot + 1u]) != 0u : ht[slot] != 0u;
        const uint64_t b = __builtin_amdgcn_ballot_w64(occ);
        const uint32_t w = (uint32_t)(base >> 5);
        if (lane == 0u) bits[w] = (uint32_t)b;
        if (lane == 1u && w + 1u < words) bits[w + 1u] = (uint32_t)(b >> 32);
    }
}

// Two 16-bit values per register, row r in half r & 1 of register r >> 1.
__device__ __forceinline__ void put16(uint32_t (&p)[16], int r, uint32_t v) { p[r >> 1] |= v << (16 * (r & 1)); }
__device__ __forceinline__ uint32_t get16(const uint32_t (&p)[16], int r) { return (p[r >> 1] >> (16 * (r & 1))) & 0xFFFFu; }

// Compact-table probe over columns of any encoding and width: one column
// decoded at a time (decode_regs, 32 registers), with only the row mask and
// the rows' brands and years (16 bits each) carried between columns.
template <bool USE_LDS>
__global__ __launch_bounds__(1024) void q21_any_kernel(const Q21Dev a) {
    extern __shared__ uint32_t lds[];
    const uint32_t* supp_bm = a.supp_bm;
    const uint32_t* part_bm = a.part_bm;
    const uint16_t* year_tab = a.year;
    if constexpr (USE_LDS) {
        uint32_t* s = lds;
        uint32_t* p = lds + a.supp_words;
        uint16_t* y = reinterpret_cast<uint16_t*>(p + a.part_words);
        for (uint32_t i = threadIdx.x; i < a.supp_words; i += blockDim.x) s[i] = a.supp_bm[i];
        for (uint32_t i = threadIdx.x; i < a.part_words; i += blockDim.x) p[i] = a.part_bm[i];
        for (uint32_t i = threadIdx.x; i < a.n_date; i += blockDim.x) y[i] = a.year[i];
        __syncthreads();
        supp_bm = s;
        part_bm = p;
        year_tab = y;
    }
    const uint32_t tail = live_row_mask((uint32_t)(a.n_rows - (a.n_vec - 1) * 1024u), a.rev.transposed != 0);
    const uint32_t km = a.keys_min;
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * 32u + (threadIdx.x >> 5);
        if (v >= a.n_vec) continue;
        uint32_t x[32];
        uint32_t m = (v == a.n_vec - 1) ? tail : 0xFFFFFFFFu;
        decode_regs(a.supp, v, x);
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            m &= ~((bit_at(supp_bm, x[r], km, a.n_supp) ^ 1u) << r);
        }
        if (!__any(m != 0)) continue;
        decode_regs(a.part, v, x);
        uint32_t b16[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            m &= ~((bit_at(part_bm, x[r], km, a.n_part) ^ 1u) << r);
        }
        if (!__any(m != 0)) continue;
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const bool live = (m >> r) & 1u;
            put16(b16, r, live ? a.brand[x[r] - km] : 0u);  // passed the part bitmap: key in range
            if ((r & 7) == 7) __builtin_amdgcn_sched_barrier(0);  // 8 rows' loads in flight
        }
        decode_regs(a.date, v, x);
        uint32_t y16[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const uint32_t k = x[r] - km;
            const bool ok = k < a.n_date;
            put16(y16, r, ok ? year_tab[k] : 0u);
            if (!ok) m &= ~(1u << r);
        }
        if (!__any(m != 0)) continue;
        decode_regs(a.rev, v, x);
#pragma unroll
        for (int r = 0; r < 32; ++r)
            if ((m >> r) & 1u) group_add(a, get16(b16, r), get16(y16, r), x[r]);
    }
}

template <typename Kernel>
int launch_resident_q21(Kernel k, int block, size_t lds, const Q21Dev& a, hipStream_t s) {
    int per_cu = 0;
    FLS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, block, lds));
    if (per_cu < 1) per_cu = 1;
    uint32_t cus = 0;
    int st = device_cus(&cus);
    if (st) return st;
    const uint32_t grid = balanced_grid((uint64_t)cus * per_cu, a.n_groups);
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), lds, s, a);
    return check_launch();
}

// The output of the paths that add to it directly: zeroed first unless the
// caller accumulates.
int zero_groups(const Q21Dev& a, uint32_t flags, hipStream_t s) {
    if (flags & FLS_Q_ACCUMULATE) return FLS_OK;
    const int st = zero_u64(reinterpret_cast<uint64_t*>(a.sums), a.n_slots, s);
    return st ? st : zero_u64(a.counts, a.n_slots, s);
}

// This call's set of partial tables: the device address of q21_copies (looked
// up once per device, with that device current) plus the next set in
// round-robin order.
int q21_set(unsigned long long** out) {
    static std::mutex mu;
    static std::vector<unsigned long long*> base;  // per device
    static std::atomic<uint32_t> next{0};
    int dev = 0;
    FLS_HIP_CHECK(hipGetDevice(&dev));
    unsigned long long* b = nullptr;
    {
        std::lock_guard<std::mutex> lock(mu);
        if ((int)base.size() <= dev) base.resize(dev + 1, nullptr);
        if (!base[dev]) {
            void* p = nullptr;
            FLS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(q21_copies)));
            base[dev] = static_cast<unsigned long long*>(p);
        }
        b = base[dev];
    }
    *out = b + (uint64_t)(next.fetch_add(1u) % kQ21Sets) * kQ21SetWords;
    return FLS_OK;
}

template <int DBW, int SBW, int PBW, int RBW, bool PHT = false>
int launch_compact(Q21Dev a, bool in_lds, size_t lds, uint32_t flags, hipStream_t s) {
    FLS_RETURN_IF(a.n_vec >> 32, FLS_ERR_ARG);  // the row queue holds a vector index in 32 bits
    auto k = in_lds ? q21_compact_kernel<DBW, SBW, PBW, RBW, true, PHT>
                    : q21_compact_kernel<DBW, SBW, PBW, RBW, false, PHT>;
    const size_t bytes = Q21Hash::kBytes + kQueueBytes + (in_lds ? lds : 0);
    FLS_HIP_CHECK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    const bool spread = a.n_slots <= kQ21MaxSlots;
    int st = spread ? q21_set(&a.copies) : zero_groups(a, flags, s);
    if (st) return st;
    if ((st = launch_resident_q21(k, 1024, bytes, a, s)) != FLS_OK || !spread) return st;
    hipLaunchKernelGGL(q21_collect_kernel, dim3((a.n_slots + 255u) / 256u), dim3(256), 0, s, a.copies, a.n_slots,
                       a.sums, a.counts, (flags & FLS_Q_ACCUMULATE) ? 1u : 0u);
    return check_launch();
}

// allow_var: the compact probe decodes any encoding (decode_regs); the
// reference-table probe reads BITPACK columns only.  Either way the four
// columns decode in one row order.
int fill_fact(Q21Dev& a, const fls_column* date, const fls_column* supp, const fls_column* part,
              const fls_column* rev, uint64_t n_rows, bool allow_var) {
    if (!date || !supp || !part || !rev) return FLS_ERR_NULL;
    int st;
    if ((st = to_dev_column(*date, n_rows, &a.date, allow_var)) != FLS_OK) return st;
    if ((st = to_dev_column(*supp, n_rows, &a.supp, allow_var)) != FLS_OK) return st;
    if ((st = to_dev_column(*part, n_rows, &a.part, allow_var)) != FLS_OK) return st;
    if ((st = to_dev_column(*rev, n_rows, &a.rev, allow_var)) != FLS_OK) return st;
    if (a.date.transposed != a.supp.transposed || a.date.transposed != a.part.transposed ||
        a.date.transposed != a.rev.transposed)
        return FLS_ERR_ARG;
    a.n_rows = n_rows;
    a.n_vec = ceil_div(n_rows, 1024);
    return FLS_OK;
}

int fill_groups(Q21Dev& a, const fls_q21_params* p, int64_t* d_sums, uint64_t* d_counts) {
    if (!p || !d_sums || !d_counts) return FLS_ERR_NULL;
    if (p->n_slots == 0 || p->n_years == 0) return FLS_ERR_ARG;
    a.year_base = p->year_base;
    a.n_years = p->n_years;
    a.n_slots = p->n_slots;
    a.slot_div = make_fastdiv(p->n_slots);
    a.sums = d_sums;
    a.counts = d_counts;
    return FLS_OK;
}

// Shared by fls_q21_build and the allocation-free entry point: clear both
// bitmaps with a kernel, then reduce the dimensions into the workspace.
int q21_build_into(const fls_q21_dims* dims, const fls_q21_params* params, char* ws, const Q21Layout& L,
                   hipStream_t s) {
    const uint32_t words = (uint32_t)(L.brand / 4u);  // both bitmaps
    hipLaunchKernelGGL(q21_zero_words, dim3(256), dim3(256), 0, s, reinterpret_cast<uint32_t*>(ws), words);
    hipLaunchKernelGGL(q21_build_compact, dim3(1024), dim3(256), 0, s, dims->s_region, dims->n_supp,
                       dims->p_category, dims->p_brand, dims->n_part, dims->d_year, dims->n_date, params->region,
                       params->category, reinterpret_cast<uint32_t*>(ws + L.supp_bm),
                       reinterpret_cast<uint32_t*>(ws + L.part_bm), reinterpret_cast<uint16_t*>(ws + L.brand),
                       reinterpret_cast<uint16_t*>(ws + L.year));
    return check_launch();
}

int q21_check_dims(const fls_q21_dims* dims) {
    FLS_RETURN_IF(!dims->s_region || !dims->p_category || !dims->p_brand || !dims->d_year, FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->n_supp || !dims->n_part || !dims->n_date, FLS_ERR_ARG);
    return FLS_OK;
}

}  // namespace

// ------------------------------------------------------- static scratch
// Device memory of the allocation-free Q2.1 entry point, which takes no
// workspace: its compact dimension tables are built here.  One copy per GPU,
// held until the process ends.  In host code `scratch` names a host-side
// stand-in, not device memory: kernels get the address static_scratch()
// looks up for the current device.
__device__ __align__(256) unsigned char scratch[FLS_Q21_STATIC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

// Device address of the current GPU's copy of `scratch`, looked up on the
// first call on each device (with that device current) and kept.
int static_scratch(void** ptr) {
    int dev = 0;
    FLS_HIP_CHECK(hipGetDevice(&dev));
    FLS_RETURN_IF(dev < 0 || dev >= kMaxDevices, FLS_ERR_NO_DEVICE);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FLS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FLS_RETURN_IF(!p, FLS_ERR_HIP);
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return FLS_OK;
}
}  // namespace
}  // namespace fls

extern "C" {

int fls_build_pht_u32(const uint32_t* d_filter, uint32_t filter_eq, const uint32_t* d_keys, const uint32_t* d_vals,
                      uint32_t n, uint32_t* d_ht, uint32_t ht_len, uint32_t keys_min, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_keys || !d_ht, FLS_ERR_NULL);
    FLS_RETURN_IF(ht_len == 0, FLS_ERR_ARG);
    const uint32_t tile = kBuildThreads * kBuildItems;
    hipLaunchKernelGGL(build_pht_kernel, dim3((n + tile - 1) / tile), dim3(kBuildThreads), 0, as_stream(stream),
                       d_filter, filter_eq, d_keys, d_vals, n, d_ht, ht_len, keys_min);
    return check_launch();
}

int fls_pht_occupancy_u32(const uint32_t* d_ht, uint32_t ht_len, uint32_t pair, uint32_t* d_bits, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!d_ht || !d_bits, FLS_ERR_NULL);
    FLS_RETURN_IF(ht_len == 0, FLS_ERR_ARG);
    const uint32_t waves = (uint32_t)ceil_div(ht_len, 64);
    const uint64_t want = ceil_div(waves, 4);
    const uint32_t grid = (uint32_t)(want < 4096 ? want : 4096);
    hipLaunchKernelGGL(pht_occupancy_kernel, dim3(grid), dim3(256), 0, as_stream(stream), d_ht, ht_len, pair ? 1u : 0u,
                       d_bits);
    return check_launch();
}

int fls_query_q21_pht(const fls_column* datekey, const fls_column* suppkey, const fls_column* partkey,
                      const fls_column* revenue, uint64_t n_rows, const fls_q21_pht* ht, const fls_q21_params* params,
                      int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!ht || !ht->ht_s || !ht->ht_p || !ht->ht_d, FLS_ERR_NULL);
    FLS_RETURN_IF(!ht->s_len || !ht->p_len || !ht->d_len, FLS_ERR_ARG);
    Q21Dev a{};
    int st = fill_fact(a, datekey, suppkey, partkey, revenue, n_rows, false);
    if (st) return st;
    hipStream_t s = as_stream(stream);
    if ((st = fill_groups(a, params, d_sums, d_counts)) != FLS_OK) return st;
    if (n_rows == 0) return zero_groups(a, flags, s);
    a.ht_s = ht->ht_s;
    a.ht_p = ht->ht_p;
    a.ht_d = ht->ht_d;
    a.s_len = ht->s_len;
    a.p_len = ht->p_len;
    a.d_len = ht->d_len;
    a.s_min = ht->s_min;
    a.p_min = ht->p_min;
    a.d_min = ht->d_min;
    const bool fast = a.date.bw == 12 && a.supp.bw == 16 && a.part.bw == 20 && a.rev.bw == 24;
    if (ht->occ_s && ht->occ_p && fast) {
        // bit-filtered probe over the same tables (SSB widths; others take the per-row probe)
        a.occ_s = ht->occ_s;
        a.occ_p = ht->occ_p;
        a.s_div = make_fastdiv(ht->s_len);
        a.p_div = make_fastdiv(ht->p_len);
        a.d_div = make_fastdiv(ht->d_len);
        a.supp_words = (ht->s_len + 31u) / 32u;
        a.part_words = (ht->p_len + 31u) / 32u;
        a.n_groups = ceil_div(a.n_vec, 32);
        const size_t lds = 4ull * (a.supp_words + a.part_words) + 16u;
        const bool in_lds = lds + Q21Hash::kBytes + kQueueBytes <= 150u * 1024u;
        return launch_compact<12, 16, 20, 24, true>(a, in_lds, lds, flags, s);
    }
    if ((st = zero_groups(a, flags, s)) != FLS_OK) return st;
    a.n_groups = ceil_div(a.n_vec, 8);
    if (fast) return launch_resident_q21(q21_pht_kernel<12, 16, 20, 24>, 256, 0, a, s);
    return launch_resident_q21(q21_pht_kernel<-1, -1, -1, -1>, 256, 0, a, s);
}

uint64_t fls_q21_workspace_bytes(uint32_t n_supp, uint32_t n_part, uint32_t n_date) {
    return fls::q21_layout(n_supp, n_part, n_date).total;
}

int fls_q21_build(const fls_q21_dims* dims, const fls_q21_params* params, void* d_workspace, uint64_t ws_bytes,
                  void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !params || !d_workspace, FLS_ERR_NULL);
    if (const int st = q21_check_dims(dims)) return st;
    const Q21Layout L = q21_layout(dims->n_supp, dims->n_part, dims->n_date);
    FLS_RETURN_IF(ws_bytes < L.total, FLS_ERR_CAPACITY);
    return q21_build_into(dims, params, static_cast<char*>(d_workspace), L, as_stream(stream));
}

int fls_query_q21(const fls_column* datekey, const fls_column* suppkey, const fls_column* partkey,
                  const fls_column* revenue, uint64_t n_rows, const fls_q21_dims* dims, const void* d_workspace,
                  const fls_q21_params* params, int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !d_workspace, FLS_ERR_NULL);
    FLS_RETURN_IF(!dims->n_supp || !dims->n_part || !dims->n_date, FLS_ERR_ARG);
    Q21Dev a{};
    int st = fill_fact(a, datekey, suppkey, partkey, revenue, n_rows, true);
    if (st) return st;
    hipStream_t s = as_stream(stream);
    if ((st = fill_groups(a, params, d_sums, d_counts)) != FLS_OK) return st;
    if (n_rows == 0) return zero_groups(a, flags, s);
    const Q21Layout L = q21_layout(dims->n_supp, dims->n_part, dims->n_date);
    const char* ws = static_cast<const char*>(d_workspace);
    a.supp_bm = reinterpret_cast<const uint32_t*>(ws + L.supp_bm);
    a.part_bm = reinterpret_cast<const uint32_t*>(ws + L.part_bm);
    a.brand = reinterpret_cast<const uint16_t*>(ws + L.brand);
    a.year = reinterpret_cast<const uint16_t*>(ws + L.year);
    a.n_supp = dims->n_supp;
    a.n_part = dims->n_part;
    a.n_date = dims->n_date;
    a.keys_min = dims->keys_min;
    a.supp_words = L.supp_words;
    a.part_words = L.part_words;
    a.n_groups = ceil_div(a.n_vec, 32);
    const size_t lds = 4ull * (L.supp_words + L.part_words) + 2ull * dims->n_date + 16u;
    const bool in_lds = lds + Q21Hash::kBytes + kQueueBytes <= 150u * 1024u;
    const bool fast = a.date.enc == FLS_ENC_BITPACK && a.supp.enc == FLS_ENC_BITPACK &&
                      a.part.enc == FLS_ENC_BITPACK && a.rev.enc == FLS_ENC_BITPACK && a.date.bw == 12 &&
                      a.supp.bw == 16 && a.part.bw == 20 && a.rev.bw == 24;
    if (fast) return launch_compact<12, 16, 20, 24>(a, in_lds, lds, flags, s);
    if ((st = zero_groups(a, flags, s)) != FLS_OK) return st;
    if (in_lds) {
        auto k = q21_any_kernel<true>;
        FLS_HIP_CHECK(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        return launch_resident_q21(k, 1024, lds, a, s);
    }
    return launch_resident_q21(q21_any_kernel<false>, 1024, 0, a, s);
}

int fls_query_q21_noalloc(const fls_column* datekey, const fls_column* suppkey, const fls_column* partkey,
                          const fls_column* revenue, uint64_t n_rows, const fls_q21_dims* dims,
                          const fls_q21_params* params, int64_t* d_sums, uint64_t* d_counts, uint32_t flags,
                          void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!dims || !params, FLS_ERR_NULL);
    if (const int st = q21_check_dims(dims)) return st;
    const Q21Layout L = q21_layout(dims->n_supp, dims->n_part, dims->n_date);
    FLS_RETURN_IF(L.total > FLS_Q21_STATIC_SCRATCH_BYTES, FLS_ERR_SCRATCH);
    Q21Dev probe{};  // every argument is checked before the first launch
    if (const int st = fill_fact(probe, datekey, suppkey, partkey, revenue, n_rows, true)) return st;
    FLS_RETURN_IF(!d_sums || !d_counts, FLS_ERR_NULL);
    FLS_RETURN_IF(params->n_slots == 0 || params->n_years == 0, FLS_ERR_ARG);
    void* ws = nullptr;
    if (const int st = static_scratch(&ws)) return st;
    if (const int st = q21_build_into(dims, params, static_cast<char*>(ws), L, as_stream(stream))) return st;
    return fls_query_q21(datekey, suppkey, partkey, revenue, n_rows, dims, ws, params, d_sums, d_counts, flags,
                         stream);
}

}  // extern "C"
