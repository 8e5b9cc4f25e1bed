// Fused FastLanes decode + filter + aggregate, Q1.1 shape
// (reference: QueryKernel_v1 / QueryKernel_v2, fastlanes/src/ssb/fls_q11.cu:29-135,
// driver query<BT,IPT>() fls_q11.cu:137-171).
//
// Design for gfx950:
//  * persistent grid-stride kernel, 256 work-items = 8 vectors per step; a
//    wave decodes two vectors (work-item t: lane t%32 of vector t/32);
//  * every column is decoded straight into registers (fused form); the
//    FastLanes delta chain of a lane only touches that lane's own rows, so the
//    orderdate rsum needs no LDS;
//  * predicates become per-row lane masks that the compiler keeps in SGPR
//    pairs, so AND-ing predicates is scalar work;
//  * the int64 revenue and the row count are reduced per wave with DPP/shuffle,
//    per block through LDS, and across the blocks inside the launch (see
//    q11_block_finish): per-block sums spread over 64 lines, a count per line,
//    the last block of the last line collecting the lines.  Two 64-bit atomics
//    per block on the output instead (the reference's form) cost 8 us of the
//    120M-row scan's 85: 2,048 same-address atomics serialise.
#include <atomic>
#include <mutex>
#include <vector>

#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_runtime.hpp"
#include "fls_columns.hpp"

namespace fls {
namespace {

constexpr int kBlock = 256;
constexpr int kVecPerBlock = kBlock / 32;

struct Q11Dev {
    ColDev date, qty, disc, price;
    uint64_t n_rows, n_vec, n_groups;
    uint32_t date_gt, date_lt, qty_lt, disc_lo, disc_hi;
    int64_t* revenue;
    uint64_t* count;
    unsigned long long* slot;  // this call's reduction slot (zero at entry and exit)
    uint32_t accumulate;
};

// Cross-block reduction state: per slot, 64 group lines of 128 bytes ([0]
// revenue, [1] count, [2] low word: blocks of the group done) and one line
// whose low word counts the groups done.  A static per-device array of 64
// slots, taken round-robin per call: calls that overlap on one device use
// different slots, up to 64 of them at once.
constexpr uint32_t kQ11Groups = 64;
constexpr uint32_t kQ11SlotWords = 16u * (kQ11Groups + 1u);
constexpr uint32_t kQ11Slots = 64;
static __device__ __align__(256) unsigned long long q11_slots[kQ11Slots][kQ11SlotWords];

// The block's sums (per-wave partials in s_acc / s_cnt) into the launch's
// reduction; every work-item calls it after the block barrier, wave 0 does
// the work.  Block b adds its sums to line b % 64 (lanes 0 and 1, one atomic
// each, their returns waited for) and then counts itself there; the block
// that completes a line counts the line; the block that completes the last
// line reads and clears every line (a lane each) and writes, or with
// FLS_Q_ACCUMULATE adds, the results.  Each count wraps to zero as it
// completes (atomicInc), so the slot is zero again when the launch ends.
__device__ __forceinline__ void q11_block_finish(const Q11Dev& a, const uint64_t* s_acc, const uint64_t* s_cnt) {
    if (threadIdx.x >= 64) return;
    const uint32_t l = threadIdx.x, grid = gridDim.x, g = blockIdx.x % kQ11Groups;
    const uint32_t n_g = (grid - 1u - g) / kQ11Groups + 1u;  // blocks of line g
    const uint32_t n_lines = grid < kQ11Groups ? grid : kQ11Groups;
    unsigned long long* line = a.slot + 16u * g;
    uint32_t last = 0;
    if (l < 2) {  // a per-lane address: one instruction, both adds in flight together
        unsigned long long v = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) v += l ? s_cnt[w] : s_acc[w];
        const unsigned long long prev = atomicAdd(&line[l], v);
        asm volatile("" : : "v"(prev) : "memory");  // performed before the count below
    }
    if (l == 0 && atomicInc(reinterpret_cast<unsigned int*>(&line[2]), n_g - 1u) == n_g - 1u)
        last = atomicInc(reinterpret_cast<unsigned int*>(a.slot + 16u * kQ11Groups), n_lines - 1u) ==
               n_lines - 1u;
    if (!__shfl(last, 0, 64)) return;
    unsigned long long r = 0, c = 0;
    if (l < n_lines) {
        r = atomicExch(a.slot + 16u * l, 0ull);
        c = atomicExch(a.slot + 16u * l + 1u, 0ull);
    }
    r = dev::wave_sum(r);
    c = dev::wave_sum(c);
    if (l == 0) {
        if (a.accumulate) {
            atomicAdd(reinterpret_cast<unsigned long long*>(a.revenue), r);
            if (a.count) atomicAdd(reinterpret_cast<unsigned long long*>(a.count), c);
        } else {
            *reinterpret_cast<unsigned long long*>(a.revenue) = r;
            if (a.count) *a.count = c;
        }
    }
}

// Generic path: any encoding per column (bit-pack + FOR, delta, RLE, mixed),
// run-time widths.  Every column is walked row by row at its own width
// (column_walk): orderdate and quantity reduce straight to row masks,
// discount is kept as 32 values (it is both filtered and multiplied), and
// price is multiplied in place during its walk -- one 32-value array live,
// not one per column.  RLE / MIXED columns stage run values in LDS.
template <bool PRED>
__global__ __launch_bounds__(kBlock) void q11_kernel(const Q11Dev a) {
    __shared__ uint32_t stage[kVecPerBlock][1024];
    uint32_t* st = stage[threadIdx.x >> 5];
    uint64_t acc = 0;
    uint32_t cnt = 0;
    const bool transposed = a.price.transposed != 0;
    const uint32_t dlo = a.date_gt + 1u, dspan = a.date_lt - a.date_gt - 1u;
    const bool ok = a.date_lt > a.date_gt && a.disc_hi >= a.disc_lo;
    const uint32_t span = a.disc_hi - a.disc_lo, qty_lt = a.qty_lt, disc_lo = a.disc_lo;
    const auto* date_rm = reinterpret_cast<const fls_rle_meta*>(a.date.aux);
    const auto* qty_rm = reinterpret_cast<const fls_rle_meta*>(a.qty.aux);
    const auto* disc_rm = reinterpret_cast<const fls_rle_meta*>(a.disc.aux);
    const auto* price_rm = reinterpret_cast<const fls_rle_meta*>(a.price.aux);
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kVecPerBlock + (threadIdx.x >> 5);
        const bool live = v < a.n_vec;  // RLE / MIXED staging needs the whole wave, so no early exit
        uint32_t m = 0;
        if (live && ok) {
            const uint64_t first = v * 1024u;
            m = first + 1024u <= a.n_rows ? 0xffffffffu : live_row_mask((uint32_t)(a.n_rows - first), transposed);
        }
        m &= column_mask(a.date, date_rm, v, st, live, [dlo, dspan](uint32_t x) { return (x - dlo) < dspan; });
        if (PRED && !__any(m != 0)) continue;
        m &= column_mask(a.qty, qty_rm, v, st, live, [qty_lt](uint32_t x) { return x < qty_lt; });
        if (PRED && !__any(m != 0)) continue;
        uint32_t d[32];
        auto keep_disc = [&](int r, uint32_t x) { d[r] = x; };
        column_walk(a.disc, disc_rm, v, st, live, keep_disc);
        if (live) {
#pragma unroll
            for (int r = 0; r < 32; ++r) {
                m &= ~((uint32_t)((d[r] - disc_lo) > span) << r);
                d[r] &= 0u - ((m >> r) & 1u);  // rows that fail contribute 0
            }
        }
        if (PRED && !__any(m != 0)) continue;
        auto mac = [&](int r, uint32_t x) { acc += (uint64_t)x * (uint64_t)d[r]; };
        column_walk(a.price, price_rm, v, st, live, mac);
        cnt += (uint32_t)__builtin_popcount(m);
    }
    __shared__ uint64_t s_acc[kBlock / 64];
    __shared__ uint64_t s_cnt[kBlock / 64];
    acc = dev::wave_sum(acc);
    const uint64_t c64 = dev::wave_sum((uint64_t)cnt);
    const int wave = threadIdx.x >> 6;
    if (dev::is_wave_leader()) {
        s_acc[wave] = acc;
        s_cnt[wave] = c64;
    }
    __syncthreads();
    q11_block_finish(a, s_acc, s_cnt);
}

// Row mask of the existing rows of this work-item (bit r = row r).
__device__ __forceinline__ uint32_t valid_mask(uint64_t v, const Q11Dev& a, bool transposed) {
    const uint64_t first = v * 1024u;
    if (first + 1024u <= a.n_rows) return 0xffffffffu;
    const uint32_t live = (uint32_t)(a.n_rows - first), l = dev::lane32();
    uint32_t m = 0;
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        const uint32_t p = l + 32u * r;
        m |= (uint32_t)((transposed ? dev::untranspose_index(p) : p) < live) << r;
    }
    return m;
}

// Fast path: compile-time widths for quantity/discount/price, FastLanes delta
// orderdate.  Live state per work-item: the packed words of the three
// fixed-width columns (QBW+DBW+PBW VGPRs), one 32-bit date mask, accumulators.
// FULL: every column's words are loaded for every vector -- issued before the
// orderdate decode, so their latency overlaps the date's dependent
// vmeta -> words chain, and the next vector's date vmeta and base are
// fetched one iteration ahead; waves whose date mask is empty skip only the
// extraction/aggregation ALU work.  This holds about 127 VGPRs: 4 waves per
// SIMD, measured faster than 5 (96 VGPRs) and than 6 (where the words spill).
// PRED: later columns are not even loaded for such waves.
template <int QBW, int DBW, int PBW, bool PRED>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void q11_fast_kernel(const Q11Dev a) {
    uint64_t acc = 0;
    uint32_t cnt = 0;
    const uint32_t l = dev::lane32();
    const uint32_t dlo = a.date_gt + 1u, dspan = a.date_lt - a.date_gt - 1u;
    const bool date_ok = a.date_lt > a.date_gt, disc_ok = a.disc_hi >= a.disc_lo;
    const uint32_t span = a.disc_hi - a.disc_lo;
    // rows of the (possibly partial) last vector, computed once: one VGPR
    const uint32_t tail = valid_mask(a.n_vec - 1, a, true);
    fls_vmeta next_m = 0;  // the date column's vmeta and base of this work-item's next vector
    uint32_t next_b = 0;
    {
        const uint64_t v0 = (uint64_t)blockIdx.x * kVecPerBlock + (threadIdx.x >> 5);
        if (v0 < a.n_vec) {
            next_m = a.date.vmeta[v0];
            next_b = a.date.bases[v0 * 32u + l];
        }
    }
    for (uint64_t g = blockIdx.x; g < a.n_groups; g += gridDim.x) {
        const uint64_t v = g * kVecPerBlock + (threadIdx.x >> 5);
        const fls_vmeta cur_m = next_m;
        const uint32_t cur_b = next_b;
        const uint64_t vn = v + (uint64_t)gridDim.x * kVecPerBlock;
        if (vn < a.n_vec) {
            next_m = a.date.vmeta[vn];
            next_b = a.date.bases[vn * 32u + l];
        }
        if (v >= a.n_vec) continue;
        uint32_t qw[QBW > 0 ? QBW : 1], dw[DBW > 0 ? DBW : 1], pw[PBW > 0 ? PBW : 1];
        if constexpr (!PRED) {
            dev::load_lane<QBW>(a.qty.words + v * (32u * QBW) + l, qw);
            dev::load_lane<DBW>(a.disc.words + v * (32u * DBW) + l, dw);
            dev::load_lane<PBW>(a.price.words + v * (32u * PBW) + l, pw);
        }
        // orderdate: its per-vector width makes this a (half-wave divergent)
        // switch; widths above 16 take the run-time walk, so at most 16 of
        // its words are live next to the other columns' words
        uint32_t m = delta_mask_pre(a.date, cur_m, cur_b, [=](uint32_t x) { return (x - dlo) < dspan; });
        if (v == a.n_vec - 1) m &= tail;
        if (!date_ok || !disc_ok) m = 0;
        if constexpr (!PRED) {
            pin_words(qw);
            pin_words(dw);
            pin_words(pw);
            if (!__any(m != 0)) continue;
        } else {
            if (!__any(m != 0)) continue;
            dev::load_lane<QBW>(a.qty.words + v * (32u * QBW) + l, qw);
            dev::load_lane<DBW>(a.disc.words + v * (32u * DBW) + l, dw);
            dev::load_lane<PBW>(a.price.words + v * (32u * PBW) + l, pw);
        }
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            // branch-free: every term is a 0/1 integer, so the row loop stays
            // one straight-line block
            const uint32_t q = dev::extract_at<QBW>(qw, r) + a.qty.base;
            const uint32_t d = dev::extract_at<DBW>(dw, r) + a.disc.base;
            const uint32_t f = ((m >> r) & 1u) & (uint32_t)(q < a.qty_lt) & (uint32_t)((d - a.disc_lo) <= span);
            const uint32_t p = dev::extract_at<PBW>(pw, r) + a.price.base;
            acc += (uint64_t)p * (uint64_t)(d * f);
            cnt += f;
        }
    }
    __shared__ uint64_t s_acc[kBlock / 64];
    __shared__ uint64_t s_cnt[kBlock / 64];
    acc = dev::wave_sum(acc);
    const uint64_t c64 = dev::wave_sum((uint64_t)cnt);
    const int wave = threadIdx.x >> 6;
    if (dev::is_wave_leader()) {
        s_acc[wave] = acc;
        s_cnt[wave] = c64;
    }
    __syncthreads();
    q11_block_finish(a, s_acc, s_cnt);
}

template <typename Kernel>
int launch_resident(Kernel k, const Q11Dev& a, hipStream_t s) {
    uint32_t grid = 0;
    int st = resident_grid(k, kBlock, a.n_groups, &grid);
    if (st) return st;
    hipLaunchKernelGGL(k, dim3(grid), dim3(kBlock), 0, s, a);
    return check_launch();
}

// This call's reduction slot: the device address of q11_slots (looked up
// once per device, with that device current) plus the next slot in
// round-robin order.
int q11_slot(unsigned long long** out) {
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
            FLS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(q11_slots)));
            base[dev] = static_cast<unsigned long long*>(p);
        }
        b = base[dev];
    }
    *out = b + (uint64_t)(next.fetch_add(1u) % kQ11Slots) * kQ11SlotWords;
    return FLS_OK;
}

int launch_generic(const Q11Dev& a, bool pred, hipStream_t s) {
    if (pred) return launch_resident(q11_kernel<true>, a, s);
    return launch_resident(q11_kernel<false>, a, s);
}

template <int QBW, int DBW, int PBW>
int launch_fast(const Q11Dev& a, bool pred, hipStream_t s) {
    if (pred) return launch_resident(q11_fast_kernel<QBW, DBW, PBW, true>, a, s);
    return launch_resident(q11_fast_kernel<QBW, DBW, PBW, false>, a, s);
}

}  // namespace
}  // namespace fls

extern "C" int fls_query_q11(const fls_column* orderdate, const fls_column* quantity, const fls_column* discount,
                             const fls_column* price, uint64_t n_rows, const fls_q11_params* p, uint32_t flags,
                             int64_t* d_revenue, uint64_t* d_count, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!orderdate || !quantity || !discount || !price || !p || !d_revenue, FLS_ERR_NULL);
    Q11Dev a{};
    int st;
    if ((st = to_dev_column(*orderdate, n_rows, &a.date, true)) != FLS_OK) return st;
    if ((st = to_dev_column(*quantity, n_rows, &a.qty, true)) != FLS_OK) return st;
    if ((st = to_dev_column(*discount, n_rows, &a.disc, true)) != FLS_OK) return st;
    if ((st = to_dev_column(*price, n_rows, &a.price, true)) != FLS_OK) return st;
    // every column must use the same row order (delta / RLE / mixed imply transposed)
    const uint32_t tr = a.price.transposed != 0;
    FLS_RETURN_IF((a.date.transposed != 0) != tr || (a.qty.transposed != 0) != tr || (a.disc.transposed != 0) != tr,
                  FLS_ERR_ARG);
    a.n_rows = n_rows;
    a.n_vec = ceil_div(n_rows, 1024);
    a.n_groups = ceil_div(a.n_vec, kVecPerBlock);
    a.date_gt = p->date_gt;
    a.date_lt = p->date_lt;
    a.qty_lt = p->qty_lt;
    a.disc_lo = p->disc_lo;
    a.disc_hi = p->disc_hi;
    a.revenue = d_revenue;
    a.count = d_count;
    a.accumulate = (flags & FLS_Q_ACCUMULATE) != 0;
    hipStream_t s = as_stream(stream);
    if (n_rows == 0) {  // no launch writes the results: zero them unless adding
        if (a.accumulate) return FLS_OK;
        if ((st = zero_u64(reinterpret_cast<uint64_t*>(d_revenue), 1, s)) != FLS_OK) return st;
        return d_count ? zero_u64(d_count, 1, s) : FLS_OK;
    }
    if ((st = q11_slot(&a.slot)) != FLS_OK) return st;
    const bool pred = (flags & FLS_Q_PRED_LOAD) != 0;
    const bool fast = a.date.enc == FLS_ENC_DELTA && a.qty.enc == FLS_ENC_BITPACK && a.disc.enc == FLS_ENC_BITPACK &&
                      a.price.enc == FLS_ENC_BITPACK && a.qty.bw == 6 && a.disc.bw == 4 && a.price.bw == 24;
    if (fast) return launch_fast<6, 4, 24>(a, pred, s);
    return launch_generic(a, pred, s);
}
