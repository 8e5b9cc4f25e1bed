// Crystal block primitives for gfx950 (wave64).  Header-only HIP.
//
// The names, template parameters and argument meaning of both reference
// flavours are kept, so a reference kernel body compiles against this header
// as written:
//   crystal/      (fastlanes/src/include/crystal/{load,pred,join,reduce,store,term}.cuh):
//                 BlockPred*<T, BLOCK_THREADS, ITEMS_PER_THREAD>, int flags;
//   crystal-opt/  (fastlanes/src/include/crystal-opt/*.cuh):
//                 BlockPred*<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD> and
//                 BlockPredLoad<T, ST, ...> with a flag type ST, the packed
//                 16 x 2 predicates, IsTerm;
//   fastlanes/    (fastlanes/src/include/fastlanes/join.cuh): the shared-result
//                 probe BlockProbeAndPHT_2_R_S.
// Both flavours coexist as overloads: their template parameter lists differ
// in kind (type vs. int) at the second position, so an explicit argument list
// selects exactly one.
//
// A "block" of BLOCK_THREADS work-items owns a tile of
// BLOCK_THREADS * ITEMS_PER_THREAD items in striped order: item j of
// work-item t is tile index t + j*BLOCK_THREADS.  The mapping onto CDNA differs
// from the reference's one-tile-per-block launch: a HIP workgroup may hold
// several tiles, work-item t of a tile being threadIdx.x % BLOCK_THREADS, so a
// FastLanes vector (BLOCK_THREADS = 32, ITEMS_PER_THREAD = 32) is one half-wave
// and a 64-wide wave serves two vectors.  Launched the reference's way (one
// tile per workgroup, blockDim.x == BLOCK_THREADS) the two agree exactly.
// BlockSum reduces inside the tile with cross-lane shuffles and, for tiles
// wider than a wave, one LDS word per wave.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {
namespace dev {

template <int BLOCK_THREADS>
__device__ __forceinline__ int tile_tid() {
    return (int)(threadIdx.x % BLOCK_THREADS);
}

// HASH(X, Y, Z) = ((X - Z) % Y)   (crystal/join.cuh:6)
template <typename K>
__device__ __forceinline__ int pht_hash(K x, int len, K keys_min) {
    return (int)((x - keys_min) % (K)len);
}

}  // namespace dev
}  // namespace fls

// The reference's hash macro (crystal/join.cuh:6, crystal-opt/join.cuh:25),
// for kernels that compute a slot themselves.
#ifndef HASH
#define HASH(X, Y, Z) ((X - Z) % Y)
#endif

// --------------------------------------------------------------- flags / load
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void InitFlags(int (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) selection_flags[i] = 1;
}

// crystal-opt/term.cuh:25-33, crystal/term.cuh:27: true when no flag is set.
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ bool IsTerm(const int (&selection_flags)[ITEMS_PER_THREAD]) {
    int any = 0;
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) any |= selection_flags[i];
    return any == 0;
}

// crystal/load.cuh:75-118, crystal-opt/load.cuh:68-97.  The Direct forms take
// the work-item index explicitly; BlockLoad passes the tile's own.
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoadDirect(const unsigned int tid, const T* block_itr,
                                                T (&items)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) items[i] = block_itr[tid + i * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoadDirect(const unsigned int tid, const T* block_itr,
                                                T (&items)[ITEMS_PER_THREAD], int num_items) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if ((int)tid + i * BLOCK_THREADS < num_items) items[i] = block_itr[tid + i * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoad(const T* block_itr, T (&items)[ITEMS_PER_THREAD], int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
        BlockLoadDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items);
    else
        BlockLoadDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items, num_items);
}

// Predicated loads: only items whose flag is set are read (crystal/load.cuh:
// 27-72 with int flags; crystal/load.cuh:170-210 and crystal-opt/load.cuh:25-65
// with a flag type ST).
template <typename T, typename ST, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoadDirect(const unsigned int tid, const T* block_itr,
                                                    T (&items)[ITEMS_PER_THREAD],
                                                    const ST (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if (selection_flags[i]) items[i] = block_itr[tid + i * BLOCK_THREADS];
}

template <typename T, typename ST, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoadDirect(const unsigned int tid, const T* block_itr,
                                                    T (&items)[ITEMS_PER_THREAD], int num_items,
                                                    const ST (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if ((int)tid + i * BLOCK_THREADS < num_items && selection_flags[i])
            items[i] = block_itr[tid + i * BLOCK_THREADS];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoadDirect(const unsigned int tid, const T* block_itr,
                                                    T (&items)[ITEMS_PER_THREAD],
                                                    const int (&selection_flags)[ITEMS_PER_THREAD]) {
    BlockPredLoadDirect<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items, selection_flags);
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoadDirect(const unsigned int tid, const T* block_itr,
                                                    T (&items)[ITEMS_PER_THREAD], int num_items,
                                                    const int (&selection_flags)[ITEMS_PER_THREAD]) {
    BlockPredLoadDirect<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items, num_items,
                                                                 selection_flags);
}

template <typename T, typename ST, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoad(const T* block_itr, T (&items)[ITEMS_PER_THREAD], int num_items,
                                              const ST (&selection_flags)[ITEMS_PER_THREAD]) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
        BlockPredLoadDirect<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items, selection_flags);
    else
        BlockPredLoadDirect<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items, num_items,
                                                                    selection_flags);
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoad(const T* block_itr, T (&items)[ITEMS_PER_THREAD], int num_items,
                                              const int (&selection_flags)[ITEMS_PER_THREAD]) {
    BlockPredLoad<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(block_itr, items, num_items, selection_flags);
}

// crystal/store.cuh:4-33, crystal-opt/store.cuh:25-69
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStoreDirect(int tid, T* block_itr, const T (&items)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) block_itr[tid + i * BLOCK_THREADS] = items[i];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStoreDirect(int tid, T* block_itr, const T (&items)[ITEMS_PER_THREAD],
                                                 int num_items) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if (tid + i * BLOCK_THREADS < num_items) block_itr[tid + i * BLOCK_THREADS] = items[i];
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStore(T* block_itr, const T (&items)[ITEMS_PER_THREAD], int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
        BlockStoreDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items);
    else
        BlockStoreDirect<T, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, block_itr, items, num_items);
}

// ------------------------------------------------------------------ predicates
// BlockPred sets flags, BlockPredAnd ANDs, BlockPredOr ORs; items at or past
// num_items keep their flag.  Generic forms take a selection functor
// (crystal/pred.cuh:13-119 with int flags; crystal-opt/pred.cuh:14-111 with a
// flag type ST for Pred and PredAnd).
namespace fls {
namespace dev {
template <typename T, typename ST, int BLOCK_THREADS, int ITEMS_PER_THREAD, typename Op, typename Combine>
__device__ __forceinline__ void block_pred_apply(int tid, const T (&items)[ITEMS_PER_THREAD],
                                                 ST (&flags)[ITEMS_PER_THREAD], int num_items, Op op,
                                                 Combine combine) {
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items) {
#pragma unroll
        for (int i = 0; i < ITEMS_PER_THREAD; ++i) flags[i] = (ST)combine((int)flags[i], op(items[i]) ? 1 : 0);
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS_PER_THREAD; ++i)
            if (tid + i * BLOCK_THREADS < num_items) flags[i] = (ST)combine((int)flags[i], op(items[i]) ? 1 : 0);
    }
}
struct SetFlag {
    __device__ int operator()(int, int p) const { return p; }
};
struct AndFlag {
    __device__ int operator()(int f, int p) const { return f && p; }
};
struct OrFlag {
    __device__ int operator()(int f, int p) const { return f || p; }
};
}  // namespace dev
}  // namespace fls

// Selection functors (crystal/pred.cuh:121-170, crystal-opt/pred.cuh:113-162).
template <typename T>
struct LessThan {
    T compare;
    __device__ __forceinline__ LessThan(T c) : compare(c) {}
    __device__ __forceinline__ bool operator()(const T& a) const { return a < compare; }
};
template <typename T>
struct GreaterThan {
    T compare;
    __device__ __forceinline__ GreaterThan(T c) : compare(c) {}
    __device__ __forceinline__ bool operator()(const T& a) const { return a > compare; }
};
template <typename T>
struct LessThanEq {
    T compare;
    __device__ __forceinline__ LessThanEq(T c) : compare(c) {}
    __device__ __forceinline__ bool operator()(const T& a) const { return a <= compare; }
};
template <typename T>
struct GreaterThanEq {
    T compare;
    __device__ __forceinline__ GreaterThanEq(T c) : compare(c) {}
    __device__ __forceinline__ bool operator()(const T& a) const { return a >= compare; }
};
template <typename T>
struct Eq {
    T compare;
    __device__ __forceinline__ Eq(T c) : compare(c) {}
    __device__ __forceinline__ bool operator()(const T& a) const { return a == compare; }
};

#define FLS_CRYSTAL_GENERIC(NAME, COMBINE)                                                                      \
    template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>                           \
    __device__ __forceinline__ void NAME##Direct(int tid, const T (&items)[ITEMS_PER_THREAD], SelectOp op,      \
                                                 int (&selection_flags)[ITEMS_PER_THREAD]) {                    \
        fls::dev::block_pred_apply<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(                                    \
            tid, items, selection_flags, BLOCK_THREADS * ITEMS_PER_THREAD, op, fls::dev::COMBINE());            \
    }                                                                                                           \
    template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>                           \
    __device__ __forceinline__ void NAME##Direct(int tid, const T (&items)[ITEMS_PER_THREAD], SelectOp op,      \
                                                 int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {     \
        fls::dev::block_pred_apply<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, selection_flags,        \
                                                                            num_items, op, fls::dev::COMBINE()); \
    }                                                                                                           \
    template <typename T, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>                           \
    __device__ __forceinline__ void NAME(const T (&items)[ITEMS_PER_THREAD], SelectOp op,                       \
                                         int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {             \
        fls::dev::block_pred_apply<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(                                    \
            fls::dev::tile_tid<BLOCK_THREADS>(), items, selection_flags, num_items, op, fls::dev::COMBINE());   \
    }

FLS_CRYSTAL_GENERIC(BlockPred, SetFlag)
FLS_CRYSTAL_GENERIC(BlockPredAnd, AndFlag)
FLS_CRYSTAL_GENERIC(BlockPredOr, OrFlag)
#undef FLS_CRYSTAL_GENERIC

// crystal-opt's flag-typed generic forms BlockPred / BlockPredAnd <T, ST, SelectOp, BT, IPT>.
#define FLS_CRYSTAL_GENERIC_ST(NAME, COMBINE)                                                                   \
    template <typename T, typename ST, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>              \
    __device__ __forceinline__ void NAME##Direct(int tid, const T (&items)[ITEMS_PER_THREAD], SelectOp op,      \
                                                 ST (&selection_flags)[ITEMS_PER_THREAD]) {                     \
        fls::dev::block_pred_apply<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD>(                                     \
            tid, items, selection_flags, BLOCK_THREADS * ITEMS_PER_THREAD, op, fls::dev::COMBINE());            \
    }                                                                                                           \
    template <typename T, typename ST, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>              \
    __device__ __forceinline__ void NAME##Direct(int tid, const T (&items)[ITEMS_PER_THREAD], SelectOp op,      \
                                                 ST (&selection_flags)[ITEMS_PER_THREAD], int num_items) {      \
        fls::dev::block_pred_apply<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, selection_flags,         \
                                                                           num_items, op, fls::dev::COMBINE()); \
    }                                                                                                           \
    template <typename T, typename ST, typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>              \
    __device__ __forceinline__ void NAME(const T (&items)[ITEMS_PER_THREAD], SelectOp op,                       \
                                         ST (&selection_flags)[ITEMS_PER_THREAD], int num_items) {              \
        fls::dev::block_pred_apply<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD>(                                     \
            fls::dev::tile_tid<BLOCK_THREADS>(), items, selection_flags, num_items, op, fls::dev::COMBINE());   \
    }

FLS_CRYSTAL_GENERIC_ST(BlockPred, SetFlag)
FLS_CRYSTAL_GENERIC_ST(BlockPredAnd, AndFlag)
#undef FLS_CRYSTAL_GENERIC_ST

// Named predicates.  Each exists as <T, BLOCK_THREADS, ITEMS_PER_THREAD> with
// int flags (crystal/pred.cuh:172-246) and as <T, ST, BLOCK_THREADS,
// ITEMS_PER_THREAD> with flags of type ST (the form crystal-opt's Q1.1 kernel
// calls, fls_q11.cu:51-63).
#define FLS_CRYSTAL_PRED(NAME, COMBINE, EXPR)                                                                   \
    template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>                                              \
    __device__ __forceinline__ void NAME(const T (&items)[ITEMS_PER_THREAD], T compare,                         \
                                         int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {             \
        fls::dev::block_pred_apply<T, int, BLOCK_THREADS, ITEMS_PER_THREAD>(                                    \
            fls::dev::tile_tid<BLOCK_THREADS>(), items, selection_flags, num_items,                             \
            [=](T x) { return EXPR; }, fls::dev::COMBINE());                                                    \
    }                                                                                                           \
    template <typename T, typename ST, int BLOCK_THREADS, int ITEMS_PER_THREAD>                                 \
    __device__ __forceinline__ void NAME(const T (&items)[ITEMS_PER_THREAD], T compare,                         \
                                         ST (&selection_flags)[ITEMS_PER_THREAD], int num_items) {              \
        fls::dev::block_pred_apply<T, ST, BLOCK_THREADS, ITEMS_PER_THREAD>(                                     \
            fls::dev::tile_tid<BLOCK_THREADS>(), items, selection_flags, num_items,                             \
            [=](T x) { return EXPR; }, fls::dev::COMBINE());                                                    \
    }

FLS_CRYSTAL_PRED(BlockPredGT, SetFlag, x > compare)
FLS_CRYSTAL_PRED(BlockPredLT, SetFlag, x < compare)
FLS_CRYSTAL_PRED(BlockPredGTE, SetFlag, x >= compare)
FLS_CRYSTAL_PRED(BlockPredLTE, SetFlag, x <= compare)
FLS_CRYSTAL_PRED(BlockPredEQ, SetFlag, x == compare)
FLS_CRYSTAL_PRED(BlockPredAndGT, AndFlag, x > compare)
FLS_CRYSTAL_PRED(BlockPredAndLT, AndFlag, x < compare)
FLS_CRYSTAL_PRED(BlockPredAndGTE, AndFlag, x >= compare)
FLS_CRYSTAL_PRED(BlockPredAndLTE, AndFlag, x <= compare)
FLS_CRYSTAL_PRED(BlockPredAndEQ, AndFlag, x == compare)
FLS_CRYSTAL_PRED(BlockPredOrEQ, OrFlag, x == compare)
#undef FLS_CRYSTAL_PRED

// ------------------------------------------------- packed 16 x 2 predicates
// crystal-opt/pred.cuh:241-459.  Each uint32 item holds two 16-bit values
// (low half first); a comparison yields 0xffff in every half where it holds
// and 0 elsewhere, except GreaterThan_int_16_2, which keeps only the lowest bit
// of each half (0x00010001 when both hold).  Only the first ITEMS_PER_THREAD/2
// words of items and flags are used.  Bounds, as the reference defines them:
// a partial tile (num_items != BLOCK_THREADS * ITEMS_PER_THREAD) updates word
// i of work-item t when t + i*BLOCK_THREADS < num_items/2 for the plain
// predicates and when t + i*BLOCK_THREADS < num_items/2 + 1 for the *X
// conjunctions.
namespace fls {
namespace dev {
__device__ __forceinline__ uint32_t halves_mask(bool lo, bool hi) {
    return (lo ? 0x0000ffffu : 0u) | (hi ? 0xffff0000u : 0u);
}
__device__ __forceinline__ int32_t s16(uint32_t v, int half) { return (int32_t)(int16_t)(uint16_t)(v >> (16 * half)); }
__device__ __forceinline__ uint32_t u16(uint32_t v, int half) { return (v >> (16 * half)) & 0xffffu; }
__device__ __forceinline__ uint32_t gt_s16x2(uint32_t a, uint32_t b) {
    return halves_mask(s16(a, 0) > s16(b, 0), s16(a, 1) > s16(b, 1));
}
__device__ __forceinline__ uint32_t lt_s16x2(uint32_t a, uint32_t b) {
    return halves_mask(s16(a, 0) < s16(b, 0), s16(a, 1) < s16(b, 1));
}
__device__ __forceinline__ uint32_t le_u16x2(uint32_t a, uint32_t b) {
    return halves_mask(u16(a, 0) <= u16(b, 0), u16(a, 1) <= u16(b, 1));
}
__device__ __forceinline__ uint32_t ge_u16x2(uint32_t a, uint32_t b) {
    return halves_mask(u16(a, 0) >= u16(b, 0), u16(a, 1) >= u16(b, 1));
}
}  // namespace dev
}  // namespace fls

// Both halves of the word hold the 16-bit constant.
__host__ __device__ constexpr uint32_t make_simd_const(uint16_t a_compare) {
    return ((uint32_t)a_compare << 16) | (uint32_t)a_compare;
}

struct GreaterThan_int_16_2 {
    uint32_t compare;
    __device__ __forceinline__ GreaterThan_int_16_2(uint32_t c) : compare(c) {}
    __device__ __forceinline__ uint32_t operator()(const uint32_t& a) const {
        return fls::dev::gt_s16x2(a, compare) & 0x00010001u;
    }
};
struct LessThan_int_16_2 {
    uint32_t compare;
    __device__ __forceinline__ LessThan_int_16_2(uint32_t c) : compare(c) {}
    __device__ __forceinline__ uint32_t operator()(const uint32_t& a) const { return fls::dev::lt_s16x2(a, compare); }
};
struct LessThanX {
    uint32_t compare;
    __device__ __forceinline__ LessThanX(uint32_t c) : compare(c) {}
    __device__ __forceinline__ uint32_t operator()(const uint32_t& a) const { return fls::dev::lt_s16x2(a, compare); }
};
struct LessThanEqX {
    uint32_t compare;
    __device__ __forceinline__ LessThanEqX(uint32_t c) : compare(c) {}
    __device__ __forceinline__ uint32_t operator()(const uint32_t& a) const { return fls::dev::le_u16x2(a, compare); }
};
struct GreaterThanEqX {
    uint32_t compare;
    __device__ __forceinline__ GreaterThanEqX(uint32_t c) : compare(c) {}
    __device__ __forceinline__ uint32_t operator()(const uint32_t& a) const { return fls::dev::ge_u16x2(a, compare); }
};

template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredDirect_int_16_2(int tid, const uint32_t (&items)[ITEMS_PER_THREAD],
                                                         SelectOp select_op,
                                                         uint32_t (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD / 2; ++i) selection_flags[i] = select_op(items[i]);
}

template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredDirect_int_16_2(int tid, const uint32_t (&items)[ITEMS_PER_THREAD],
                                                         SelectOp select_op,
                                                         uint32_t (&selection_flags)[ITEMS_PER_THREAD],
                                                         int num_items) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD / 2; ++i)
        if (tid + i * BLOCK_THREADS < num_items / 2) selection_flags[i] = select_op(items[i]);
}

template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPred_int_16_2(const uint32_t (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                                   uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
        BlockPredDirect_int_16_2<SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, select_op, selection_flags);
    else
        BlockPredDirect_int_16_2<SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, select_op, selection_flags,
                                                                            num_items);
}

// The four-argument BlockPredDirect_int_16_2 is the reference's dispatcher
// (crystal-opt/pred.cuh:306-319), identical to BlockPred_int_16_2.
template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredDirect_int_16_2(const uint32_t (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                                         uint32_t (&selection_flags)[ITEMS_PER_THREAD],
                                                         int num_items) {
    BlockPred_int_16_2<SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(items, select_op, selection_flags, num_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredGT_int_16_2(const uint32_t (&items)[ITEMS_PER_THREAD], uint32_t compare,
                                                     uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    BlockPred_int_16_2<GreaterThan_int_16_2, BLOCK_THREADS, ITEMS_PER_THREAD>(items, GreaterThan_int_16_2(compare),
                                                                              selection_flags, num_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLT_int_16_2(const uint32_t (&items)[ITEMS_PER_THREAD], uint32_t compare,
                                                     uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    BlockPred_int_16_2<LessThan_int_16_2, BLOCK_THREADS, ITEMS_PER_THREAD>(items, LessThan_int_16_2(compare),
                                                                           selection_flags, num_items);
}

template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndDirectX(int tid, const uint32_t (&items)[ITEMS_PER_THREAD],
                                                    SelectOp select_op,
                                                    uint32_t (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD / 2; ++i) selection_flags[i] &= select_op(items[i]);
}

template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndDirectX(int tid, const uint32_t (&items)[ITEMS_PER_THREAD],
                                                    SelectOp select_op, uint32_t (&selection_flags)[ITEMS_PER_THREAD],
                                                    int num_items) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD / 2; ++i)
        if (tid + i * BLOCK_THREADS < num_items / 2 + 1) selection_flags[i] &= select_op(items[i]);
}

template <typename SelectOp, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndX(const uint32_t (&items)[ITEMS_PER_THREAD], SelectOp select_op,
                                              uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items)
        BlockPredAndDirectX<SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, select_op, selection_flags);
    else
        BlockPredAndDirectX<SelectOp, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, select_op, selection_flags,
                                                                       num_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndLTX(const uint32_t (&items)[ITEMS_PER_THREAD], uint32_t compare,
                                                uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    BlockPredAndX<LessThanX, BLOCK_THREADS, ITEMS_PER_THREAD>(items, LessThanX(compare), selection_flags, num_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndLTEX(const uint32_t (&items)[ITEMS_PER_THREAD], uint32_t compare,
                                                 uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    BlockPredAndX<LessThanEqX, BLOCK_THREADS, ITEMS_PER_THREAD>(items, LessThanEqX(compare), selection_flags,
                                                                num_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredAndGTEX(const uint32_t (&items)[ITEMS_PER_THREAD], uint32_t compare,
                                                 uint32_t (&selection_flags)[ITEMS_PER_THREAD], int num_items) {
    BlockPredAndX<GreaterThanEqX, BLOCK_THREADS, ITEMS_PER_THREAD>(items, GreaterThanEqX(compare), selection_flags,
                                                                   num_items);
}

// -------------------------------------------------------- perfect hash tables
// PHT_1 (crystal/join.cuh:8-72): slot = HASH(key); a non-zero slot is a hit.
// A stored key of 0 is indistinguishable from an empty slot, so callers keep
// keys >= 1 (the FOR base of a dense 0-based key column is 1).
// The probes of a thread are issued kProbeBatch at a time before any result is
// used (an item that is not probed reads slot 0 and ignores it): written as
// one conditional load per item, every probe waits for the one before it.
constexpr int kProbeBatch = 8;

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_1(const K (&items)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht, int ht_len,
                                                   K keys_min, int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
#pragma unroll
    for (int i0 = 0; i0 < ITEMS_PER_THREAD; i0 += kProbeBatch) {
        constexpr int B = ITEMS_PER_THREAD < kProbeBatch ? ITEMS_PER_THREAD : kProbeBatch;
        K slot[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int i = i0 + j;
            const bool live = i < ITEMS_PER_THREAD && tid + i * BLOCK_THREADS < num_items && selection_flags[i];
            slot[j] = ht[live ? fls::dev::pht_hash(items[i], ht_len, keys_min) : 0];
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int i = i0 + j;
            if (i < ITEMS_PER_THREAD && tid + i * BLOCK_THREADS < num_items && selection_flags[i])
                selection_flags[i] = slot[j] != 0;
        }
    }
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_1(const K (&items)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht, int ht_len,
                                                   int num_items) {
    BlockProbeAndPHT_1<K, BLOCK_THREADS, ITEMS_PER_THREAD>(items, selection_flags, ht, ht_len, (K)0, num_items);
}

// PHT_2 (crystal/join.cuh:74-149): slot = [key, value] read as one 64-bit
// word; a hit yields value = slot >> 32.
template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_2(const K (&items)[ITEMS_PER_THREAD], V (&res)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht, int ht_len,
                                                   K keys_min, int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
#pragma unroll
    for (int i0 = 0; i0 < ITEMS_PER_THREAD; i0 += kProbeBatch) {
        constexpr int B = ITEMS_PER_THREAD < kProbeBatch ? ITEMS_PER_THREAD : kProbeBatch;
        uint64_t slot[B];
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int i = i0 + j;
            const bool live = i < ITEMS_PER_THREAD && tid + i * BLOCK_THREADS < num_items && selection_flags[i];
            const int h = live ? fls::dev::pht_hash(items[i], ht_len, keys_min) : 0;
            slot[j] = *reinterpret_cast<const uint64_t*>(&ht[h << 1]);
        }
#pragma unroll
        for (int j = 0; j < B; ++j) {
            const int i = i0 + j;
            if (i < ITEMS_PER_THREAD && tid + i * BLOCK_THREADS < num_items && selection_flags[i]) {
                if (slot[j] != 0) res[i] = (V)(slot[j] >> 32);
                else selection_flags[i] = 0;
            }
        }
    }
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_2(const K (&items)[ITEMS_PER_THREAD], V (&res)[ITEMS_PER_THREAD],
                                                   int (&selection_flags)[ITEMS_PER_THREAD], const K* ht, int ht_len,
                                                   int num_items) {
    BlockProbeAndPHT_2<K, V, BLOCK_THREADS, ITEMS_PER_THREAD>(items, res, selection_flags, ht, ht_len, (K)0,
                                                              num_items);
}

// Shared-result PHT_2 (fastlanes/join.cuh:6-81): the value of item j of
// work-item t goes to res[t + j*BLOCK_THREADS], a tile-sized array (in LDS in
// the reference's Q3.1 kernel) instead of a register.
template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_2_R_S(const K (&items)[ITEMS_PER_THREAD],
                                                       V (&res)[BLOCK_THREADS * ITEMS_PER_THREAD],
                                                       int (&selection_flags)[ITEMS_PER_THREAD], const K* ht,
                                                       int ht_len, K keys_min, int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    V vals[ITEMS_PER_THREAD];
    BlockProbeAndPHT_2<K, V, BLOCK_THREADS, ITEMS_PER_THREAD>(items, vals, selection_flags, ht, ht_len, keys_min,
                                                              num_items);
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if (tid + i * BLOCK_THREADS < num_items && selection_flags[i]) res[tid + i * BLOCK_THREADS] = vals[i];
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockProbeAndPHT_2_R_S(const K (&items)[ITEMS_PER_THREAD],
                                                       V (&res)[BLOCK_THREADS * ITEMS_PER_THREAD],
                                                       int (&selection_flags)[ITEMS_PER_THREAD], const K* ht,
                                                       int ht_len, int num_items) {
    BlockProbeAndPHT_2_R_S<K, V, BLOCK_THREADS, ITEMS_PER_THREAD>(items, res, selection_flags, ht, ht_len, (K)0,
                                                                  num_items);
}

// Builds (crystal/join.cuh:151-275): claim the slot with a CAS from 0, then
// (PHT_2) store the value beside it.
template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_1(const K (&keys)[ITEMS_PER_THREAD],
                                                         const int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, K keys_min, int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) {
        if (tid + i * BLOCK_THREADS < num_items && selection_flags[i]) {
            const int h = fls::dev::pht_hash(keys[i], ht_len, keys_min);
            atomicCAS(&ht[h], (K)0, keys[i]);
        }
    }
}

template <typename K, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_1(const K (&keys)[ITEMS_PER_THREAD],
                                                         const int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, int num_items) {
    BlockBuildSelectivePHT_1<K, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, selection_flags, ht, ht_len, (K)0, num_items);
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_2(const K (&keys)[ITEMS_PER_THREAD],
                                                         const V (&res)[ITEMS_PER_THREAD],
                                                         const int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, K keys_min, int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) {
        if (tid + i * BLOCK_THREADS < num_items && selection_flags[i]) {
            const int h = fls::dev::pht_hash(keys[i], ht_len, keys_min);
            atomicCAS(&ht[h << 1], (K)0, keys[i]);
            ht[(h << 1) + 1] = (K)res[i];
        }
    }
}

template <typename K, typename V, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockBuildSelectivePHT_2(const K (&keys)[ITEMS_PER_THREAD],
                                                         const V (&res)[ITEMS_PER_THREAD],
                                                         const int (&selection_flags)[ITEMS_PER_THREAD], K* ht,
                                                         int ht_len, int num_items) {
    BlockBuildSelectivePHT_2<K, V, BLOCK_THREADS, ITEMS_PER_THREAD>(keys, res, selection_flags, ht, ht_len, (K)0,
                                                                    num_items);
}

// ------------------------------------------------------------------- BlockSum
// crystal-opt/reduce.cuh:25-75: sum of one value per work-item over the tile.
// Every work-item of the tile receives the total (the reference guarantees it
// to work-item 0 only).  `shared` needs one T per wave of the tile when
// BLOCK_THREADS > 64; the whole workgroup must call it when the tile spans
// waves (it synchronises the workgroup).
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ T BlockSum(T item, T* shared) {
    static_assert(BLOCK_THREADS <= 64 || BLOCK_THREADS % 64 == 0, "tile must be part of, or whole, waves");
    T val = item;
    constexpr int width = BLOCK_THREADS < 64 ? BLOCK_THREADS : 64;
#pragma unroll
    for (int off = width / 2; off > 0; off >>= 1) val += __shfl_xor(val, off, 64);
    if constexpr (BLOCK_THREADS > 64) {
        constexpr int waves = BLOCK_THREADS / 64;
        const int tile_wave = (int)((threadIdx.x % BLOCK_THREADS) >> 6);
        const int tile_base = (int)((threadIdx.x / BLOCK_THREADS) * waves);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) shared[tile_base + tile_wave] = val;
        __syncthreads();
        T total = 0;
#pragma unroll
        for (int w = 0; w < waves; ++w) total += shared[tile_base + w];
        val = total;
    }
    return val;
}

template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ T BlockSum(const T (&items)[ITEMS_PER_THREAD], T* shared) {
    T thread_sum = 0;
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) thread_sum += items[i];
    return BlockSum<T, BLOCK_THREADS, ITEMS_PER_THREAD>(thread_sum, shared);
}
