// Crystal block primitives for gfx950 (wave64).  Header-only HIP.
//
// Same names, template parameters and semantics as the reference's
// primitives (fastlanes/src/include/crystal/{load,pred,join,reduce,store}.cuh,
// crystal-opt/*.cuh): a "block" of BLOCK_THREADS work-items owns a tile of
// BLOCK_THREADS * ITEMS_PER_THREAD items in striped order, item j of
// work-item t being tile index t + j*BLOCK_THREADS.
//
// The difference is the mapping onto CDNA hardware: a HIP workgroup may hold
// several tiles.  Work-item t of a tile is threadIdx.x % BLOCK_THREADS, so a
// FastLanes vector (BLOCK_THREADS = 32, ITEMS_PER_THREAD = 32) is one
// half-wave and a 64-wide wave serves two vectors.  BlockSum reduces inside
// the tile with cross-lane shuffles and, for tiles wider than a wave, one LDS
// word per wave.
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

// --------------------------------------------------------------- flags / load
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void InitFlags(int (&selection_flags)[ITEMS_PER_THREAD]) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) selection_flags[i] = 1;
}

// crystal/load.cuh:27-118
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockLoad(const T* block_itr, T (&items)[ITEMS_PER_THREAD], int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items) {
#pragma unroll
        for (int i = 0; i < ITEMS_PER_THREAD; ++i) items[i] = block_itr[tid + i * BLOCK_THREADS];
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS_PER_THREAD; ++i)
            if (tid + i * BLOCK_THREADS < num_items) items[i] = block_itr[tid + i * BLOCK_THREADS];
    }
}

// crystal-opt/load.cuh:25-97: load only the items whose flag is set
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockPredLoad(const T* block_itr, T (&items)[ITEMS_PER_THREAD], int num_items,
                                              const int (&selection_flags)[ITEMS_PER_THREAD]) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if (tid + i * BLOCK_THREADS < num_items && selection_flags[i]) items[i] = block_itr[tid + i * BLOCK_THREADS];
}

// crystal/store.cuh:24
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void BlockStore(T* block_itr, const T (&items)[ITEMS_PER_THREAD], int num_items) {
    const int tid = fls::dev::tile_tid<BLOCK_THREADS>();
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if (tid + i * BLOCK_THREADS < num_items) block_itr[tid + i * BLOCK_THREADS] = items[i];
}

// ------------------------------------------------------------------ predicates
// crystal/pred.cuh:6-246: BlockPred<OP> sets flags, BlockPredAnd<OP> ANDs,
// BlockPredOr<OP> ORs; items beyond num_items are left untouched.
namespace fls {
namespace dev {
template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD, typename Op, typename Combine>
__device__ __forceinline__ void block_pred_apply(const T (&items)[ITEMS_PER_THREAD],
                                                 int (&flags)[ITEMS_PER_THREAD], int num_items, Op op,
                                                 Combine combine) {
    const int tid = tile_tid<BLOCK_THREADS>();
    if (BLOCK_THREADS * ITEMS_PER_THREAD == num_items) {
#pragma unroll
        for (int i = 0; i < ITEMS_PER_THREAD; ++i) flags[i] = combine(flags[i], op(items[i]) ? 1 : 0);
    } else {
#pragma unroll
        for (int i = 0; i < ITEMS_PER_THREAD; ++i)
            if (tid + i * BLOCK_THREADS < num_items) flags[i] = combine(flags[i], op(items[i]) ? 1 : 0);
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

#define FLS_CRYSTAL_PRED(NAME, COMBINE, EXPR)                                                                   \
    template <typename T, int BLOCK_THREADS, int ITEMS_PER_THREAD>                                              \
    __device__ __forceinline__ void NAME(const T (&items)[ITEMS_PER_THREAD], T compare,                         \
                                         int (&selection_flags)[ITEMS_PER_THREAD], int num_items) {             \
        fls::dev::block_pred_apply<T, BLOCK_THREADS, ITEMS_PER_THREAD>(                                         \
            items, selection_flags, num_items, [=](T x) { return EXPR; }, fls::dev::COMBINE());                 \
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

// ------------------------------------------------------------------- BlockSum
// crystal-opt/reduce.cuh:25-75: sum of one value per work-item over the tile.
// Every work-item of the tile receives the total.  `shared` needs one T per
// wave of the tile when BLOCK_THREADS > 64; the whole workgroup must call it
// when the tile spans waves (it synchronises the workgroup).
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
