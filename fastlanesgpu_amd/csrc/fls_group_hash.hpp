// Per-block LDS hash table of group slots for the group-by probes.
//
// A block meets far fewer groups than the slot space holds (SSB Q2.1: 40
// brands x 7 years of one category among 7,000 slots; Q3.1: 150 of 4,375), so
// qualifying rows are summed here and each block adds its groups to the
// global table once.  Open addressing over 2^BITS entries with linear
// probing: key = slot + 1 (0 = empty, slot < 2^32 - 1), claimed with
// atomicCAS; a slot that finds no entry within kProbe probes goes to the
// global table directly, so any number of groups per block stays exact.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace fls {

template <uint32_t BITS>
struct GroupHash {
    static constexpr uint32_t kSize = 1u << BITS;
    static constexpr uint32_t kProbe = 32;
    static constexpr size_t kBytes = (size_t)kSize * (8u + 8u + 4u);

    unsigned long long *sum, *cnt;
    uint32_t* key;

    // base: 8-byte aligned LDS; every work-item of the block calls this, and
    // a barrier follows before the first add
    __device__ __forceinline__ void init(void* base) {
        sum = static_cast<unsigned long long*>(base);
        cnt = sum + kSize;
        key = reinterpret_cast<uint32_t*>(cnt + kSize);
        for (uint32_t i = threadIdx.x; i < kSize; i += blockDim.x) {
            key[i] = 0u;
            sum[i] = 0ull;
            cnt[i] = 0ull;
        }
    }

    __device__ __forceinline__ void add(uint32_t slot, uint64_t value, int64_t* __restrict g_sum,
                                        uint64_t* __restrict g_cnt) const {
        const uint32_t k = slot + 1u;
        uint32_t i = (slot * 0x9E3779B1u) >> (32u - BITS);
#pragma unroll 1
        for (uint32_t probe = 0; probe < kProbe; ++probe) {
            uint32_t cur = __atomic_load_n(&key[i], __ATOMIC_RELAXED);
            if (cur == 0u) {
                const uint32_t old = atomicCAS(&key[i], 0u, k);
                cur = old == 0u ? k : old;
            }
            if (cur == k) {
                atomicAdd(&sum[i], (unsigned long long)value);
                atomicAdd(&cnt[i], 1ull);
                return;
            }
            i = (i + 1u) & (kSize - 1u);
        }
        atomicAdd(reinterpret_cast<unsigned long long*>(&g_sum[slot]), (unsigned long long)value);
        atomicAdd(reinterpret_cast<unsigned long long*>(&g_cnt[slot]), 1ull);
    }

    // Every work-item of the block, after its last add: each claimed entry is
    // added to the global table.
    __device__ __forceinline__ void flush(int64_t* __restrict g_sum, uint64_t* __restrict g_cnt) const {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < kSize; i += blockDim.x) {
            const uint32_t k = key[i];
            if (k) {
                atomicAdd(reinterpret_cast<unsigned long long*>(&g_sum[k - 1u]), sum[i]);
                atomicAdd(reinterpret_cast<unsigned long long*>(&g_cnt[k - 1u]), cnt[i]);
            }
        }
    }
};

}  // namespace fls
