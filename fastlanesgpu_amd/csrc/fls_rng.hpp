// Counter-based generator shared by host and device code: splitmix64 mixing
// (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014) of a per-(seed, stream) key plus the element counter.  Every
// synthetic value is a pure function of (seed, stream, index).  Internal.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fls {
namespace rng {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream * 0x9E3779B97F4A7C15ull));
}

// uniform in [0, range), range <= 2^32
__host__ __device__ __forceinline__ uint32_t uniform(uint64_t key, uint64_t i, uint64_t range) {
    const uint64_t hi = mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull) >> 32;
    return (uint32_t)((hi * range) >> 32);
}

// Streams of the batched-segments workload (one key per segment and role).
constexpr uint64_t kSegParams = 1000;
constexpr uint64_t kSegValues = 100000;
constexpr uint64_t kSegRuns = 200000;
// Streams of the selector's training vectors.
constexpr uint64_t kTrainFamily = 300001;
constexpr uint64_t kTrainValues = 300002;
constexpr uint64_t kTrainNoise = 300003;
constexpr uint64_t kTrainRuns = 300004;

}  // namespace rng
}  // namespace fls
