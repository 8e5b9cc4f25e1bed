// Per-vector encoding classes and their exact encoded sizes (mixed column
// format).  Internal; host- and device-usable.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fls/fls.h"
#include "fls_pack.hpp"

namespace fls {

enum : uint32_t { MIX_BITPACK = FLS_MIX_BITPACK, MIX_FOR = FLS_MIX_FOR, MIX_DELTA = FLS_MIX_DELTA, MIX_RLE = FLS_MIX_RLE };

// Largest payload of one vector in 32-bit words: RLE with 32 index bases, a
// 1-bit index and 1024 runs of 32-bit values (32 + 32 + 1024).
constexpr uint64_t kMixMaxWords = 1088;

// Statistics from which every encoding's width and size follow.
struct VecStats {
    uint32_t w_bp;     // width of the largest value (bit-pack)
    uint32_t vmin;     // smallest value (FOR base, RLE value base)
    uint32_t w_for;    // width of max - min (FOR; also RLE run values)
    uint32_t w_delta;  // width of the largest FastLanes delta
    uint32_t n_runs;   // runs in the vector
    uint32_t ibw;      // 1 when some run starts inside a chain (RLE index width)
};

// Encoded bytes of one vector: payload words + 16 bytes of metadata
// (vmeta + aux), see fls.h FLS_ENC_MIXED.
__host__ __device__ inline uint32_t mix_bytes(uint32_t enc, const VecStats& s) {
    switch (enc) {
        case MIX_BITPACK: return 128u * s.w_bp + 16u;
        case MIX_FOR: return 128u * s.w_for + 16u;
        case MIX_DELTA: return 128u + 128u * s.w_delta + 16u;
        default: return 128u + 128u * s.ibw + 4u * ((s.n_runs * s.w_for + 31u) / 32u) + 16u;
    }
}

// 10 width buckets: 0 | 1-2 | 3-4 | 5-6 | 7-8 | 9-12 | 13-16 | 17-20 | 21-24 | 25-32
__host__ __device__ inline uint32_t width_bucket(uint32_t w) {
    return w == 0 ? 0u : w <= 8 ? (w + 1u) / 2u : w <= 24 ? 4u + (w - 5u) / 4u : 9u;
}
__host__ __device__ inline uint32_t mix_class(uint32_t enc, uint32_t w) { return enc * 10u + width_bucket(w); }
__host__ __device__ inline uint32_t mix_class_enc(uint32_t c) { return c / 10u; }

}  // namespace fls
