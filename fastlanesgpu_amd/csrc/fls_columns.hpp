// Device view of an fls_column and the per-work-item column decoders shared by
// the fused query kernels.  Internal to libfls.so.
#pragma once

#include "fls/device/unpack.hpp"
#include "fls_runtime.hpp"

namespace fls {

struct ColDev {
    uint32_t enc, bw, base, transposed;
    const uint32_t* words;
    const fls_vmeta* vmeta;
    const uint32_t* bases;
};

// Validate a host-side column descriptor for n_rows rows and build its device
// view.  Fixed-width columns are bounds-checked against n_words.
int to_dev_column(const fls_column& c, uint64_t n_rows, ColDev* out, bool allow_delta);

// Grid size for a persistent kernel: min(work, CUs * blocks_per_cu).
int persistent_grid(uint64_t work, uint32_t blocks_per_cu, uint32_t* grid);

// Number of compute units of the current device (cached per device).
int device_cus(uint32_t* cus);

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

}  // namespace fls
