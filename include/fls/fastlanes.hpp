// C++ drop-in for the reference's host-side FastLanes interface, implemented
// over the C ABI of libfls.so (fls.h).  Existing host code that calls
//   generated::pack::fallback::scalar::pack(in, out, bw)        (fls_gen/pack/pack.hpp:8-23)
//   generated::transpose::fallback::scalar::transpose_i/o(in, out)  (fls_gen/transpose/transpose.hpp:8-20)
//   generated::unrsum::fallback::scalar::unrsum(in, out)       (fls_gen/unrsum/unrsum.hpp:8-11)
// compiles unchanged against this header and links libfls.so instead of the
// generated pack.cpp / transpose.cpp / unrsum.cpp.
//
// The device side (unpack_device, d_rsum_32's in-register form rsum_lane,
// BlockLoad / BlockPred* / BlockProbeAndPHT_* / BlockBuildSelectivePHT_* /
// BlockSum) is header-only HIP in fls/device/*.hpp.
#pragma once

#include <stdexcept>
#include <string>

#include "fls/fls.h"

namespace fls {

inline void check(int status, const char* what) {
    if (status != FLS_OK) throw std::runtime_error(std::string(what) + ": " + fls_status_string(status));
}

}  // namespace fls

namespace generated {
namespace pack {
namespace fallback {
namespace scalar {
inline void pack(const uint32_t* __restrict in, uint32_t* __restrict out, uint8_t bw) {
    fls::check(fls_pack_u32(in, out, bw), "pack");
}
inline void pack(const int32_t* __restrict in, int32_t* __restrict out, uint8_t bw) {
    pack(reinterpret_cast<const uint32_t*>(in), reinterpret_cast<uint32_t*>(out), bw);
}
}  // namespace scalar
}  // namespace fallback
}  // namespace pack

namespace transpose {
namespace fallback {
namespace scalar {
inline void transpose_i(const uint32_t* __restrict in, uint32_t* __restrict out) {
    fls::check(fls_transpose_i_u32(in, out), "transpose_i");
}
inline void transpose_i(const int32_t* __restrict in, int32_t* __restrict out) {
    transpose_i(reinterpret_cast<const uint32_t*>(in), reinterpret_cast<uint32_t*>(out));
}
// The true inverse of transpose_i (the reference's u32 transpose_o repeats
// transpose_i's permutation).
inline void transpose_o(const uint32_t* __restrict in, uint32_t* __restrict out) {
    fls::check(fls_transpose_o_u32(in, out), "transpose_o");
}
}  // namespace scalar
}  // namespace fallback
}  // namespace transpose

namespace unrsum {
namespace fallback {
namespace scalar {
inline void unrsum(const uint32_t* in, uint32_t* out) { fls::check(fls_unrsum_u32(in, out), "unrsum"); }
}  // namespace scalar
}  // namespace fallback
}  // namespace unrsum
}  // namespace generated

// Grid-level decode under the reference's kernel name: the reference launches
// unpack_global<<<n_vec, 32>>>(in, out, bw) (unpack.cuh:3436); here it is a
// host call that picks the gfx950 launch shape itself.
inline void unpack_global(const uint32_t* d_in, uint32_t* d_out, uint64_t n_vec, uint8_t bw, void* stream = nullptr) {
    fls::check(fls_unpack_u32(d_in, d_out, n_vec, bw, stream), "unpack_global");
}
