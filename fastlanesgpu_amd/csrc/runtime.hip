// Status strings, HIP error capture and tiny utility kernels of the C ABI.
#include "fls_runtime.hpp"

namespace fls {

static thread_local int g_last_hip_error = 0;

int record_hip_error(hipError_t e) {
    g_last_hip_error = static_cast<int>(e);
    return FLS_ERR_HIP;
}

__global__ void zero_u64_kernel(uint64_t* __restrict d, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0;
}

int zero_u64(uint64_t* d, uint64_t n, hipStream_t stream) {
    if (n == 0) return FLS_OK;
    uint32_t blocks = (uint32_t)ceil_div(n, 256);
    zero_u64_kernel<<<blocks, 256, 0, stream>>>(d, n);
    return check_launch();
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16-byte-per-lane streaming copy, grid-stride.
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict src, u32x4* __restrict dst, uint64_t n16) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

}  // namespace fls

extern "C" {

int fls_copy_bytes(const void* d_src, void* d_dst, uint64_t n_bytes, void* stream) {
    if (n_bytes == 0) return FLS_OK;
    if (!d_src || !d_dst) return FLS_ERR_NULL;
    if ((n_bytes & 15u) || (((uintptr_t)d_src | (uintptr_t)d_dst) & 15u)) return FLS_ERR_ARG;
    const uint64_t n16 = n_bytes / 16u;
    uint64_t blocks = fls::ceil_div(n16, 256);
    if (blocks > 256u * 16u) blocks = 256u * 16u;
    hipLaunchKernelGGL(fls::copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, fls::as_stream(stream),
                       static_cast<const fls::u32x4*>(d_src), static_cast<fls::u32x4*>(d_dst), n16);
    return fls::check_launch();
}

const char* fls_status_string(int status) {
    switch (status) {
        case FLS_OK: return "ok";
        case FLS_ERR_BW: return "bit width outside 0..32";
        case FLS_ERR_SIZE: return "count or size out of range";
        case FLS_ERR_NULL: return "required pointer is NULL";
        case FLS_ERR_HIP: return "HIP runtime error";
        case FLS_ERR_ENCODING: return "unknown or unsupported encoding";
        case FLS_ERR_ARG: return "invalid argument";
        case FLS_ERR_CAPACITY: return "output buffer too small";
        case FLS_ERR_NO_DEVICE: return "no usable gfx950 device";
        default: return "unknown status";
    }
}

int fls_last_hip_error(void) { return fls::g_last_hip_error; }

const char* fls_version(void) { return "fls-mi355x 0.1"; }

}  // extern "C"
