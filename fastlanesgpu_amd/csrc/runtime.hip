// Status strings, HIP error capture and tiny utility kernels of the C ABI.
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
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

// 16-byte-per-lane streaming copy, one element per work-item over a grid of
// every element: on MI355X this reached 6.5 TB/s read + write from 16 MB to
// 1 GB, where grid-stride loops over a capped grid (2-32 blocks per CU,
// 256-1024 work-items, 4 or 8 loads in flight, normal or non-temporal) and
// one contiguous range per block stayed at 4.0-5.8 TB/s.
__global__ __launch_bounds__(256) void copy_kernel(const u32x4* __restrict src, u32x4* __restrict dst, uint64_t n16) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Read-only stream: sum of all 32-bit words mod 2^64 (the achievable read
// bandwidth that decode kernels are compared with).  1024 work-items per
// block, two blocks per CU, eight independent non-temporal 16-byte loads in
// flight per work-item: the best of a sweep over block size (256-1024), grid
// (1-32 blocks per CU), loads in flight (4-16) and cache policy on MI355X
// (6.6 TB/s; normal loads peak near 6.3 TB/s).  Block reduction, one atomic
// per block.
constexpr int kReadBlock = 1024, kReadInflight = 8;

__global__ __launch_bounds__(kReadBlock) void read_kernel(const u32x4* __restrict src, uint64_t n16,
                                                          unsigned long long* __restrict sum) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t acc = 0;
    for (; i + (kReadInflight - 1) * stride < n16; i += kReadInflight * stride) {
        u32x4 v[kReadInflight];
#pragma unroll
        for (int k = 0; k < kReadInflight; ++k) v[k] = __builtin_nontemporal_load(src + i + k * stride);
#pragma unroll
        for (int k = 0; k < kReadInflight; ++k) acc += (uint64_t)v[k].x + v[k].y + v[k].z + v[k].w;
    }
    for (; i < n16; i += stride) {
        const u32x4 a = __builtin_nontemporal_load(src + i);
        acc += (uint64_t)a.x + a.y + a.z + a.w;
    }
    __shared__ uint64_t part[kReadBlock / 64];
    acc = dev::wave_sum(acc);
    if (dev::is_wave_leader()) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < kReadBlock / 64; ++w) t += part[w];
        atomicAdd(sum, (unsigned long long)t);
    }
}

}  // namespace fls

extern "C" {

int fls_copy_bytes(const void* d_src, void* d_dst, uint64_t n_bytes, void* stream) {
    if (n_bytes == 0) return FLS_OK;
    if (!d_src || !d_dst) return FLS_ERR_NULL;
    if ((n_bytes & 15u) || (((uintptr_t)d_src | (uintptr_t)d_dst) & 15u)) return FLS_ERR_ARG;
    const uint64_t n16 = n_bytes / 16u;
    const uint64_t blocks = fls::ceil_div(n16, 256);
    if (blocks > 0x7fffffffu) return FLS_ERR_SIZE;
    hipLaunchKernelGGL(fls::copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, fls::as_stream(stream),
                       static_cast<const fls::u32x4*>(d_src), static_cast<fls::u32x4*>(d_dst), n16);
    return fls::check_launch();
}

int fls_read_bytes(const void* d_src, uint64_t n_bytes, uint64_t* d_sum, void* stream) {
    if (!d_sum) return FLS_ERR_NULL;
    if (const int rc = fls::zero_u64(d_sum, 1, fls::as_stream(stream))) return rc;
    if (n_bytes == 0) return FLS_OK;
    if (!d_src) return FLS_ERR_NULL;
    if ((n_bytes & 15u) || ((uintptr_t)d_src & 15u)) return FLS_ERR_ARG;
    const uint64_t n16 = n_bytes / 16u;
    uint32_t blocks = 0;
    if (const int rc = fls::persistent_grid(fls::ceil_div(n16, fls::kReadBlock), 2, &blocks)) return rc;
    hipLaunchKernelGGL(fls::read_kernel, dim3(blocks), dim3(fls::kReadBlock), 0, fls::as_stream(stream),
                       static_cast<const fls::u32x4*>(d_src), n16, reinterpret_cast<unsigned long long*>(d_sum));
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
        case FLS_ERR_FORMAT: return "serialized column is malformed (magic, version, sizes, checksum or vector metadata)";
        case FLS_ERR_SCRATCH: return "the work exceeds the library's static scratch array (allocation-free entry point)";
        default: return "unknown status";
    }
}

int fls_last_hip_error(void) { return fls::g_last_hip_error; }

const char* fls_version(void) { return "fls-mi355x 0.1"; }

}  // extern "C"
