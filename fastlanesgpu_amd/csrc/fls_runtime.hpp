// Shared host-side plumbing of the C ABI: status codes, HIP error capture,
// launch helpers.  Internal to libfls.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fls/fls.h"

namespace fls {

// Remember the HIP error (thread-local) and map it to FLS_ERR_HIP.
int record_hip_error(hipError_t e);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Check a launch: picks up both launch-configuration errors and earlier
// asynchronous faults reported at launch time.
inline int check_launch() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FLS_OK : record_hip_error(e);
}

__host__ __device__ inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// Zero n 64-bit words on the stream with a kernel (not hipMemsetAsync, so the
// call is safe inside captured graphs).
int zero_u64(uint64_t* d, uint64_t n, hipStream_t stream);

}  // namespace fls

#define FLS_RETURN_IF(cond, code) \
    do {                          \
        if (cond) return (code);  \
    } while (0)

#define FLS_HIP_CHECK(expr)                                      \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return ::fls::record_hip_error(_e); \
    } while (0)
