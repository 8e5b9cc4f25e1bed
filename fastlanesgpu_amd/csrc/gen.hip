// Synthetic workload generators, computed on the GPU straight into HBM.
//
// Every value is a pure function of (seed, stream, row): splitmix64 mixing
// (Steele, Lea, Flood, OOPSLA 2014) of a per-(seed, stream) key plus the row
// counter, so any row range can be produced independently and every consumer
// (CPU checker included) agrees on the data without moving it.
#include "fls_runtime.hpp"

namespace fls {
namespace {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream * 0x9E3779B97F4A7C15ull));
}

// uniform in [0, range), range <= 2^32
__device__ __forceinline__ uint32_t uniform(uint64_t key, uint64_t i, uint64_t range) {
    const uint64_t hi = mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull) >> 32;
    return (uint32_t)((hi * range) >> 32);
}

// Column streams (shared with every checker of this data).
enum : uint64_t {
    ST_C1 = 1,
    ST_QTY = 11, ST_DISC = 12, ST_PRICE = 13,
    ST_DFK = 21, ST_SFK = 22, ST_PFK = 23, ST_CFK = 24, ST_REV = 25,
    ST_SREG = 31, ST_SNAT = 32, ST_PCAT = 33, ST_PBRAND = 34, ST_CREG = 35, ST_CNAT = 36,
    ST_SCOST = 26, ST_PMFGR = 37
};

constexpr uint32_t kDays = 2557;  // 1992-01-01 .. 1998-12-31

// days since 1992-01-01 -> yyyymmdd (proleptic Gregorian)
__host__ __device__ inline uint32_t yyyymmdd(uint32_t day) {
    uint32_t y = 1992;
    for (;;) {
        const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
        const uint32_t len = leap ? 366u : 365u;
        if (day < len) break;
        day -= len;
        ++y;
    }
    const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
    const uint8_t md[12] = {31, (uint8_t)(leap ? 29 : 28), 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    uint32_t m = 0;
    while (day >= md[m]) {
        day -= md[m];
        ++m;
    }
    return y * 10000u + (m + 1u) * 100u + day + 1u;
}

__global__ void gen_c1_kernel(uint64_t key, uint64_t n, uint64_t range, uint32_t* __restrict out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = uniform(key, i, range);
}

__global__ void gen_c2_kernel(uint64_t seed, uint64_t n, uint32_t* __restrict od, uint32_t* __restrict q,
                              uint32_t* __restrict d, uint32_t* __restrict p) {
    const uint64_t kq = stream_key(seed, ST_QTY), kd = stream_key(seed, ST_DISC), kp = stream_key(seed, ST_PRICE);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        od[i] = yyyymmdd((uint32_t)((i * kDays) / n));  // sorted, uniform over the days
        q[i] = 1u + uniform(kq, i, 50);
        d[i] = uniform(kd, i, 11);
        p[i] = 90000u + uniform(kp, i, 10000000u - 90000u);
    }
}

__global__ void gen_c3_fact_kernel(uint64_t seed, uint64_t n, uint32_t* __restrict dfk, uint32_t* __restrict sfk,
                                   uint32_t* __restrict pfk, uint32_t* __restrict cfk, uint32_t* __restrict rev) {
    const uint64_t k1 = stream_key(seed, ST_DFK), k2 = stream_key(seed, ST_SFK), k3 = stream_key(seed, ST_PFK),
                   k4 = stream_key(seed, ST_CFK), k5 = stream_key(seed, ST_REV);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        dfk[i] = uniform(k1, i, 2556);
        sfk[i] = uniform(k2, i, 40000);
        pfk[i] = uniform(k3, i, 800000);
        if (cfk) cfk[i] = uniform(k4, i, 600000);
        rev[i] = uniform(k5, i, 10000000);
    }
}

// SSB's hierarchies.  Nation U{1..25}, its region (nation - 1) / 5 + 1.
__global__ void gen_geo_kernel(uint64_t k_nation, uint32_t n, uint32_t* __restrict region,
                               uint32_t* __restrict nation) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t na = 1u + uniform(k_nation, i, 25u);
        nation[i] = na;
        region[i] = (na - 1u) / 5u + 1u;
    }
}

// Category U{1..25}, brand 40 (category - 1) + U{1..40}.
__global__ void gen_part_kernel(uint64_t k_cat, uint64_t k_brand, uint32_t n, uint32_t* __restrict category,
                                uint32_t* __restrict brand) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = 1u + uniform(k_cat, i, 25u);
        category[i] = c;
        brand[i] = (c - 1u) * 40u + 1u + uniform(k_brand, i, 40u);
    }
}

__global__ void gen_year_kernel(uint32_t n, uint32_t* __restrict year) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        year[i] = yyyymmdd(i) / 10000u;
}

// the columns the Q3.1 / Q4.1 shapes add: lo_supplycost U[60000, 120000),
// p_mfgr 1 + U{0..4}
__global__ void gen_star_extra_kernel(uint64_t seed, uint64_t n, uint32_t* __restrict cost, uint32_t n_part,
                                      uint32_t* __restrict mfgr) {
    const uint64_t kc = stream_key(seed, ST_SCOST), km = stream_key(seed, ST_PMFGR);
    const uint64_t m = n > n_part ? n : n_part;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < n) cost[i] = 60000u + uniform(kc, i, 60000);
        if (i < n_part) mfgr[i] = 1u + uniform(km, i, 5);
    }
}

uint32_t grid_for(uint64_t n) {
    const uint64_t g = ceil_div(n, 256);
    return (uint32_t)(g < 65536 ? (g ? g : 1) : 65536);
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_gen_c1(uint64_t seed, uint64_t n, uint32_t bw, uint32_t* d_out, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(bw > 32, FLS_ERR_BW);
    if (!n) return FLS_OK;
    FLS_RETURN_IF(!d_out, FLS_ERR_NULL);
    hipLaunchKernelGGL(gen_c1_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), stream_key(seed, ST_C1), n,
                       (uint64_t)1 << bw, d_out);
    return check_launch();
}

int fls_gen_c2(uint64_t seed, uint64_t n, uint32_t* d_orderdate, uint32_t* d_quantity, uint32_t* d_discount,
               uint32_t* d_price, void* stream) {
    using namespace fls;
    if (!n) return FLS_OK;
    FLS_RETURN_IF(!d_orderdate || !d_quantity || !d_discount || !d_price, FLS_ERR_NULL);
    hipLaunchKernelGGL(gen_c2_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), seed, n, d_orderdate,
                       d_quantity, d_discount, d_price);
    return check_launch();
}

int fls_gen_c3_fact(uint64_t seed, uint64_t n, uint32_t* d_datefk, uint32_t* d_suppfk, uint32_t* d_partfk,
                    uint32_t* d_custfk, uint32_t* d_revenue, void* stream) {
    using namespace fls;
    if (!n) return FLS_OK;
    FLS_RETURN_IF(!d_datefk || !d_suppfk || !d_partfk || !d_revenue, FLS_ERR_NULL);
    hipLaunchKernelGGL(gen_c3_fact_kernel, dim3(grid_for(n)), dim3(256), 0, as_stream(stream), seed, n, d_datefk,
                       d_suppfk, d_partfk, d_custfk, d_revenue);
    return check_launch();
}

int fls_gen_star_extra(uint64_t seed, uint64_t n, uint32_t* d_supplycost, uint32_t n_part, uint32_t* d_p_mfgr,
                       void* stream) {
    using namespace fls;
    if (!n && !n_part) return FLS_OK;
    FLS_RETURN_IF((n && !d_supplycost) || (n_part && !d_p_mfgr), FLS_ERR_NULL);
    hipLaunchKernelGGL(gen_star_extra_kernel, dim3(grid_for(n > n_part ? n : n_part)), dim3(256), 0, as_stream(stream),
                       seed, n, d_supplycost, n_part, d_p_mfgr);
    return check_launch();
}

int fls_gen_c3_dims(uint64_t seed, uint32_t n_date, uint32_t n_supp, uint32_t n_part, uint32_t n_cust,
                    uint32_t* d_year, uint32_t* d_s_region, uint32_t* d_s_nation, uint32_t* d_p_category,
                    uint32_t* d_p_brand, uint32_t* d_c_region, uint32_t* d_c_nation, void* stream) {
    using namespace fls;
    hipStream_t s = as_stream(stream);
    if (n_date) {
        FLS_RETURN_IF(!d_year, FLS_ERR_NULL);
        hipLaunchKernelGGL(gen_year_kernel, dim3(grid_for(n_date)), dim3(256), 0, s, n_date, d_year);
    }
    if (n_supp) {
        FLS_RETURN_IF(!d_s_region || !d_s_nation, FLS_ERR_NULL);
        hipLaunchKernelGGL(gen_geo_kernel, dim3(grid_for(n_supp)), dim3(256), 0, s, stream_key(seed, ST_SNAT), n_supp,
                           d_s_region, d_s_nation);
    }
    if (n_part) {
        FLS_RETURN_IF(!d_p_category || !d_p_brand, FLS_ERR_NULL);
        hipLaunchKernelGGL(gen_part_kernel, dim3(grid_for(n_part)), dim3(256), 0, s, stream_key(seed, ST_PCAT),
                           stream_key(seed, ST_PBRAND), n_part, d_p_category, d_p_brand);
    }
    if (n_cust) {
        FLS_RETURN_IF(!d_c_region || !d_c_nation, FLS_ERR_NULL);
        hipLaunchKernelGGL(gen_geo_kernel, dim3(grid_for(n_cust)), dim3(256), 0, s, stream_key(seed, ST_CNAT), n_cust,
                           d_c_region, d_c_nation);
    }
    return check_launch();
}

}  // extern "C"
