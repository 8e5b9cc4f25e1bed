// Kernels written the way the reference writes its query and decode kernels
// (fastlanes/src/ssb/fls_q11.cu:29-135, fls_q21.cu:21-263, fls_q31.cu:27-94 and
// 96-160, fastlanes/src/delta_shared_memory.cu:24-60, delta_global_memory.cu:
// 25-40): the same primitive calls with the same explicit template arguments,
// the same tile arithmetic (one tile per workgroup of BLOCK_THREADS
// work-items), compiled against include/fls/device/*.hpp alone.  They exist
// to show that the device headers are a drop-in for the reference's and to
// instantiate every primitive under test; the library's own kernels are the
// fast path.  Each launcher returns 0 or the HIP error code.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fls/device/crystal.hpp"
#include "fls/device/unpack.hpp"

#define FLS_EXPORT extern "C" __attribute__((visibility("default")))

static int launch_status() { return (int)hipGetLastError(); }

// Tile bookkeeping shared by every query kernel below (fls_q11.cu:43-47).
template <int TILE>
__device__ __forceinline__ int tile_items(int n) {
    const int num_tiles = (n + TILE - 1) / TILE;
    return (int)blockIdx.x == num_tiles - 1 ? n - (int)blockIdx.x * TILE : TILE;
}

// ------------------------------------------------------------- Q1.1 shapes
struct Q11Meta {
    int n_rows;
    int orderdate_bw, quantity_bw, discount_bw, price_bw;
    int orderdate_min;  // frame of reference of the packed orderdate
};

// QueryKernel_v1 (fls_q11.cu:29-81): four packed columns, crystal-opt forms.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void q11_v1(const int* enc_orderdate, const int* enc_discount, const int* enc_quantity,
                       const int* enc_price, Q11Meta m, unsigned long long* revenue) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int items2[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    long long sum = 0;
    const int num_tile_items = tile_items<TILE>(m.n_rows);

    unpack_device(enc_orderdate + blockIdx.x * m.orderdate_bw * 32, items, m.orderdate_bw);
    BlockPredGT<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 19930000 - m.orderdate_min, selection_flags,
                                                           num_tile_items);
    BlockPredAndLT<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 19940000 - m.orderdate_min, selection_flags,
                                                              num_tile_items);
    unpack_device(enc_quantity + blockIdx.x * m.quantity_bw * 32, items, m.quantity_bw);
    BlockPredAndLT<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 25, selection_flags, num_tile_items);
    unpack_device(enc_discount + blockIdx.x * m.discount_bw * 32, items, m.discount_bw);
    BlockPredAndGTE<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 1, selection_flags, num_tile_items);
    BlockPredAndLTE<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 3, selection_flags, num_tile_items);
    unpack_device(enc_price + blockIdx.x * m.price_bw * 32, items2, m.price_bw);
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if ((int)threadIdx.x + BLOCK_THREADS * i < num_tile_items && selection_flags[i])
            sum += (long long)items[i] * items2[i];
    __syncthreads();
    static __shared__ long long buffer[32];
    const unsigned long long aggregate = BlockSum<long long, BLOCK_THREADS, ITEMS_PER_THREAD>(sum, buffer);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(revenue, aggregate);
}

// QueryKernel_v2 (fls_q11.cu:83-135): the price column is plain and read
// with BlockPredLoad<T, ST, ...> for the rows that passed.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void q11_v2(const int* enc_orderdate, const int* enc_discount, const int* enc_quantity, const int* price,
                       Q11Meta m, unsigned long long* revenue) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int items2[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    long long sum = 0;
    const int tile_offset = blockIdx.x * TILE;
    const int num_tile_items = tile_items<TILE>(m.n_rows);

    unpack_device(enc_orderdate + blockIdx.x * m.orderdate_bw * 32, items, m.orderdate_bw);
    BlockPredGT<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 19930000 - m.orderdate_min, selection_flags,
                                                           num_tile_items);
    BlockPredAndLT<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 19940000 - m.orderdate_min, selection_flags,
                                                              num_tile_items);
    unpack_device(enc_quantity + blockIdx.x * m.quantity_bw * 32, items, m.quantity_bw);
    BlockPredAndLT<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 25, selection_flags, num_tile_items);
    unpack_device(enc_discount + blockIdx.x * m.discount_bw * 32, items, m.discount_bw);
    BlockPredAndGTE<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 1, selection_flags, num_tile_items);
    BlockPredAndLTE<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, 3, selection_flags, num_tile_items);
    BlockPredLoad<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(price + tile_offset, items2, num_tile_items,
                                                             selection_flags);
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if ((int)threadIdx.x + BLOCK_THREADS * i < num_tile_items && selection_flags[i])
            sum += (long long)items[i] * items2[i];
    __syncthreads();
    static __shared__ long long buffer[32];
    const unsigned long long aggregate = BlockSum<long long, BLOCK_THREADS, ITEMS_PER_THREAD>(sum, buffer);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(revenue, aggregate);
}

FLS_EXPORT int rs_q11(int version, const int* od, const int* disc, const int* qty, const int* price, int n_rows,
                      int od_bw, int qty_bw, int disc_bw, int price_bw, int od_min, unsigned long long* d_revenue) {
    const Q11Meta m{n_rows, od_bw, qty_bw, disc_bw, price_bw, od_min};
    const int tiles = (n_rows + 1023) / 1024;
    if (version == 1) q11_v1<32, 32><<<tiles, 32>>>(od, disc, qty, price, m, d_revenue);
    else q11_v2<32, 32><<<tiles, 32>>>(od, disc, qty, price, m, d_revenue);
    return launch_status();
}

// ------------------------------------------------------------- Q2.1 shapes
// Dimension builds (fls_q21.cu:197-263): crystal/ three-parameter forms.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void build_hashtable_s(const int* filter_col, const int* dim_key, int num_tuples, int* hash_table,
                                  int num_slots, int filter_eq) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    const int tile_offset = blockIdx.x * TILE;
    const int num_tile_items = tile_items<TILE>(num_tuples);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(filter_col + tile_offset, items, num_tile_items);
    BlockPredEQ<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, filter_eq, selection_flags, num_tile_items);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(dim_key + tile_offset, items, num_tile_items);
    BlockBuildSelectivePHT_1<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, selection_flags, hash_table, num_slots,
                                                                   num_tile_items);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void build_hashtable_p(const int* filter_col, const int* dim_key, const int* dim_val, int num_tuples,
                                  int* hash_table, int num_slots, int filter_eq) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int items2[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    const int tile_offset = blockIdx.x * TILE;
    const int num_tile_items = tile_items<TILE>(num_tuples);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(filter_col + tile_offset, items, num_tile_items);
    BlockPredEQ<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, filter_eq, selection_flags, num_tile_items);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(dim_key + tile_offset, items, num_tile_items);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(dim_val + tile_offset, items2, num_tile_items);
    BlockBuildSelectivePHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, items2, selection_flags, hash_table,
                                                                        num_slots, num_tile_items);
}

// Date build with a key minimum; the year range is a conjunction (Q3.1's
// build_hashtable_d shape, fls_q31.cu:349-372).
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void build_hashtable_d(const int* dim_key, const int* dim_val, int num_tuples, int* hash_table,
                                  int num_slots, int val_min, int year_lo, int year_hi) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int items2[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    const int tile_offset = blockIdx.x * TILE;
    const int num_tile_items = tile_items<TILE>(num_tuples);
    InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(selection_flags);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(dim_val + tile_offset, items2, num_tile_items);
    BlockPredGTE<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items2, year_lo, selection_flags, num_tile_items);
    BlockPredAndLTE<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items2, year_hi, selection_flags, num_tile_items);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(dim_key + tile_offset, items, num_tile_items);
    BlockBuildSelectivePHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, items2, selection_flags, hash_table,
                                                                        num_slots, val_min, num_tile_items);
}

struct Q21Args {
    const int *lo_orderdate, *lo_partkey, *lo_suppkey, *lo_revenue;
    int lo_len;
    const int* ht_s;
    int s_len;
    const int* ht_p;
    int p_len;
    const int* ht_d;
    int d_len, d_min;
    int suppkey_bw, partkey_bw, orderdate_bw, revenue_bw;
    int* res;
};

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__device__ __forceinline__ void q21_group(const Q21Args& a, const int (&flags)[ITEMS_PER_THREAD],
                                          const int (&brand)[ITEMS_PER_THREAD], const int (&year)[ITEMS_PER_THREAD],
                                          const int (&revenue)[ITEMS_PER_THREAD], int num_tile_items) {
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) {
        if ((int)threadIdx.x + BLOCK_THREADS * i < num_tile_items && flags[i]) {
            const int hash = (brand[i] * 7 + (year[i] - 1992)) % ((1998 - 1992 + 1) * (5 * 5 * 40));
            a.res[hash * 4] = year[i];
            a.res[hash * 4 + 1] = brand[i];
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.res[hash * 4 + 2]), (long long)revenue[i]);
        }
    }
}

// probe_v1 (fls_q21.cu:21-78): four packed columns, whole-vector decode.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void q21_probe_v1(Q21Args a) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    int brand[ITEMS_PER_THREAD];
    int year[ITEMS_PER_THREAD];
    int revenue[ITEMS_PER_THREAD];
    const int num_tile_items = tile_items<TILE>(a.lo_len);
    InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(selection_flags);
    unpack_device(a.lo_suppkey + blockIdx.x * a.suppkey_bw * 32, items, a.suppkey_bw);
    BlockProbeAndPHT_1<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, selection_flags, a.ht_s, a.s_len, num_tile_items);
    unpack_device(a.lo_partkey + blockIdx.x * a.partkey_bw * 32, items, a.partkey_bw);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, brand, selection_flags, a.ht_p, a.p_len,
                                                                  num_tile_items);
    unpack_device(a.lo_orderdate + blockIdx.x * a.orderdate_bw * 32, items, a.orderdate_bw);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, year, selection_flags, a.ht_d, a.d_len,
                                                                  a.d_min, num_tile_items);
    unpack_device(a.lo_revenue + blockIdx.x * a.revenue_bw * 32, revenue, a.revenue_bw);
    q21_group<BLOCK_THREADS, ITEMS_PER_THREAD>(a, selection_flags, brand, year, revenue, num_tile_items);
}

// probe_v2 / probe_v3 (fls_q21.cu:80-195): quarter-vector decode of the keys
// (tile = 256 rows, quarter q of vector v at word 8*bw*(4v+q)), the plain
// revenue column read whole (v2) or for the surviving rows only (v3).
template <int BLOCK_THREADS, int ITEMS_PER_THREAD, bool PRED_LOAD>
__global__ void q21_probe_quarter(Q21Args a) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    int brand[ITEMS_PER_THREAD];
    int year[ITEMS_PER_THREAD];
    int revenue[ITEMS_PER_THREAD];
    const int tile_offset = blockIdx.x * TILE;
    const int num_tile_items = tile_items<TILE>(a.lo_len);
    InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(selection_flags);
    unpack_8_at_a_time::unpack_device(a.lo_suppkey + blockIdx.x * a.suppkey_bw * 8, items, a.suppkey_bw);
    BlockProbeAndPHT_1<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, selection_flags, a.ht_s, a.s_len, num_tile_items);
    unpack_8_at_a_time::unpack_device(a.lo_partkey + blockIdx.x * a.partkey_bw * 8, items, a.partkey_bw);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, brand, selection_flags, a.ht_p, a.p_len,
                                                                  num_tile_items);
    unpack_8_at_a_time::unpack_device(a.lo_orderdate + blockIdx.x * a.orderdate_bw * 8, items, a.orderdate_bw);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, year, selection_flags, a.ht_d, a.d_len,
                                                                  a.d_min, num_tile_items);
    if constexpr (PRED_LOAD)
        BlockPredLoad<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(a.lo_revenue + tile_offset, revenue, num_tile_items,
                                                                 selection_flags);
    else
        BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(a.lo_revenue + tile_offset, revenue, num_tile_items);
    q21_group<BLOCK_THREADS, ITEMS_PER_THREAD>(a, selection_flags, brand, year, revenue, num_tile_items);
}

// Builds all three tables, then probes.  Tables and res must be zero.
FLS_EXPORT int rs_q21(int version, const int* lo_orderdate, const int* lo_partkey, const int* lo_suppkey,
                      const int* lo_revenue, int lo_len, int suppkey_bw, int partkey_bw, int orderdate_bw,
                      int revenue_bw, const int* s_region, const int* s_suppkey, int s_len, int region,
                      const int* p_category, const int* p_partkey, const int* p_brand, int p_len, int category,
                      const int* d_datekey, const int* d_year, int d_len, int d_min, int* ht_s, int* ht_p, int* ht_d,
                      int* res) {
    constexpr int BT = 32;
    const int ipt = version == 1 ? 32 : 8;
    const int tile = BT * ipt;
    auto grid = [&](int n) { return (n + tile - 1) / tile; };
    if (version == 1) {
        build_hashtable_s<32, 32><<<grid(s_len), BT>>>(s_region, s_suppkey, s_len, ht_s, s_len, region);
        build_hashtable_p<32, 32><<<grid(p_len), BT>>>(p_category, p_partkey, p_brand, p_len, ht_p, p_len, category);
        build_hashtable_d<32, 32><<<grid(d_len), BT>>>(d_datekey, d_year, d_len, ht_d, d_len, d_min, 0, 1 << 30);
    } else {
        build_hashtable_s<32, 8><<<grid(s_len), BT>>>(s_region, s_suppkey, s_len, ht_s, s_len, region);
        build_hashtable_p<32, 8><<<grid(p_len), BT>>>(p_category, p_partkey, p_brand, p_len, ht_p, p_len, category);
        build_hashtable_d<32, 8><<<grid(d_len), BT>>>(d_datekey, d_year, d_len, ht_d, d_len, d_min, 0, 1 << 30);
    }
    const Q21Args a{lo_orderdate, lo_partkey, lo_suppkey, lo_revenue, lo_len, ht_s, s_len, ht_p, p_len,
                    ht_d, d_len, d_min, suppkey_bw, partkey_bw, orderdate_bw, revenue_bw, res};
    if (version == 1) q21_probe_v1<32, 32><<<grid(lo_len), BT>>>(a);
    else if (version == 2) q21_probe_quarter<32, 8, false><<<grid(lo_len), BT>>>(a);
    else q21_probe_quarter<32, 8, true><<<grid(lo_len), BT>>>(a);
    return launch_status();
}

// ------------------------------------------------------------- Q3.1 shapes
struct Q31Args {
    const int *lo_orderdate, *lo_custkey, *lo_suppkey, *lo_revenue;
    int lo_len;
    const int* ht_s;
    int s_len;
    const int* ht_c;
    int c_len;
    const int* ht_d;
    int d_len, d_min;
    int suppkey_bw, custkey_bw, orderdate_bw, revenue_bw;
    int* res;
};

__device__ __forceinline__ void q31_add(const Q31Args& a, int s_nation, int c_nation, int year, int revenue) {
    const int hash = (s_nation * 25 * 7 + c_nation * 7 + (year - 1992)) % ((1998 - 1992 + 1) * 25 * 25);
    a.res[hash * 6] = year;
    a.res[hash * 6 + 1] = c_nation;
    a.res[hash * 6 + 2] = s_nation;
    atomicAdd(reinterpret_cast<unsigned long long*>(&a.res[hash * 6 + 4]), (long long)revenue);
}

// probe_v1 (fls_q31.cu:27-94): each key column decoded into LDS in the
// "normal" order, loaded back striped, probed into registers.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void q31_probe_v1(Q31Args a) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    int c_nation[ITEMS_PER_THREAD];
    int s_nation[ITEMS_PER_THREAD];
    int year[ITEMS_PER_THREAD];
    int revenue[ITEMS_PER_THREAD];
    static __shared__ int unpacked[1024];
    const int num_tile_items = tile_items<TILE>(a.lo_len);
    InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(selection_flags);
    unpack_device(a.lo_suppkey + blockIdx.x * a.suppkey_bw * 32, unpacked, a.suppkey_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(unpacked, items, num_tile_items);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, s_nation, selection_flags, a.ht_s, a.s_len,
                                                                  num_tile_items);
    unpack_device(a.lo_custkey + blockIdx.x * a.custkey_bw * 32, unpacked, a.custkey_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(unpacked, items, num_tile_items);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, c_nation, selection_flags, a.ht_c, a.c_len,
                                                                  num_tile_items);
    unpack_device(a.lo_orderdate + blockIdx.x * a.orderdate_bw * 32, unpacked, a.orderdate_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(unpacked, items, num_tile_items);
    BlockProbeAndPHT_2<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, year, selection_flags, a.ht_d, a.d_len,
                                                                  a.d_min, num_tile_items);
    unpack_device(a.lo_revenue + blockIdx.x * a.revenue_bw * 32, unpacked, a.revenue_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(unpacked, revenue, num_tile_items);
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if ((int)threadIdx.x + BLOCK_THREADS * i < num_tile_items && selection_flags[i])
            q31_add(a, s_nation[i], c_nation[i], year[i], revenue[i]);
}

// probe_v2 (fls_q31.cu:96-160): probe results land in tile-sized LDS arrays
// through BlockProbeAndPHT_2_R_S.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void q31_probe_v2(Q31Args a) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    int items[ITEMS_PER_THREAD];
    int selection_flags[ITEMS_PER_THREAD];
    static __shared__ int shared_c_nation[TILE];
    static __shared__ int shared_s_nation[TILE];
    static __shared__ int shared_year[TILE];
    static __shared__ int shared_revenue[TILE];
    const int num_tile_items = tile_items<TILE>(a.lo_len);
    InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(selection_flags);
    unpack_device(a.lo_suppkey + blockIdx.x * a.suppkey_bw * 32, shared_revenue, a.suppkey_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(shared_revenue, items, num_tile_items);
    BlockProbeAndPHT_2_R_S<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, shared_s_nation, selection_flags,
                                                                      a.ht_s, a.s_len, num_tile_items);
    unpack_device(a.lo_custkey + blockIdx.x * a.custkey_bw * 32, shared_revenue, a.custkey_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(shared_revenue, items, num_tile_items);
    BlockProbeAndPHT_2_R_S<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, shared_c_nation, selection_flags,
                                                                      a.ht_c, a.c_len, num_tile_items);
    unpack_device(a.lo_orderdate + blockIdx.x * a.orderdate_bw * 32, shared_revenue, a.orderdate_bw);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(shared_revenue, items, num_tile_items);
    BlockProbeAndPHT_2_R_S<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, shared_year, selection_flags, a.ht_d,
                                                                      a.d_len, a.d_min, num_tile_items);
    unpack_device(a.lo_revenue + blockIdx.x * a.revenue_bw * 32, shared_revenue, a.revenue_bw);
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) {
        const int s = BLOCK_THREADS * i + (int)threadIdx.x;
        if (s < num_tile_items && selection_flags[i])
            q31_add(a, shared_s_nation[s], shared_c_nation[s], shared_year[s], shared_revenue[s]);
    }
}

FLS_EXPORT int rs_q31(int version, const int* lo_orderdate, const int* lo_custkey, const int* lo_suppkey,
                      const int* lo_revenue, int lo_len, int suppkey_bw, int custkey_bw, int orderdate_bw,
                      int revenue_bw, const int* s_region, const int* s_suppkey, const int* s_nation, int s_len,
                      int region_s, const int* c_region, const int* c_custkey, const int* c_nation, int c_len,
                      int region_c, const int* d_datekey, const int* d_year, int d_len, int d_min, int year_lo,
                      int year_hi, int* ht_s, int* ht_c, int* ht_d, int* res) {
    auto grid = [](int n) { return (n + 1023) / 1024; };
    build_hashtable_p<32, 32><<<grid(s_len), 32>>>(s_region, s_suppkey, s_nation, s_len, ht_s, s_len, region_s);
    build_hashtable_p<32, 32><<<grid(c_len), 32>>>(c_region, c_custkey, c_nation, c_len, ht_c, c_len, region_c);
    build_hashtable_d<32, 32><<<grid(d_len), 32>>>(d_datekey, d_year, d_len, ht_d, d_len, d_min, year_lo, year_hi);
    const Q31Args a{lo_orderdate, lo_custkey, lo_suppkey, lo_revenue, lo_len, ht_s, s_len, ht_c, c_len,
                    ht_d, d_len, d_min, suppkey_bw, custkey_bw, orderdate_bw, revenue_bw, res};
    if (version == 1) q31_probe_v1<32, 32><<<grid(lo_len), 32>>>(a);
    else q31_probe_v2<32, 32><<<grid(lo_len), 32>>>(a);
    return launch_status();
}

// ------------------------------------------------------ delta (d_rsum_32)
// delta_shared_memory.cu:24-60: decode into LDS, running sums into LDS, then
// BlockLoad + BlockSum; delta_global_memory.cu:25-40: running sums straight
// to the output column (transposed order).
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void delta_shared(const uint32_t* base_col, const uint32_t* packed, int bw, int n_tup,
                             unsigned long long* total) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    uint32_t items[ITEMS_PER_THREAD];
    static __shared__ uint32_t unpacked[1024];
    static __shared__ uint32_t rsumed[1024];
    long long sum = 0;
    const int num_tile_items = tile_items<TILE>(n_tup);
    unpack_device(packed + blockIdx.x * bw * 32, unpacked, bw);
    d_rsum_32(unpacked, rsumed, base_col + blockIdx.x * 32);
    BlockLoad<uint32_t, BLOCK_THREADS, ITEMS_PER_THREAD>(rsumed, items, num_tile_items);
#pragma unroll
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if ((int)threadIdx.x + BLOCK_THREADS * i < num_tile_items) sum += items[i];
    __syncthreads();
    static __shared__ long long buffer[32];
    const unsigned long long aggregate = BlockSum<long long, BLOCK_THREADS, ITEMS_PER_THREAD>(sum, buffer);
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, aggregate);
}

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void delta_global(const uint32_t* base_col, const uint32_t* packed, int bw, uint32_t* out) {
    static __shared__ uint32_t unpacked[1024];
    unpack_device(packed + blockIdx.x * bw * 32, unpacked, bw);
    d_rsum_32(unpacked, out + blockIdx.x * BLOCK_THREADS * ITEMS_PER_THREAD, base_col + blockIdx.x * 32);
}

FLS_EXPORT int rs_delta(const uint32_t* bases, const uint32_t* packed, int bw, int n_vec, int n_tup, uint32_t* out,
                        unsigned long long* total) {
    delta_global<32, 32><<<n_vec, 32>>>(bases, packed, bw, out);
    delta_shared<32, 32><<<n_vec, 32>>>(bases, packed, bw, n_tup, total);
    return launch_status();
}

// ------------------------------------------------------ decode entry points
// Every per-width name of the three reference namespaces, called by name.
#define RS_FOR_EACH_BW(X)                                                                                        \
    X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) \
        X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28) X(29) X(30) X(31) X(32)

// mode 0: fused::unpack_<bw>... into registers, stored with load_registers;
// mode 1: normal::unpack_<bw>... straight into the output;
// mode 2: fused::unpack(in, out, bw) and normal::unpack(in, out, bw), XORed
//         (the two must agree, so the result is 0 when they do).
__global__ void named_widths(const uint32_t* in, uint32_t* out, int bw, int mode) {
    const uint32_t* vin = in + blockIdx.x * bw * 32;
    uint32_t* vout = out + blockIdx.x * 1024;
    uint32_t regs[32];
    if (mode == 0) {
        switch (bw) {
#define RS_FUSED(B) \
    case B: generated::unpack::cuda::fused::unpack_##B##bw_32ow_32crw_1uf(vin, regs); break;
            RS_FOR_EACH_BW(RS_FUSED)
#undef RS_FUSED
        }
        load_registers<uint32_t, 32, 32>((int)threadIdx.x, vout, regs);
    } else if (mode == 1) {
        switch (bw) {
#define RS_NORMAL(B) \
    case B: generated::unpack::cuda::normal::unpack_##B##bw_32ow_32crw_1uf(vin, vout); break;
            RS_FOR_EACH_BW(RS_NORMAL)
#undef RS_NORMAL
        }
    } else {
        generated::unpack::cuda::normal::unpack(vin, vout, (uint8_t)bw);
        generated::unpack::cuda::fused::unpack(vin, regs, (uint8_t)bw);
#pragma unroll
        for (int r = 0; r < 32; ++r) vout[32 * r + threadIdx.x] ^= regs[r];
    }
}

// Quarter q of every vector through unpack_8_at_a_time's per-width names
// (mode 0) or its unpack_device (mode 1), rows written to out[256*t + 32*r + lane]
// for tile t = 4v + q.
__global__ void named_quarters(const uint32_t* in, uint32_t* out, int bw, int mode) {
    const uint32_t* qin = in + blockIdx.x * bw * 8;
    uint32_t rows[8];
    if (mode == 0) {
        switch (bw) {
#define RS_Q(B) \
    case B: unpack_8_at_a_time::unpack_##B##bw_32ow_32crw_1uf(qin, rows); break;
            RS_Q(0) RS_Q(4) RS_Q(8) RS_Q(12) RS_Q(16) RS_Q(20) RS_Q(24) RS_Q(28) RS_Q(32)
#undef RS_Q
            default:
                for (int r = 0; r < 8; ++r) rows[r] = 0xdeadbeefu;
        }
    } else {
        unpack_8_at_a_time::unpack_device(reinterpret_cast<const int32_t*>(qin), reinterpret_cast<int32_t*>(rows),
                                          (uint8_t)bw);
    }
    load_registers<uint32_t, 32, 8>((int)threadIdx.x, out + blockIdx.x * 256, rows);
}

FLS_EXPORT int rs_unpack_named(const uint32_t* in, uint32_t* out, int bw, int n_vec, int mode) {
    named_widths<<<n_vec, 32>>>(in, out, bw, mode);
    return launch_status();
}

FLS_EXPORT int rs_unpack_quarters(const uint32_t* in, uint32_t* out, int bw, int n_vec, int mode) {
    named_quarters<<<4 * n_vec, 32>>>(in, out, bw, mode);
    return launch_status();
}

// unpack_global as the reference launches it (one 32-wide block per vector)
// and with block_threads work-items (a multiple of 32 dividing 32 * n_vec).
FLS_EXPORT int rs_unpack_global(const uint32_t* in, uint32_t* out, int bw, int n_vec, int block_threads) {
    if (block_threads % 32 || (32 * n_vec) % block_threads) return -1;
    unpack_global<<<32 * n_vec / block_threads, block_threads>>>(in, out, (uint8_t)bw);
    return launch_status();
}

// ------------------------------------------------------ primitive catalogue
// Every named predicate, both template forms, the generic functor forms,
// BlockPredLoad / BlockLoad / BlockStore (+ Direct), IsTerm and both BlockSum
// overloads, on tiles of BLOCK_THREADS x ITEMS_PER_THREAD items.  Output
// block p (p = 0..N_PRED_OUT-1) of out_flags holds the flags of test p in
// row order; out_vals receives BlockPredLoad'ed values (0 where not loaded);
// out_sums[t] the tile's BlockSum of the items and out_term[t] IsTerm of the
// final EQ flags.
constexpr int N_PRED_OUT = 22;

template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void prims(const int* vals, const int* init_flags, int n, int c, int* out_flags, int* out_vals,
                      long long* out_sums, int* out_term) {
    constexpr int TILE = BLOCK_THREADS * ITEMS_PER_THREAD;
    const int off = blockIdx.x * TILE;
    const int num = tile_items<TILE>(n);
    int items[ITEMS_PER_THREAD] = {};
    int f0[ITEMS_PER_THREAD] = {};
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(vals + off, items, num);
    BlockLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(init_flags + off, f0, num);
    int f[ITEMS_PER_THREAD];
    uint16_t g[ITEMS_PER_THREAD];
    int p = 0;
    auto emit = [&](const int (&fl)[ITEMS_PER_THREAD]) {
        BlockStore<int, BLOCK_THREADS, ITEMS_PER_THREAD>(out_flags + (size_t)p * n + off, fl, num);
        ++p;
    };
    auto emit16 = [&](const uint16_t (&gl)[ITEMS_PER_THREAD]) {
        int t[ITEMS_PER_THREAD];
        for (int i = 0; i < ITEMS_PER_THREAD; ++i) t[i] = gl[i];
        emit(t);
    };
#define RS_SET3(NAME)                                                                  \
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) f[i] = f0[i];                           \
    NAME<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, num);                      \
    emit(f);
#define RS_SET4(NAME)                                                                  \
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) g[i] = (uint16_t)f0[i];                 \
    NAME<int, uint16_t, BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, g, num);            \
    emit16(g);
    // 0-4: set forms (crystal), 5-9: set forms (crystal-opt, uint16 flags)
    RS_SET3(BlockPredGT) RS_SET3(BlockPredLT) RS_SET3(BlockPredGTE) RS_SET3(BlockPredLTE) RS_SET3(BlockPredEQ)
    RS_SET4(BlockPredGT) RS_SET4(BlockPredLT) RS_SET4(BlockPredGTE) RS_SET4(BlockPredLTE) RS_SET4(BlockPredEQ)
    // 10-14: And forms (crystal), 15-19 And forms (crystal-opt)
    RS_SET3(BlockPredAndGT) RS_SET3(BlockPredAndLT) RS_SET3(BlockPredAndGTE) RS_SET3(BlockPredAndLTE)
    RS_SET3(BlockPredAndEQ)
    RS_SET4(BlockPredAndGT) RS_SET4(BlockPredAndLT) RS_SET4(BlockPredAndGTE) RS_SET4(BlockPredAndLTE)
    RS_SET4(BlockPredAndEQ)
    // 20: OrEQ (both forms must agree; the crystal-opt result is emitted)
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) f[i] = f0[i];
    BlockPredOrEQ<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, num);
    int h[ITEMS_PER_THREAD];
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) h[i] = f0[i];
    BlockPredOrEQ<int, int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, h, num);
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) h[i] = (h[i] == f[i]) ? h[i] : 7;
    emit(h);
    // 21: generic functor forms composed: (x < c) && (x >= c/2) || x == c,
    // via BlockPred<T, SelectOp>, BlockPredAnd<T, ST, SelectOp>, BlockPredOr
    // and their Direct forms (both must agree; mismatch shows as 7).
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) f[i] = 0;
    BlockPred<int, LessThan<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(items, LessThan<int>(c), f, num);
    BlockPredAnd<int, int, GreaterThanEq<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(items, GreaterThanEq<int>(c / 2), f,
                                                                               num);
    BlockPredOr<int, Eq<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(items, Eq<int>(c), f, num);
    const int tid = (int)threadIdx.x % BLOCK_THREADS;
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) h[i] = 0;
    BlockPredDirect<int, int, LessThan<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, LessThan<int>(c), h, num);
    BlockPredAndDirect<int, GreaterThanEq<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items,
                                                                                 GreaterThanEq<int>(c / 2), h, num);
    BlockPredOrDirect<int, Eq<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(tid, items, Eq<int>(c), h, num);
    int gt[ITEMS_PER_THREAD] = {};
    BlockPred<int, int, GreaterThan<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(items, GreaterThan<int>(c), gt, num);
    BlockPredAnd<int, LessThanEq<int>, BLOCK_THREADS, ITEMS_PER_THREAD>(items, LessThanEq<int>(c), gt, num);
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) f[i] = (h[i] == f[i] && gt[i] == 0) ? f[i] : 7;
    emit(f);
#undef RS_SET3
#undef RS_SET4
    // predicated loads of the items under the initial flags, both forms
    int loaded[ITEMS_PER_THREAD] = {};
    int loaded2[ITEMS_PER_THREAD] = {};
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) g[i] = (uint16_t)f0[i];
    BlockPredLoad<int, BLOCK_THREADS, ITEMS_PER_THREAD>(vals + off, loaded, num, f0);
    BlockPredLoad<int, uint16_t, BLOCK_THREADS, ITEMS_PER_THREAD>(vals + off, loaded2, num, g);
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) loaded[i] = loaded[i] == loaded2[i] ? loaded[i] : -7;
    BlockStore<int, BLOCK_THREADS, ITEMS_PER_THREAD>(out_vals + off, loaded, num);
    // BlockSum of the live items (array overload) and of their per-item sum
    long long wide[ITEMS_PER_THREAD];
    long long mine = 0;
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) {
        wide[i] = tid + i * BLOCK_THREADS < num ? (long long)items[i] : 0;
        mine += wide[i];
    }
    __shared__ long long sh[64];
    const long long s1 = BlockSum<long long, BLOCK_THREADS, ITEMS_PER_THREAD>(wide, sh);
    const long long s2 = BlockSum<long long, BLOCK_THREADS, ITEMS_PER_THREAD>(mine, sh);
    // IsTerm over the EQ flags of the tile: work-item flags OR-ed over the tile
    int eq[ITEMS_PER_THREAD] = {};
    InitFlags<BLOCK_THREADS, ITEMS_PER_THREAD>(eq);
    BlockPredEQ<int, BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, eq, num);
    for (int i = 0; i < ITEMS_PER_THREAD; ++i)
        if (tid + i * BLOCK_THREADS >= num) eq[i] = 0;
    const int none_here = IsTerm<int, BLOCK_THREADS, ITEMS_PER_THREAD>(eq) ? 0 : 1;
    const int some = BlockSum<int, BLOCK_THREADS, ITEMS_PER_THREAD>(none_here, reinterpret_cast<int*>(sh + 32));
    if (tid == 0) {
        out_sums[blockIdx.x] = s1 == s2 ? s1 : -1;
        out_term[blockIdx.x] = some == 0 ? 1 : 0;
    }
}

FLS_EXPORT int rs_prims(int block_threads, int items_per_thread, const int* vals, const int* init_flags, int n, int c,
                        int* out_flags, int* out_vals, long long* out_sums, int* out_term) {
    const int tile = block_threads * items_per_thread;
    const int grid = (n + tile - 1) / tile;
#define RS_PRIMS(BT, IPT)                                                                                          \
    if (block_threads == BT && items_per_thread == IPT) {                                                        \
        prims<BT, IPT><<<grid, BT>>>(vals, init_flags, n, c, out_flags, out_vals, out_sums, out_term);           \
        return launch_status();                                                                                  \
    }
    RS_PRIMS(32, 32) RS_PRIMS(32, 8) RS_PRIMS(64, 16) RS_PRIMS(128, 4) RS_PRIMS(256, 2)
#undef RS_PRIMS
    return -1;
}

FLS_EXPORT int rs_n_pred_out(void) { return N_PRED_OUT; }

// ------------------------------------------------------ packed 16 x 2
// out[k] block (k = 0..5) holds, per word (row order of the words, ITEMS/2
// per work-item), the flag word after: 0 GT_int_16_2, 1 LT_int_16_2, 2 AndLTX,
// 3 AndLTEX, 4 AndGTEX (each And applied to init), 5 BlockPred_int_16_2 with
// GreaterThanEqX and BlockPredDirect_int_16_2 (4-argument) with LessThanEqX
// ANDed together.
template <int BLOCK_THREADS, int ITEMS_PER_THREAD>
__global__ void simd16(const uint32_t* words, const uint32_t* init, int n_items, uint32_t c, uint32_t* out,
                       int n_words_total) {
    constexpr int HALF = ITEMS_PER_THREAD / 2;
    const int tid = (int)threadIdx.x;
    const int base = blockIdx.x * BLOCK_THREADS * HALF;
    uint32_t items[ITEMS_PER_THREAD] = {};
    uint32_t f[ITEMS_PER_THREAD];
    for (int i = 0; i < HALF; ++i) items[i] = words[base + tid + i * BLOCK_THREADS];
    auto put = [&](int k, const uint32_t (&fl)[ITEMS_PER_THREAD]) {
        for (int i = 0; i < HALF; ++i) out[(size_t)k * n_words_total + base + tid + i * BLOCK_THREADS] = fl[i];
    };
    auto reset = [&]() {
        for (int i = 0; i < ITEMS_PER_THREAD; ++i) f[i] = i < HALF ? init[base + tid + i * BLOCK_THREADS] : 0u;
    };
    reset();
    BlockPredGT_int_16_2<BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, n_items);
    put(0, f);
    reset();
    BlockPredLT_int_16_2<BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, n_items);
    put(1, f);
    reset();
    BlockPredAndLTX<BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, n_items);
    put(2, f);
    reset();
    BlockPredAndLTEX<BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, n_items);
    put(3, f);
    reset();
    BlockPredAndGTEX<BLOCK_THREADS, ITEMS_PER_THREAD>(items, c, f, n_items);
    put(4, f);
    reset();
    uint32_t f2[ITEMS_PER_THREAD];
    for (int i = 0; i < ITEMS_PER_THREAD; ++i) f2[i] = f[i];
    BlockPred_int_16_2<GreaterThanEqX, BLOCK_THREADS, ITEMS_PER_THREAD>(items, GreaterThanEqX(c), f, n_items);
    BlockPredDirect_int_16_2<LessThanEqX, BLOCK_THREADS, ITEMS_PER_THREAD>(items, LessThanEqX(c), f2, n_items);
    for (int i = 0; i < HALF; ++i) f[i] &= f2[i];
    put(5, f);
}

// n_words = n_tiles * BLOCK_THREADS * ITEMS_PER_THREAD / 2 words; n_items is
// the num_items argument every tile sees.
FLS_EXPORT int rs_simd16(int items_per_thread, const uint32_t* words, const uint32_t* init, int n_tiles, int n_items,
                         uint32_t c, uint32_t* out) {
    const int n_words = n_tiles * 32 * items_per_thread / 2;
    if (items_per_thread == 32) simd16<32, 32><<<n_tiles, 32>>>(words, init, n_items, c, out, n_words);
    else if (items_per_thread == 8) simd16<32, 8><<<n_tiles, 32>>>(words, init, n_items, c, out, n_words);
    else return -1;
    return launch_status();
}
