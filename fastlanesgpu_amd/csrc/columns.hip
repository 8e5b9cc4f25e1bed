// Column descriptor validation and launch-shape helpers.
#include "fls_columns.hpp"

#include <mutex>
#include <utility>
#include <vector>

namespace fls {

int device_cus(uint32_t* cus) {
    static int cached[64] = {0};
    int dev = 0;
    FLS_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return FLS_ERR_NO_DEVICE;
    if (!cached[dev]) {
        int n = 0;
        FLS_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        if (n <= 0) return FLS_ERR_NO_DEVICE;
        cached[dev] = n;
    }
    *cus = (uint32_t)cached[dev];
    return FLS_OK;
}

int blocks_per_cu(const void* kernel, int block, size_t lds, int* per_cu) {
    struct Key {
        const void* k;
        int block;
        size_t lds;
        int dev;
    };
    static std::mutex mu;
    static std::vector<std::pair<Key, int>> cache;
    int dev = 0;
    FLS_HIP_CHECK(hipGetDevice(&dev));
    {
        std::lock_guard<std::mutex> g(mu);
        for (const auto& e : cache)
            if (e.first.k == kernel && e.first.block == block && e.first.lds == lds && e.first.dev == dev) {
                *per_cu = e.second;
                return FLS_OK;
            }
    }
    int n = 0;
    FLS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, block, lds));
    if (n < 1) n = 1;
    std::lock_guard<std::mutex> g(mu);
    cache.push_back({Key{kernel, block, lds, dev}, n});
    *per_cu = n;
    return FLS_OK;
}

uint32_t balanced_grid(uint64_t max_blocks, uint64_t work) {
    if (max_blocks == 0) max_blocks = 1;
    if (work <= max_blocks) return (uint32_t)(work ? work : 1);
    const uint64_t rounds = ceil_div(work, max_blocks);
    return (uint32_t)ceil_div(work, rounds);
}

int resident_grid_of(const void* kernel, int block, size_t lds, uint64_t work, uint32_t* grid) {
    int per_cu = 0;
    int st = blocks_per_cu(kernel, block, lds, &per_cu);
    if (st) return st;
    uint32_t cus = 0;
    if ((st = device_cus(&cus)) != FLS_OK) return st;
    *grid = balanced_grid((uint64_t)cus * (uint64_t)per_cu, work);
    return FLS_OK;
}

int persistent_grid(uint64_t work, uint32_t blocks_per_cu, uint32_t* grid) {
    uint32_t cus = 0;
    int st = device_cus(&cus);
    if (st) return st;
    *grid = balanced_grid((uint64_t)cus * blocks_per_cu, work);
    return FLS_OK;
}

int to_dev_column(const fls_column& c, uint64_t n_rows, ColDev* out, bool allow_var) {
    const uint64_t n_vec = ceil_div(n_rows, 1024);
    out->enc = c.encoding;
    out->bw = c.bw;
    out->base = c.base;
    out->transposed = c.transposed;
    out->words = c.words;
    out->vmeta = c.vmeta;
    out->bases = c.bases;
    out->aux = c.aux;
    switch (c.encoding) {
        case FLS_ENC_BITPACK:
            if (c.bw > 32) return FLS_ERR_BW;
            if (n_rows && c.bw && !c.words) return FLS_ERR_NULL;
            if (c.n_words < n_vec * 32u * c.bw) return FLS_ERR_SIZE;
            return FLS_OK;
        case FLS_ENC_DELTA:
            if (!allow_var) return FLS_ERR_ENCODING;
            if (n_rows && (!c.vmeta || !c.bases)) return FLS_ERR_NULL;
            if (n_rows && c.n_words && !c.words) return FLS_ERR_NULL;
            out->transposed = 1;
            return FLS_OK;
        case FLS_ENC_RLE:
            if (!allow_var) return FLS_ERR_ENCODING;
            if (n_rows && (!c.vmeta || !c.bases || !c.aux || !c.words)) return FLS_ERR_NULL;
            out->transposed = 1;
            return FLS_OK;
        case FLS_ENC_MIXED:
            if (!allow_var) return FLS_ERR_ENCODING;
            if (n_rows && (!c.vmeta || !c.aux || !c.words)) return FLS_ERR_NULL;
            out->transposed = 1;
            return FLS_OK;
        default:
            return FLS_ERR_ENCODING;
    }
}

}  // namespace fls

extern "C" uint64_t fls_column_bytes(const fls_column* c, uint64_t n_rows) {
    if (!c) return 0;
    const uint64_t n_vec = (n_rows + 1023u) / 1024u;
    switch (c->encoding) {
        case FLS_ENC_BITPACK: return n_vec * 32u * c->bw * 4u;
        case FLS_ENC_DELTA: return c->n_words * 4u + n_vec * (sizeof(fls_vmeta) + 32u * 4u);
        case FLS_ENC_RLE: return c->n_words * 4u + n_vec * (sizeof(fls_vmeta) + sizeof(fls_rle_meta) + 32u * 4u);
        case FLS_ENC_MIXED: return c->n_words * 4u + n_vec * 16u;
        default: return c->n_words * 4u;
    }
}
