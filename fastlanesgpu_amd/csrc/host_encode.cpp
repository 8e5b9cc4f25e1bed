// Host (CPU) encoders of the C ABI: FastLanes pack, transpose, unrsum, and
// their column-level drivers (bit-packing, FastLanes delta).  Encoding is host
// work in the reference as well (fastlanes/src/pack.cpp, transpose.cpp,
// unrsum.cpp, called per vector from every driver, e.g. fls_q11.cu:207-223).
//
// Built with the host compiler and OpenMP; vectors are encoded in parallel.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

#include "fls/fls.h"

namespace {

// transpose_i moves k = 64a + 8b + c to a + 16*rev3(b) + 128*c
// (fastlanes/src/transpose.cpp:4110-5135).
struct TransposeTable {
    std::array<uint16_t, 1024> to;  // to[k] = transposed position of k
    TransposeTable() {
        for (unsigned k = 0; k < 1024; ++k) {
            unsigned a = k >> 6, b = (k >> 3) & 7u, c = k & 7u;
            unsigned rb = ((b & 1u) << 2) | (b & 2u) | (b >> 2);
            to[k] = (uint16_t)(a + 16u * rb + 128u * c);
        }
    }
};
const TransposeTable kT;

// Register-row order of the FastLanes delta chain (rsum.cuh:8-106).
constexpr unsigned rsum_row(unsigned j) {
    return 4u * (j % 8u) + (j / 8u == 0 ? 0u : j / 8u == 1 ? 2u : j / 8u == 2 ? 1u : 3u);
}

// Lane-at-a-time bit packing: each lane's 32 rows are streamed through a
// 64-bit accumulator and flushed one 32-bit word at a time.
void pack_vec(const uint32_t* __restrict in, uint32_t* __restrict out, unsigned bw) {
    if (bw == 0) return;
    if (bw == 32) {
        memcpy(out, in, 4096);
        return;
    }
    const uint32_t mask = (1u << bw) - 1u;
    for (unsigned l = 0; l < 32; ++l) {
        uint64_t acc = 0;
        unsigned nbits = 0, w = 0;
        for (unsigned r = 0; r < 32; ++r) {
            acc |= (uint64_t)(in[32u * r + l] & mask) << nbits;
            nbits += bw;
            if (nbits >= 32) {
                out[32u * w + l] = (uint32_t)acc;
                ++w;
                acc >>= 32;
                nbits -= 32;
            }
        }
    }
}

void transpose_i(const uint32_t* __restrict in, uint32_t* __restrict out) {
    for (unsigned k = 0; k < 1024; ++k) out[kT.to[k]] = in[k];
}

void transpose_o(const uint32_t* __restrict in, uint32_t* __restrict out) {
    for (unsigned k = 0; k < 1024; ++k) out[k] = in[kT.to[k]];
}

// unrsum (unrsum.cpp:120-256): per lane, the first chain element becomes 0
// and every later one the difference to its chain predecessor.
void unrsum(const uint32_t* __restrict in, uint32_t* __restrict out) {
    for (unsigned l = 0; l < 32; ++l) {
        uint32_t prev = in[l];
        out[l] = 0;
        for (unsigned j = 1; j < 32; ++j) {
            const unsigned p = l + 32u * rsum_row(j);
            out[p] = in[p] - prev;
            prev = in[p];
        }
    }
}

unsigned width_of(uint32_t x) { return x ? 32u - (unsigned)__builtin_clz(x) : 0u; }

// Copy vector v of an n-value column into a full 1024-value buffer, padding
// the tail with `pad_last ? last value : 0`.
void load_vec(const uint32_t* in, uint64_t n, uint64_t v, bool pad_last, uint32_t* buf) {
    const uint64_t begin = v * 1024u;
    const uint64_t cnt = std::min<uint64_t>(1024u, n - begin);
    memcpy(buf, in + begin, cnt * 4u);
    const uint32_t pad = (pad_last && cnt) ? buf[cnt - 1] : 0u;
    for (uint64_t i = cnt; i < 1024u; ++i) buf[i] = pad;
}

}  // namespace

extern "C" {

int fls_pack_u32(const uint32_t* in, uint32_t* out, uint32_t bw) {
    if (bw > 32) return FLS_ERR_BW;
    if (bw && (!in || !out)) return FLS_ERR_NULL;
    pack_vec(in, out, bw);
    return FLS_OK;
}

int fls_transpose_i_u32(const uint32_t* in, uint32_t* out) {
    if (!in || !out) return FLS_ERR_NULL;
    transpose_i(in, out);
    return FLS_OK;
}

int fls_transpose_o_u32(const uint32_t* in, uint32_t* out) {
    if (!in || !out) return FLS_ERR_NULL;
    transpose_o(in, out);
    return FLS_OK;
}

int fls_unrsum_u32(const uint32_t* in, uint32_t* out) {
    if (!in || !out) return FLS_ERR_NULL;
    unrsum(in, out);
    return FLS_OK;
}

uint64_t fls_packed_words(uint64_t n, uint32_t bw) { return ((n + 1023u) / 1024u) * 32u * bw; }

int fls_pack_column_u32(const uint32_t* in, uint64_t n, uint32_t bw, uint32_t flags, uint32_t* out) {
    if (bw > 32) return FLS_ERR_BW;
    if (n == 0 || bw == 0) return FLS_OK;
    if (!in || !out) return FLS_ERR_NULL;
    const int64_t n_vec = (int64_t)((n + 1023u) / 1024u);
    const bool tr = (flags & FLS_ENC_FLAG_TRANSPOSE) != 0;
#pragma omp parallel
    {
        std::vector<uint32_t> buf(1024), tb(1024);
#pragma omp for schedule(static)
        for (int64_t v = 0; v < n_vec; ++v) {
            load_vec(in, n, (uint64_t)v, false, buf.data());
            const uint32_t* src = buf.data();
            if (tr) {
                transpose_i(buf.data(), tb.data());
                src = tb.data();
            }
            pack_vec(src, out + (uint64_t)v * 32u * bw, bw);
        }
    }
    return FLS_OK;
}

int fls_delta_encode_column_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint64_t out_cap,
                                fls_vmeta* vmeta, uint32_t* bases, uint64_t* out_words) {
    if (n == 0) {
        if (out_words) *out_words = 0;
        return FLS_OK;
    }
    if (!in || !vmeta || !bases) return FLS_ERR_NULL;
    const int64_t n_vec = (int64_t)((n + 1023u) / 1024u);
    // pass 1: per-vector width and bases
    std::vector<uint8_t> bws((size_t)n_vec);
#pragma omp parallel
    {
        std::vector<uint32_t> buf(1024), t(1024), d(1024);
#pragma omp for schedule(static)
        for (int64_t v = 0; v < n_vec; ++v) {
            load_vec(in, n, (uint64_t)v, true, buf.data());
            transpose_i(buf.data(), t.data());
            memcpy(bases + (uint64_t)v * 32u, t.data(), 128);
            unrsum(t.data(), d.data());
            uint32_t mx = 0;
            for (unsigned i = 0; i < 1024; ++i) mx |= d[i];
            bws[(size_t)v] = (uint8_t)width_of(mx);
        }
    }
    // offsets (exclusive scan of 32*bw)
    uint64_t total = 0;
    for (int64_t v = 0; v < n_vec; ++v) {
        vmeta[v] = FLS_VMETA(total, bws[(size_t)v]);
        total += 32u * bws[(size_t)v];
    }
    if (out_words) *out_words = total;
    if (total > out_cap) return FLS_ERR_CAPACITY;
    if (total && !out) return FLS_ERR_NULL;
    // pass 2: pack the differences
#pragma omp parallel
    {
        std::vector<uint32_t> buf(1024), t(1024), d(1024);
#pragma omp for schedule(static)
        for (int64_t v = 0; v < n_vec; ++v) {
            const unsigned bw = FLS_VMETA_BW(vmeta[v]);
            if (!bw) continue;
            load_vec(in, n, (uint64_t)v, true, buf.data());
            transpose_i(buf.data(), t.data());
            unrsum(t.data(), d.data());
            pack_vec(d.data(), out + FLS_VMETA_OFFSET(vmeta[v]), bw);
        }
    }
    return FLS_OK;
}

}  // extern "C"
