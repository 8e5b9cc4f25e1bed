// Host encoders of the tile-based competitor formats (Shanbhag et al.;
// reference tile_based/): GPU-FOR (bin_pack, tile_based_bench_bitpack.cu:
// 6-83), GPU-DFOR (deltaBinPack, tile_based_bench_delta.cu:8-112) and GPU-RFOR
// (rleBinPack, tile_based/src/rlebinpack.cpp:9-146), word-compatible with
// them: a 4-word header {block size, 4, total count, first value}, blocks of
// 128 values (FOR, and DFOR's 4 per 512-value tile behind the tile's first
// value) or one value block and one run-length block per 512-value tile
// (RFOR); starts[] holds each block's word offset and the end.  One width per
// FOR block, as the benches encode ("simple binpack"); LSB-first packing, a
// run of n values at width w taking max(1, ceil(n w / 32)) words.  Width 32 is
// one value per word (the reference loops shift by 32 there, which C++ leaves
// undefined).  Inputs are padded to whole blocks with the last value.
// OpenMP: block sizes, an exclusive scan, then the blocks written in parallel.
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "fls/fls.h"

namespace {

inline uint32_t width_of(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

inline uint64_t run_words(uint64_t n, uint32_t bw) {
    const uint64_t w = bw >= 32 ? n : (n * bw + 31) / 32;
    return w ? w : 1;
}

// n values at width bw from a fresh word (out zeroed by the caller).
void put_run(uint32_t* out, const uint32_t* v, uint32_t n, uint32_t bw) {
    if (bw >= 32) {
        memcpy(out, v, 4ull * n);
        return;
    }
    if (bw == 0) return;
    const uint32_t mask = (1u << bw) - 1u;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t bit = (uint64_t)i * bw;
        const uint32_t w = (uint32_t)(bit >> 5), s = (uint32_t)(bit & 31u), x = v[i] & mask;
        out[w] |= x << s;
        if (s + bw > 32) out[w + 1] |= x >> (32 - s);
    }
}

struct Padded {
    const uint32_t* in;
    uint64_t n;
    uint32_t operator[](uint64_t i) const { return in[i < n ? i : n - 1]; }
};

// FOR block of 128 values: {min (0 if zero_min), widths, 4 miniblocks}.
struct ForBlock {
    uint32_t mn, bw;
};

ForBlock for_plan(const uint32_t* v, bool zero_min) {
    uint32_t mn = v[0];
    for (int i = 1; i < 128; ++i) mn = std::min(mn, v[i]);
    if (zero_min) mn = 0;
    uint32_t bw = 0;
    for (int i = 0; i < 128; ++i) bw = std::max(bw, width_of(v[i] - mn));
    return {mn, bw};
}

uint64_t for_words(uint32_t bw) { return 2 + 4 * run_words(32, bw); }

void for_write(uint32_t* out, const uint32_t* v, ForBlock p) {
    uint32_t d[128];
    for (int i = 0; i < 128; ++i) d[i] = v[i] - p.mn;
    out[0] = p.mn;
    out[1] = p.bw | (p.bw << 8) | (p.bw << 16) | (p.bw << 24);
    const uint64_t mw = run_words(32, p.bw);
    memset(out + 2, 0, 4ull * 4 * mw);
    for (int m = 0; m < 4; ++m) put_run(out + 2 + m * mw, d + 32 * m, 32, p.bw);
}

void header(uint32_t* out, uint32_t block, uint64_t n, uint32_t first) {
    out[0] = block;
    out[1] = 4;
    out[2] = (uint32_t)n;
    out[3] = first;
}

// exclusive scan of sizes into starts (offset from `base`), returns the end
uint64_t scan_into(const std::vector<uint64_t>& sizes, uint64_t base, uint32_t* starts) {
    uint64_t o = base;
    for (size_t i = 0; i < sizes.size(); ++i) {
        starts[i] = (uint32_t)o;
        o += sizes[i];
    }
    starts[sizes.size()] = (uint32_t)o;
    return o;
}

}  // namespace

extern "C" {

uint64_t fls_tile_max_words(uint64_t n) { return 8 + ((n + 511) / 512) * (4 * (2 + 4 * 32) + 8) + n; }

int fls_tile_for_encode_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts, uint64_t* n_words) {
    if (!out || !starts || !n_words || (n && !in)) return FLS_ERR_NULL;
    if (fls_tile_max_words(n) > 0xffffffffull) return FLS_ERR_SIZE;  // offsets are 32-bit words
    const Padded x{in, n};
    header(out, 128, n, n ? in[0] : 0);
    const int64_t nb = (int64_t)((n + 127) / 128);
    std::vector<ForBlock> plan(nb);
    std::vector<uint64_t> sizes(nb);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        uint32_t v[128];
        for (int i = 0; i < 128; ++i) v[i] = x[128 * (uint64_t)b + i];
        plan[b] = for_plan(v, false);
        sizes[b] = for_words(plan[b].bw);
    }
    *n_words = scan_into(sizes, 4, starts);
#pragma omp parallel for schedule(static)
    for (int64_t b = 0; b < nb; ++b) {
        uint32_t v[128];
        for (int i = 0; i < 128; ++i) v[i] = x[128 * (uint64_t)b + i];
        for_write(out + starts[b], v, plan[b]);
    }
    return FLS_OK;
}

int fls_tile_dfor_encode_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts, uint64_t* n_words) {
    if (!out || !starts || !n_words || (n && !in)) return FLS_ERR_NULL;
    if (fls_tile_max_words(n) > 0xffffffffull) return FLS_ERR_SIZE;
    const Padded x{in, n};
    header(out, 128, n, n ? in[0] : 0);
    const int64_t nt = (int64_t)((n + 511) / 512);
    std::vector<ForBlock> plan(4 * nt);
    std::vector<uint64_t> sizes(nt);
    auto deltas = [&](int64_t t, uint32_t* d) {
        const uint64_t s = 512 * (uint64_t)t;
        d[0] = 0;
        for (int i = 1; i < 512; ++i) d[i] = x[s + i] - x[s + i - 1];
    };
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        uint32_t d[512];
        deltas(t, d);
        uint64_t sz = 1;  // the tile's first value
        for (int b = 0; b < 4; ++b) {
            plan[4 * t + b] = for_plan(d + 128 * b, true);
            sz += for_words(plan[4 * t + b].bw);
        }
        sizes[t] = sz;
    }
    std::vector<uint32_t> tile_start(nt + 1);
    *n_words = scan_into(sizes, 4, tile_start.data());
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        uint32_t d[512];
        deltas(t, d);
        uint64_t o = tile_start[t];
        out[o++] = x[512 * (uint64_t)t];
        for (int b = 0; b < 4; ++b) {
            starts[4 * t + b] = (uint32_t)o;
            for_write(out + o, d + 128 * b, plan[4 * t + b]);
            o += for_words(plan[4 * t + b].bw);
        }
    }
    starts[4 * nt] = (uint32_t)*n_words;
    return FLS_OK;
}

int fls_tile_rfor_encode_u32(const uint32_t* in, uint64_t n, uint32_t* values, uint32_t* val_starts,
                             uint32_t* run_lengths, uint32_t* rl_starts, uint64_t* n_val_words,
                             uint64_t* n_rl_words) {
    if (!values || !val_starts || !run_lengths || !rl_starts || !n_val_words || !n_rl_words || (n && !in))
        return FLS_ERR_NULL;
    if (fls_tile_max_words(n) > 0xffffffffull) return FLS_ERR_SIZE;
    const Padded x{in, n};
    header(values, 512, n, n ? in[0] : 0);
    header(run_lengths, 512, n, n ? in[0] : 0);
    const int64_t nt = (int64_t)((n + 511) / 512);
    struct Runs {
        uint32_t count, vmin, rmin, vbw, rbw;
    };
    std::vector<Runs> plan(nt);
    std::vector<uint64_t> vsz(nt), rsz(nt);
    // the tile's runs: values and lengths (count <= 512)
    auto runs = [&](int64_t t, uint32_t* val, uint32_t* rl) {
        const uint64_t s = 512 * (uint64_t)t;
        uint32_t count = 0, run = 1;
        val[0] = x[s];
        for (int i = 1; i < 512; ++i) {
            if (x[s + i] != x[s + i - 1]) {
                rl[count++] = run;
                val[count] = x[s + i];
                run = 1;
            } else {
                ++run;
            }
        }
        rl[count++] = run;
        return count;
    };
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        uint32_t val[512], rl[512];
        Runs p{};
        p.count = runs(t, val, rl);
        p.vmin = *std::min_element(val, val + p.count);
        p.rmin = *std::min_element(rl, rl + p.count);
        for (uint32_t i = 0; i < p.count; ++i) {
            p.vbw = std::max(p.vbw, width_of(val[i] - p.vmin));
            p.rbw = std::max(p.rbw, width_of(rl[i] - p.rmin));
        }
        plan[t] = p;
        vsz[t] = 3 + run_words(p.count, p.vbw);
        rsz[t] = 3 + run_words(p.count, p.rbw);
    }
    *n_val_words = scan_into(vsz, 4, val_starts);
    *n_rl_words = scan_into(rsz, 4, rl_starts);
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        uint32_t val[512], rl[512];
        const Runs p = plan[t];
        runs(t, val, rl);
        for (uint32_t i = 0; i < p.count; ++i) {
            val[i] -= p.vmin;
            rl[i] -= p.rmin;
        }
        uint32_t* vb = values + val_starts[t];
        uint32_t* rb = run_lengths + rl_starts[t];
        vb[0] = p.vmin;
        rb[0] = p.rmin;
        vb[1] = p.vbw | (p.vbw << 8) | (p.vbw << 16) | (p.vbw << 24);
        rb[1] = p.rbw | (p.rbw << 8) | (p.rbw << 16) | (p.rbw << 24);
        vb[2] = rb[2] = p.count;
        memset(vb + 3, 0, 4ull * run_words(p.count, p.vbw));
        memset(rb + 3, 0, 4ull * run_words(p.count, p.rbw));
        put_run(vb + 3, val, p.count, p.vbw);
        put_run(rb + 3, rl, p.count, p.rbw);
    }
    return FLS_OK;
}

}  // extern "C"
