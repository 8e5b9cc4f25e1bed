/* CPU restatement of the tile-based competitor formats of Shanbhag et al.
 * (reference tile_based/): GPU-FOR, GPU-DFOR and GPU-RFOR.  TEST
 * INFRASTRUCTURE ONLY -- the checker of the device decoders, never the thing
 * measured or shipped.
 *
 * Layout shared by the three (reference encoders cited per function): a
 * 4-word header {block_size, miniblock_count, total_count, first value},
 * then blocks; starts[b] is the word offset of block b and starts[n_blocks]
 * the end.  Packing inside a block is linear LSB-first.  The reference
 * encoders shift a 32-bit value by 32 when a width is 32 (undefined in C++);
 * here a width-32 block is packed linearly, which the reference decoders read
 * correctly.  Inputs are padded to whole blocks / tiles with the last value.
 *
 * Pinning: GPU-RFOR word for word against the reference's own rleBinPack
 * (tile_based/src/rlebinpack.cpp:9-146, built into oracle/_ref) for inputs
 * whose widths stay below 32; GPU-FOR and GPU-DFOR (encoders inside CUDA
 * benches, tile_based_bench_bitpack.cu:6-83 and tile_based_bench_delta.cu:
 * 8-112, which need nvcc and CUB: unbuildable here) by hand-computed blocks
 * and round trips; every decoder by round trips over all widths 0..32. */
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static unsigned width_of(uint32_t v) {
    unsigned b = 0;
    while (b < 32 && (v >> b) != 0) ++b;
    return b;
}

/* Append n values at width bw, LSB-first, starting at a fresh word; returns
 * the words written: max(1, ceil(n*bw/32)).  For bw < 32 this is the
 * reference loop (tile_based_bench_bitpack.cu:63-78, rlebinpack.cpp:104-130);
 * at bw == 32 that loop's straddle branch shifts a 32-bit value by 32, so
 * width 32 is defined here as value i in word i. */
static uint64_t put_linear(uint32_t* out, const uint32_t* v, unsigned n, unsigned bw) {
    if (bw >= 32) {
        for (unsigned i = 0; i < n; ++i) out[i] = v[i];
        if (n == 0) out[0] = 0;
        return n ? n : 1;
    }
    uint64_t o = 0;
    unsigned shift = 0;
    out[0] = 0;
    for (unsigned i = 0; i < n; ++i) {
        const uint32_t x = v[i] & ((1u << bw) - 1u);
        if (shift + bw > 32) {
            if (shift != 32) out[o] |= x << shift;
            ++o;
            shift = (shift + bw) & 31u;
            out[o] = x >> (bw - shift);
        } else {
            out[o] |= shift < 32 ? x << shift : 0u;
            shift += bw;
        }
    }
    return o + 1;
}

static uint32_t get_linear(const uint32_t* w, unsigned k, unsigned bw) {
    if (bw == 0) return 0;
    const uint64_t bit = (uint64_t)k * bw;
    const uint64_t i = bit >> 5;
    const unsigned s = (unsigned)(bit & 31u);
    uint64_t x = w[i];
    if (s + bw > 32) x |= (uint64_t)w[i + 1] << 32;
    return (uint32_t)((x >> s) & (bw >= 32 ? 0xffffffffull : ((1ull << bw) - 1ull)));
}

static void header(uint32_t* out, uint32_t block, uint32_t n, uint32_t first) {
    out[0] = block;
    out[1] = 4;
    out[2] = n;
    out[3] = first;
}

/* One FOR block of 128 values at one width (the "simple binpack" the benches
 * use: every miniblock at the block's widest): {min, widths, 4 miniblocks}. */
static uint64_t for_block(uint32_t* out, const uint32_t* v, int force_zero_min) {
    uint32_t mn = v[0], d[128];
    for (int i = 1; i < 128; ++i)
        if (v[i] < mn) mn = v[i];
    if (force_zero_min) mn = 0;
    unsigned bw = 0;
    for (int i = 0; i < 128; ++i) {
        d[i] = v[i] - mn;
        const unsigned b = width_of(d[i]);
        if (b > bw) bw = b;
    }
    out[0] = mn;
    out[1] = bw | (bw << 8) | (bw << 16) | (bw << 24);
    uint64_t o = 2;
    for (int m = 0; m < 4; ++m) o += put_linear(out + o, d + 32 * m, 32, bw);
    return o;
}

static uint32_t padded(const uint32_t* in, uint64_t n, uint64_t i) { return in[i < n ? i : n - 1]; }

/* GPU-FOR (bin_pack, tile_based_bench_bitpack.cu:6-83). */
uint64_t orc_tile_for_encode(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts) {
    header(out, 128, (uint32_t)n, n ? in[0] : 0);
    uint64_t o = 4;
    const uint64_t nb = (n + 127) / 128;
    uint32_t v[128];
    for (uint64_t b = 0; b < nb; ++b) {
        for (int i = 0; i < 128; ++i) v[i] = padded(in, n, 128 * b + i);
        starts[b] = (uint32_t)o;
        o += for_block(out + o, v, 0);
    }
    starts[nb] = (uint32_t)o;
    return o;
}

/* decodeElement (binpack_kernel.cuh:3-18): miniblock m starts at the sum of
 * the widths before it. */
static uint32_t for_element(const uint32_t* blk, unsigned e) {
    const unsigned m = e / 32, idx = e % 32;
    const uint32_t widths = blk[1];
    unsigned off = 0;
    for (unsigned j = 0; j < m; ++j) off += (widths >> (8 * j)) & 255u;
    const unsigned bw = (widths >> (8 * m)) & 255u;
    return blk[0] + get_linear(blk + 2 + off, idx, bw);
}

void orc_tile_for_decode(const uint32_t* data, const uint32_t* starts, uint64_t n, uint32_t* out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = for_element(data + starts[i / 128], (unsigned)(i % 128));
}

/* GPU-DFOR (deltaBinPack, tile_based_bench_delta.cu:8-112): tiles of 512,
 * each {first value, 4 FOR blocks of the deltas with min forced to 0}; the
 * tile's first delta is 0. */
uint64_t orc_tile_dfor_encode(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts) {
    header(out, 128, (uint32_t)n, n ? in[0] : 0);
    uint64_t o = 4;
    const uint64_t nt = (n + 511) / 512;
    uint32_t d[512];
    for (uint64_t t = 0; t < nt; ++t) {
        const uint64_t s = 512 * t;
        out[o++] = padded(in, n, s);
        d[0] = 0;
        for (int i = 1; i < 512; ++i) d[i] = padded(in, n, s + i) - padded(in, n, s + i - 1);
        for (int b = 0; b < 4; ++b) {
            starts[4 * t + b] = (uint32_t)o;
            o += for_block(out + o, d + 128 * b, 1);
        }
    }
    starts[4 * nt] = (uint32_t)o;
    return o;
}

/* LoadDBinPack (deltabinpack_kernel.cuh:39-103): decode the tile's deltas,
 * put the first value in element 0, inclusive sum. */
void orc_tile_dfor_decode(const uint32_t* data, const uint32_t* starts, uint64_t n, uint32_t* out) {
    const uint64_t nt = (n + 511) / 512;
    for (uint64_t t = 0; t < nt; ++t) {
        uint32_t acc = 0;
        for (unsigned e = 0; e < 512 && 512 * t + e < n; ++e) {
            const uint32_t* blk = data + starts[4 * t + e / 128];
            const uint32_t d = e == 0 ? data[starts[4 * t] - 1] : for_element(blk, e % 128);
            acc += d;
            out[512 * t + e] = acc;
        }
    }
}

/* GPU-RFOR (rleBinPack, rlebinpack.cpp:9-146): tiles of 512; per tile one
 * block of run values and one of run lengths, each {min, width x4, count,
 * linear packing}. */
void orc_tile_rfor_encode(const uint32_t* in, uint64_t n, uint32_t* values, uint32_t* val_starts,
                          uint32_t* run_lengths, uint32_t* rl_starts, uint64_t* n_val_words, uint64_t* n_rl_words) {
    header(values, 512, (uint32_t)n, n ? in[0] : 0);
    header(run_lengths, 512, (uint32_t)n, n ? in[0] : 0);
    uint64_t vo = 4, ro = 4;
    const uint64_t nt = (n + 511) / 512;
    uint32_t val[512], rl[512];
    for (uint64_t t = 0; t < nt; ++t) {
        const uint64_t s = 512 * t;
        unsigned count = 0, run = 1;
        val[0] = padded(in, n, s);
        for (int i = 1; i < 512; ++i) {
            if (padded(in, n, s + i) != padded(in, n, s + i - 1)) {
                rl[count++] = run;
                val[count] = padded(in, n, s + i);
                run = 1;
            } else {
                ++run;
            }
        }
        rl[count++] = run;
        val_starts[t] = (uint32_t)vo;
        rl_starts[t] = (uint32_t)ro;
        uint32_t mv = val[0], mr = rl[0];
        for (unsigned i = 1; i < count; ++i) {
            if (val[i] < mv) mv = val[i];
            if (rl[i] < mr) mr = rl[i];
        }
        unsigned vbw = 0, rbw = 0;
        for (unsigned i = 0; i < count; ++i) {
            val[i] -= mv;
            rl[i] -= mr;
            if (width_of(val[i]) > vbw) vbw = width_of(val[i]);
            if (width_of(rl[i]) > rbw) rbw = width_of(rl[i]);
        }
        values[vo] = mv;
        run_lengths[ro] = mr;
        values[vo + 1] = vbw | (vbw << 8) | (vbw << 16) | (vbw << 24);
        run_lengths[ro + 1] = rbw | (rbw << 8) | (rbw << 16) | (rbw << 24);
        values[vo + 2] = count;
        run_lengths[ro + 2] = count;
        vo += 3 + put_linear(values + vo + 3, val, count, vbw);
        ro += 3 + put_linear(run_lengths + ro + 3, rl, count, rbw);
    }
    val_starts[nt] = (uint32_t)vo;
    rl_starts[nt] = (uint32_t)ro;
    *n_val_words = vo;
    *n_rl_words = ro;
}

/* LoadRBinPack (rlebinpack_kernel.cuh:20-146): decode the runs, expand them
 * over the tile's 512 positions. */
void orc_tile_rfor_decode(const uint32_t* values, const uint32_t* val_starts, const uint32_t* run_lengths,
                          const uint32_t* rl_starts, uint64_t n, uint32_t* out) {
    const uint64_t nt = (n + 511) / 512;
    for (uint64_t t = 0; t < nt; ++t) {
        const uint32_t* vb = values + val_starts[t];
        const uint32_t* rb = run_lengths + rl_starts[t];
        const unsigned count = vb[2];
        uint64_t p = 512 * t;
        for (unsigned r = 0; r < count; ++r) {
            const uint32_t v = vb[0] + get_linear(vb + 3, r, vb[1] & 255u);
            const uint32_t len = rb[0] + get_linear(rb + 3, r, rb[1] & 255u);
            for (uint32_t k = 0; k < len && p < n && p < 512 * (t + 1); ++k) out[p++] = v;
        }
    }
}

/* Words the encoders may write for n values (capacity for the callers). */
uint64_t orc_tile_max_words(uint64_t n) { return 8 + ((n + 511) / 512) * (4 * (2 + 4 * 32) + 8) + n; }
