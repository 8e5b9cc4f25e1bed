/*
 * Same-run CPU baselines (TEST INFRASTRUCTURE: loaded only by the tests and
 * by bench.py's cpu_baseline legs; the product library never uses it).
 *
 * Each function is the oracle's scalar decode of encoded vectors
 * (orc_decode_vector: per-lane word-window extraction as in
 * fastlanes/src/include/fls_gen/unpack/unpack.cuh:877-971, FastLanes rsum as in
 * fls_gen/rsum/rsum.cuh:3-107) followed by the scalar evaluation of the
 * matching plain oracle function (orc_q21 / orc_q31 / orc_q41 / the
 * batched-segments range count / orc_exhaustive_class), parallelised with
 * OpenMP over vectors or segments.  Pinned in tests/test_cpu_baseline.py
 * against those plain functions on the same values.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* original row of stored position k inside a vector */
static uint64_t row_of(const orc_col* c, uint64_t v, unsigned k, const uint32_t* inv) {
    return v * 1024 + (c->transposed ? inv[k] : k);
}

static void inverse_transpose(uint32_t* inv) {
    uint32_t id[1024], t[1024];
    for (unsigned i = 0; i < 1024; ++i) id[i] = i;
    orc_transpose_i_u32(id, t); /* t[p] = original index at stored position p */
    for (unsigned p = 0; p < 1024; ++p) inv[p] = t[p];
}

uint64_t orc_c1_encoded(const uint32_t* words, uint64_t n_values, unsigned bw, uint32_t* out, int n_threads) {
    const uint64_t n_vec = (n_values + 1023) / 1024;
    uint64_t sum = 0;
#pragma omp parallel for num_threads(n_threads) reduction(+ : sum) schedule(static)
    for (int64_t vv = 0; vv < (int64_t)n_vec; ++vv) {
        const uint64_t v = (uint64_t)vv;
        uint32_t d[1024];
        orc_unpack_lanewise_u32(words + v * 32 * bw, d, bw);
        for (unsigned k = 0; k < 1024 && v * 1024 + k < n_values; ++k) {
            out[v * 1024 + k] = d[k];
            sum += (v * 1024 + k + 1) * (uint64_t)d[k];
        }
    }
    return sum;
}

#define Q21_SLOTS 7000u

void orc_q21_encoded(const orc_col* date, const orc_col* supp, const orc_col* part, const orc_col* rev,
                     uint64_t n_rows, uint64_t v_begin, uint64_t v_end, const uint32_t* d_year,
                     const uint32_t* s_region, const uint32_t* p_category, const uint32_t* p_brand, uint32_t n_date,
                     uint32_t n_supp, uint32_t n_part, uint32_t keys_min, uint32_t cat, uint32_t reg, int n_threads,
                     int64_t* sums, uint64_t* counts) {
    uint32_t inv[1024];
    inverse_transpose(inv);
    memset(sums, 0, sizeof(int64_t) * Q21_SLOTS);
    memset(counts, 0, sizeof(uint64_t) * Q21_SLOTS);
#pragma omp parallel num_threads(n_threads)
    {
        int64_t* ls = calloc(Q21_SLOTS, sizeof(int64_t));
        uint64_t* lc = calloc(Q21_SLOTS, sizeof(uint64_t));
        uint32_t a[1024], b[1024], p[1024], r[1024];
#pragma omp for schedule(static)
        for (int64_t vv = (int64_t)v_begin; vv < (int64_t)v_end; ++vv) {
            const uint64_t v = (uint64_t)vv;
            orc_decode_vector(supp, v, b);
            orc_decode_vector(part, v, p);
            orc_decode_vector(date, v, a);
            orc_decode_vector(rev, v, r);
            for (unsigned k = 0; k < 1024; ++k) {
                if (row_of(rev, v, k, inv) >= n_rows) continue;
                const uint32_t si = b[k] - keys_min, pi = p[k] - keys_min, di = a[k] - keys_min;
                if (si >= n_supp || pi >= n_part || di >= n_date) continue;
                if (s_region[si] != reg || p_category[pi] != cat) continue;
                const uint32_t slot = (p_brand[pi] * 7 + (d_year[di] - 1992)) % Q21_SLOTS; /* fls_q21.cu:71 */
                ls[slot] += (int64_t)r[k];
                lc[slot] += 1;
            }
        }
#pragma omp critical
        for (unsigned i = 0; i < Q21_SLOTS; ++i) {
            sums[i] += ls[i];
            counts[i] += lc[i];
        }
        free(ls);
        free(lc);
    }
}

void orc_q31_encoded(const orc_col* date, const orc_col* cust, const orc_col* supp, const orc_col* rev,
                     uint64_t n_rows, uint64_t v_begin, uint64_t v_end, const uint32_t* d_year,
                     const uint32_t* s_region, const uint32_t* s_nation, const uint32_t* c_region,
                     const uint32_t* c_nation, uint32_t n_date, uint32_t n_supp, uint32_t n_cust, uint32_t keys_min,
                     uint32_t reg_c, uint32_t reg_s, uint32_t year_lo, uint32_t year_hi, int n_threads,
                     int64_t* sums, uint64_t* counts) {
    uint32_t inv[1024];
    inverse_transpose(inv);
    memset(sums, 0, sizeof(int64_t) * 4375);
    memset(counts, 0, sizeof(uint64_t) * 4375);
#pragma omp parallel num_threads(n_threads)
    {
        int64_t* ls = calloc(4375, sizeof(int64_t));
        uint64_t* lc = calloc(4375, sizeof(uint64_t));
        uint32_t a[1024], c[1024], s[1024], r[1024];
#pragma omp for schedule(static)
        for (int64_t vv = (int64_t)v_begin; vv < (int64_t)v_end; ++vv) {
            const uint64_t v = (uint64_t)vv;
            orc_decode_vector(supp, v, s);
            orc_decode_vector(cust, v, c);
            orc_decode_vector(date, v, a);
            orc_decode_vector(rev, v, r);
            for (unsigned k = 0; k < 1024; ++k) {
                if (row_of(rev, v, k, inv) >= n_rows) continue;
                const uint32_t si = s[k] - keys_min, ci = c[k] - keys_min, di = a[k] - keys_min;
                if (si >= n_supp || ci >= n_cust || di >= n_date) continue;
                if (s_region[si] != reg_s || c_region[ci] != reg_c) continue;
                const uint32_t year = d_year[di];
                if (year < year_lo || year > year_hi) continue;
                const uint32_t slot = (s_nation[si] * 25 * 7 + c_nation[ci] * 7 + (year - 1992)) % 4375;
                ls[slot] += (int64_t)r[k];
                lc[slot] += 1;
            }
        }
#pragma omp critical
        for (unsigned i = 0; i < 4375; ++i) {
            sums[i] += ls[i];
            counts[i] += lc[i];
        }
        free(ls);
        free(lc);
    }
}

void orc_q41_encoded(const orc_col* date, const orc_col* cust, const orc_col* supp, const orc_col* part,
                     const orc_col* rev, const orc_col* cost, uint64_t n_rows, uint64_t v_begin, uint64_t v_end,
                     const uint32_t* d_year, const uint32_t* s_region, const uint32_t* c_region,
                     const uint32_t* c_nation, const uint32_t* p_mfgr, uint32_t n_date, uint32_t n_supp,
                     uint32_t n_cust, uint32_t n_part, uint32_t keys_min, uint32_t reg_c, uint32_t reg_s,
                     uint32_t mfgr_lo, uint32_t mfgr_hi, int n_threads, int64_t* sums, uint64_t* counts) {
    uint32_t inv[1024];
    inverse_transpose(inv);
    memset(sums, 0, sizeof(int64_t) * 175);
    memset(counts, 0, sizeof(uint64_t) * 175);
#pragma omp parallel num_threads(n_threads)
    {
        int64_t ls[175] = {0};
        uint64_t lc[175] = {0};
        uint32_t a[1024], c[1024], s[1024], p[1024], r[1024], x[1024];
#pragma omp for schedule(static)
        for (int64_t vv = (int64_t)v_begin; vv < (int64_t)v_end; ++vv) {
            const uint64_t v = (uint64_t)vv;
            orc_decode_vector(supp, v, s);
            orc_decode_vector(cust, v, c);
            orc_decode_vector(part, v, p);
            orc_decode_vector(date, v, a);
            orc_decode_vector(rev, v, r);
            orc_decode_vector(cost, v, x);
            for (unsigned k = 0; k < 1024; ++k) {
                if (row_of(rev, v, k, inv) >= n_rows) continue;
                const uint32_t si = s[k] - keys_min, ci = c[k] - keys_min, pi = p[k] - keys_min,
                               di = a[k] - keys_min;
                if (si >= n_supp || ci >= n_cust || pi >= n_part || di >= n_date) continue;
                if (s_region[si] != reg_s || c_region[ci] != reg_c) continue;
                if (p_mfgr[pi] < mfgr_lo || p_mfgr[pi] > mfgr_hi) continue;
                const uint32_t slot = (c_nation[ci] * 7 + (d_year[di] - 1992)) % 175;
                ls[slot] += (int64_t)r[k] - (int64_t)x[k];
                lc[slot] += 1;
            }
        }
#pragma omp critical
        for (unsigned i = 0; i < 175; ++i) {
            sums[i] += ls[i];
            counts[i] += lc[i];
        }
    }
}

void orc_c4_encoded(const orc_col* segs, const uint64_t* n_values, const uint32_t* lo, const uint32_t* hi,
                    uint32_t n_seg, int n_threads, uint64_t* counts) {
    uint32_t inv[1024];
    inverse_transpose(inv);
    for (uint32_t s = 0; s < n_seg; ++s) counts[s] = 0;
    /* (segment, vector) pairs flattened so long segments spread over threads */
    uint64_t* first = malloc(sizeof(uint64_t) * (n_seg + 1));
    first[0] = 0;
    for (uint32_t s = 0; s < n_seg; ++s) first[s + 1] = first[s] + (n_values[s] + 1023) / 1024;
#pragma omp parallel num_threads(n_threads)
    {
        uint64_t* lc = calloc(n_seg ? n_seg : 1, sizeof(uint64_t));
        uint32_t d[1024];
        uint32_t s = 0;
#pragma omp for schedule(static)
        for (int64_t gg = 0; gg < (int64_t)first[n_seg]; ++gg) {
            const uint64_t g = (uint64_t)gg;
            while (first[s + 1] <= g) ++s;
            while (first[s] > g) --s;
            const uint64_t v = g - first[s];
            orc_decode_vector(&segs[s], v, d);
            uint64_t c = 0;
            for (unsigned k = 0; k < 1024; ++k)
                if (row_of(&segs[s], v, k, inv) < n_values[s] && d[k] >= lo[s] && d[k] <= hi[s]) ++c;
            lc[s] += c;
        }
#pragma omp critical
        for (uint32_t i = 0; i < n_seg; ++i) counts[i] += lc[i];
        free(lc);
    }
    free(first);
}

void orc_exhaustive_classes(const uint32_t* in, uint64_t n_vec, uint8_t* cls, int n_threads) {
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 64)
    for (int64_t v = 0; v < (int64_t)n_vec; ++v) cls[v] = (uint8_t)orc_exhaustive_class(in + (uint64_t)v * 1024);
}
