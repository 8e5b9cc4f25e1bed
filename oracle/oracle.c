/*
 * CPU oracle for the FastLanes decode + Crystal query hot path.
 * TEST INFRASTRUCTURE ONLY: the checker for the product library, never part of it.
 * Declarations, citations and the pinning status of each function: oracle.h.
 *
 * Style is intentionally naive (one bit at a time, one row at a time) so that
 * it shares no structure with the product's word-parallel encoders/decoders.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ===================================================================== layout
 * The FastLanes vector: 1024 values as 32 lanes x 32 rows; value index 32*r+l
 * is row r of lane l.  Lane l's values form an LSB-first bit stream, row r at
 * bits [r*bw, r*bw+bw); bit p of lane l lives in word 32*(p/32)+l, bit p%32.
 * (fastlanes/src/pack.cpp:2970-3123 writes exactly these words; the reference
 * masks every input value to bw bits first, pack.cpp:2975.)                  */
void orc_pack_u32(const uint32_t* in, uint32_t* out, unsigned bw) {
    if (bw == 0) return; /* pack_0bit_32ow writes nothing */
    memset(out, 0, sizeof(uint32_t) * 32u * bw);
    for (unsigned l = 0; l < 32; ++l) {
        for (unsigned r = 0; r < 32; ++r) {
            uint64_t v = in[32 * r + l];
            for (unsigned b = 0; b < bw; ++b) {
                if ((v >> b) & 1u) {
                    unsigned p = r * bw + b;
                    out[32 * (p / 32) + l] |= 1u << (p % 32);
                }
            }
        }
    }
}

void orc_unpack_u32(const uint32_t* in, uint32_t* out, unsigned bw) {
    for (unsigned l = 0; l < 32; ++l) {
        for (unsigned r = 0; r < 32; ++r) {
            uint32_t v = 0;
            for (unsigned b = 0; b < bw; ++b) {
                unsigned p = r * bw + b;
                uint32_t bit = (in[32 * (p / 32) + l] >> (p % 32)) & 1u;
                v |= bit << b;
            }
            out[32 * r + l] = v;
        }
    }
}

/* =================================================================== transpose
 * transpose_i (transpose.cpp:4110-5135) moves input k = 64a + 8b + c
 * (a<16, b<8, c<8) to output a + 16*rev3(b) + 128*c, where rev3 reverses the
 * three bits of b (k=8 -> 64, k=16 -> 32, k=32 -> 16, k=64 -> 1).           */
static unsigned rev3(unsigned b) { return ((b & 1u) << 2) | (b & 2u) | ((b >> 2) & 1u); }

static unsigned transpose_pos(unsigned k) {
    unsigned a = k / 64, b = (k / 8) % 8, c = k % 8;
    return a + 16 * rev3(b) + 128 * c;
}

/* original index of transposed position p (inverse of transpose_pos; rev3 is
 * its own inverse) */
static unsigned transpose_pos_inv(unsigned p) {
    unsigned a = p % 16, rb = (p / 16) % 8, c = p / 128;
    return 64 * a + 8 * rev3(rb) + c;
}

void orc_transpose_i_u32(const uint32_t* in, uint32_t* out) {
    for (unsigned k = 0; k < 1024; ++k) out[transpose_pos(k)] = in[k];
}

void orc_transpose_o_u32(const uint32_t* in, uint32_t* out) {
    for (unsigned k = 0; k < 1024; ++k) out[k] = in[transpose_pos(k)];
}

/* ======================================================================= rsum
 * rsum.cuh:8-106 walks, for lane l, positions l + O_j with
 * O = 0,128,...,896, 64,...,960, 32,...,928, 96,...,992.                    */
unsigned orc_rsum_offset(unsigned j) {
    static const unsigned grp[4] = {0, 64, 32, 96};
    return 128 * (j % 8) + grp[j / 8];
}

void orc_unrsum_u32(const uint32_t* in, uint32_t* out) {
    for (unsigned l = 0; l < 32; ++l) {
        out[l + orc_rsum_offset(0)] = 0; /* unrsum.cpp:130 */
        for (unsigned j = 1; j < 32; ++j) {
            unsigned cur = l + orc_rsum_offset(j), prev = l + orc_rsum_offset(j - 1);
            out[cur] = in[cur] - in[prev];
        }
    }
}

void orc_rsum_u32(const uint32_t* in, uint32_t* out, const uint32_t* base) {
    for (unsigned l = 0; l < 32; ++l) {
        uint32_t acc = base[l];
        for (unsigned j = 0; j < 32; ++j) {
            unsigned p = l + orc_rsum_offset(j);
            acc += in[p];
            out[p] = acc;
        }
    }
}

/* ========================================================================= bw */
unsigned orc_bit_width(uint32_t x) {
    unsigned w = 0;
    while (x) {
        ++w;
        x >>= 1;
    }
    return w;
}

unsigned orc_find_bw4(const uint32_t* in, unsigned n) {
    uint32_t mx = 0;
    for (unsigned i = 0; i < n; ++i)
        if (in[i] > mx) mx = in[i];
    unsigned w = orc_bit_width(mx);
    return (w + 3) / 4 * 4; /* ssb_utils.h:226-236: 0 stays 0, else next multiple of 4 */
}

/* ====================================================================== delta */
unsigned orc_delta_encode_u32(const uint32_t* in, uint32_t* packed, uint32_t* bases) {
    uint32_t t[1024], d[1024];
    orc_transpose_i_u32(in, t);
    for (unsigned l = 0; l < 32; ++l) bases[l] = t[l];
    orc_unrsum_u32(t, d);
    uint32_t mx = 0;
    for (unsigned i = 0; i < 1024; ++i)
        if (d[i] > mx) mx = d[i];
    unsigned bw = orc_bit_width(mx);
    orc_pack_u32(d, packed, bw);
    return bw;
}

void orc_delta_decode_u32(const uint32_t* packed, unsigned bw, const uint32_t* bases, uint32_t* out_t) {
    uint32_t d[1024];
    orc_unpack_u32(packed, d, bw);
    orc_rsum_u32(d, out_t, bases);
}

/* ======================================================================== RLE */
uint32_t orc_linear_extract(const uint32_t* words, uint64_t k, unsigned bw) {
    uint32_t v = 0;
    for (unsigned b = 0; b < bw; ++b) {
        uint64_t p = k * bw + b;
        v |= ((words[p / 32] >> (p % 32)) & 1u) << b;
    }
    return v;
}

static void linear_put(uint32_t* words, uint64_t k, unsigned bw, uint32_t v) {
    for (unsigned b = 0; b < bw; ++b) {
        uint64_t p = k * bw + b;
        if ((v >> b) & 1u) words[p / 32] |= 1u << (p % 32);
    }
}

void orc_rle_encode_u32(const uint32_t* in, unsigned n, orc_rle_vec* hdr, uint32_t* values_out,
                        uint32_t* idx_out) {
    uint32_t full[1024], idx[1024], runv[1024];
    unsigned nr = 0;
    for (unsigned i = 0; i < 1024; ++i) full[i] = in[i < n ? i : (n ? n - 1 : 0)];
    if (n == 0)
        for (unsigned i = 0; i < 1024; ++i) full[i] = 0;
    for (unsigned i = 0; i < 1024; ++i) {
        if (i == 0 || full[i] != full[i - 1]) runv[nr++] = full[i];
        idx[i] = nr - 1;
    }
    uint32_t mn = runv[0], mx = runv[0];
    for (unsigned k = 0; k < nr; ++k) {
        if (runv[k] < mn) mn = runv[k];
        if (runv[k] > mx) mx = runv[k];
    }
    hdr->n_runs = nr;
    hdr->vbase = mn;
    hdr->vbw = orc_bit_width(mx - mn);
    memset(values_out, 0, sizeof(uint32_t) * ((nr * hdr->vbw + 31) / 32));
    for (unsigned k = 0; k < nr; ++k) linear_put(values_out, k, hdr->vbw, runv[k] - mn);
    hdr->ibw = orc_delta_encode_u32(idx, idx_out, hdr->bases);
}

void orc_rle_decode_u32(const orc_rle_vec* hdr, const uint32_t* values, const uint32_t* idx,
                        uint32_t* out_t) {
    uint32_t id_t[1024];
    orc_delta_decode_u32(idx, hdr->ibw, hdr->bases, id_t);
    for (unsigned p = 0; p < 1024; ++p) out_t[p] = hdr->vbase + orc_linear_extract(values, id_t[p], hdr->vbw);
}

/* ================================================================== checksum */
uint64_t orc_checksum_u32(const uint32_t* v, uint64_t n) {
    uint64_t s = 0;
    for (uint64_t i = 0; i < n; ++i) s += (i + 1) * (uint64_t)v[i];
    return s;
}

/* ======================================================================= RNG
 * splitmix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
 * number generators", OOPSLA 2014), keyed by (seed, stream) and counter i.   */
uint64_t orc_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_rand64(uint64_t seed, uint64_t stream, uint64_t i) {
    uint64_t key = orc_mix64(seed ^ (stream * 0x9E3779B97F4A7C15ULL));
    return orc_mix64(key + (i + 1) * 0x9E3779B97F4A7C15ULL);
}

/* uniform in [0, range), range <= 2^32: high 32 bits scaled by range */
uint32_t orc_uniform(uint64_t seed, uint64_t stream, uint64_t i, uint64_t range) {
    uint64_t hi = orc_rand64(seed, stream, i) >> 32;
    return (uint32_t)((hi * range) >> 32);
}

static int is_leap(unsigned y) { return (y % 4 == 0 && y % 100 != 0) || y % 400 == 0; }

uint32_t orc_yyyymmdd(uint32_t day) {
    static const unsigned mdays[12] = {31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    unsigned y = 1992;
    for (;;) {
        unsigned len = is_leap(y) ? 366 : 365;
        if (day < len) break;
        day -= len;
        ++y;
    }
    unsigned m = 0;
    for (;;) {
        unsigned len = mdays[m] + (m == 1 && is_leap(y) ? 1 : 0);
        if (day < len) break;
        day -= len;
        ++m;
    }
    return y * 10000 + (m + 1) * 100 + (day + 1);
}

/* Streams: one per synthetic column, fixed so every consumer agrees. */
enum {
    ST_C1 = 1,
    ST_QTY = 11, ST_DISC = 12, ST_PRICE = 13,
    ST_DFK = 21, ST_SFK = 22, ST_PFK = 23, ST_CFK = 24, ST_REV = 25,
    ST_SREG = 31, ST_SNAT = 32, ST_PCAT = 33, ST_PBRAND = 34, ST_CREG = 35, ST_CNAT = 36,
    ST_SCOST = 26, ST_PMFGR = 37
};

void orc_gen_c1(uint64_t seed, uint64_t n, unsigned bw, uint32_t* v) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) v[i] = orc_uniform(seed, ST_C1, (uint64_t)i, 1ULL << bw);
}

#define C2_DAYS 2557u /* 1992-01-01 .. 1998-12-31 */

void orc_gen_c2(uint64_t seed, uint64_t n, uint32_t* orderdate, uint32_t* quantity, uint32_t* discount,
                uint32_t* price) {
    uint32_t table[C2_DAYS];
    for (unsigned d = 0; d < C2_DAYS; ++d) table[d] = orc_yyyymmdd(d);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint64_t i = (uint64_t)ii;
        /* sorted and uniform over the days: row i falls on day floor(i*D/n) */
        orderdate[i] = table[(i * C2_DAYS) / n];
        quantity[i] = 1 + orc_uniform(seed, ST_QTY, i, 50);
        discount[i] = orc_uniform(seed, ST_DISC, i, 11);
        price[i] = 90000 + orc_uniform(seed, ST_PRICE, i, 10000000 - 90000);
    }
}

void orc_gen_c3_fact(uint64_t seed, uint64_t n, uint32_t* datefk, uint32_t* suppfk, uint32_t* partfk,
                     uint32_t* custfk, uint32_t* revenue) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint64_t i = (uint64_t)ii;
        datefk[i] = orc_uniform(seed, ST_DFK, i, 2556);
        suppfk[i] = orc_uniform(seed, ST_SFK, i, 40000);
        partfk[i] = orc_uniform(seed, ST_PFK, i, 800000);
        if (custfk) custfk[i] = orc_uniform(seed, ST_CFK, i, 600000);
        revenue[i] = orc_uniform(seed, ST_REV, i, 10000000);
    }
}

void orc_gen_c3_dims(uint64_t seed, uint32_t n_date, uint32_t n_supp, uint32_t n_part, uint32_t n_cust,
                     uint32_t* d_year, uint32_t* s_region, uint32_t* s_nation, uint32_t* p_category,
                     uint32_t* p_brand, uint32_t* c_region, uint32_t* c_nation) {
    for (uint32_t d = 0; d < n_date; ++d) d_year[d] = orc_yyyymmdd(d) / 10000;
    /* SSB's hierarchies: a nation lies in one region (5 nations each), a brand
     * in one category (40 brands each, brand1 = category x 40 + 1..40) */
    for (uint32_t s = 0; s < n_supp; ++s) {
        s_nation[s] = 1 + orc_uniform(seed, ST_SNAT, s, 25);
        s_region[s] = (s_nation[s] - 1) / 5 + 1;
    }
    for (uint32_t p = 0; p < n_part; ++p) {
        p_category[p] = 1 + orc_uniform(seed, ST_PCAT, p, 25);
        p_brand[p] = (p_category[p] - 1) * 40 + 1 + orc_uniform(seed, ST_PBRAND, p, 40);
    }
    for (uint32_t c = 0; c < n_cust; ++c) {
        c_nation[c] = 1 + orc_uniform(seed, ST_CNAT, c, 25);
        c_region[c] = (c_nation[c] - 1) / 5 + 1;
    }
}

void orc_gen_star_extra(uint64_t seed, uint64_t n, uint32_t* supplycost, uint32_t n_part, uint32_t* p_mfgr) {
    for (uint64_t i = 0; i < n; ++i) supplycost[i] = 60000u + orc_uniform(seed, ST_SCOST, i, 60000);
    for (uint32_t p = 0; p < n_part; ++p) p_mfgr[p] = 1u + orc_uniform(seed, ST_PMFGR, p, 5);
}

/* ==================================================================== queries */
int64_t orc_q11(const uint32_t* orderdate, const uint32_t* quantity, const uint32_t* discount,
                const uint32_t* price, uint64_t n, uint32_t date_gt, uint32_t date_lt, uint32_t qty_lt,
                uint32_t disc_lo, uint32_t disc_hi, uint64_t* count) {
    int64_t sum = 0;
    uint64_t cnt = 0;
#pragma omp parallel for reduction(+ : sum, cnt) schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        uint64_t i = (uint64_t)ii;
        if (orderdate[i] > date_gt && orderdate[i] < date_lt && quantity[i] < qty_lt && discount[i] >= disc_lo &&
            discount[i] <= disc_hi) {
            sum += (int64_t)price[i] * (int64_t)discount[i];
            cnt += 1;
        }
    }
    if (count) *count = cnt;
    return sum;
}

#define Q21_SLOTS 7000u

void orc_q21(const uint32_t* datefk, const uint32_t* suppfk, const uint32_t* partfk, const uint32_t* revenue,
             uint64_t n, const uint32_t* d_year, const uint32_t* s_region, const uint32_t* p_category,
             const uint32_t* p_brand, uint32_t cat, uint32_t reg, int64_t* sums, uint64_t* counts) {
    memset(sums, 0, sizeof(int64_t) * Q21_SLOTS);
    memset(counts, 0, sizeof(uint64_t) * Q21_SLOTS);
    for (uint64_t i = 0; i < n; ++i) {
        if (s_region[suppfk[i]] != reg) continue;
        if (p_category[partfk[i]] != cat) continue;
        uint32_t year = d_year[datefk[i]];
        uint32_t brand = p_brand[partfk[i]];
        uint32_t slot = (brand * 7 + (year - 1992)) % Q21_SLOTS; /* fls_q21.cu:71 */
        sums[slot] += (int64_t)revenue[i];
        counts[slot] += 1;
    }
}

/* Q3.1 shape, fls_q31.cu:28-94 (probe_v1) and 300-372 (builds): suppliers of
 * s_region == reg_s give s_nation, customers of c_region == reg_c give
 * c_nation, dates with year_lo <= d_year <= year_hi give the year; slot
 * (s_nation*25*7 + c_nation*7 + year - 1992) % 4375 (fls_q31.cu:84).
 * Foreign keys are 0-based row indices of the dimension columns.          */
void orc_q31(const uint32_t* datefk, const uint32_t* custfk, const uint32_t* suppfk, const uint32_t* revenue,
             uint64_t n, const uint32_t* d_year, const uint32_t* s_region, const uint32_t* s_nation,
             const uint32_t* c_region, const uint32_t* c_nation, uint32_t reg_c, uint32_t reg_s, uint32_t year_lo,
             uint32_t year_hi, int64_t* sums, uint64_t* counts) {
    memset(sums, 0, sizeof(int64_t) * 4375);
    memset(counts, 0, sizeof(uint64_t) * 4375);
    for (uint64_t i = 0; i < n; ++i) {
        if (s_region[suppfk[i]] != reg_s) continue;
        if (c_region[custfk[i]] != reg_c) continue;
        const uint32_t year = d_year[datefk[i]];
        if (year < year_lo || year > year_hi) continue;
        const uint32_t slot = (s_nation[suppfk[i]] * 25 * 7 + c_nation[custfk[i]] * 7 + (year - 1992)) % 4375;
        sums[slot] += (int64_t)revenue[i];
        counts[slot] += 1;
    }
}

/* Q4.1 shape, fls_q41.cu:27-107 (probe_v1) and 186-271 (builds): suppliers of
 * s_region == reg_s and parts with mfgr_lo <= p_mfgr <= mfgr_hi pass, customers
 * of c_region == reg_c give c_nation, every date gives its year; slot
 * (c_nation*7 + year - 1992) % 175 (fls_q41.cu:97), sum of revenue - supplycost. */
void orc_q41(const uint32_t* datefk, const uint32_t* custfk, const uint32_t* suppfk, const uint32_t* partfk,
             const uint32_t* revenue, const uint32_t* supplycost, uint64_t n, const uint32_t* d_year,
             const uint32_t* s_region, const uint32_t* c_region, const uint32_t* c_nation, const uint32_t* p_mfgr,
             uint32_t reg_c, uint32_t reg_s, uint32_t mfgr_lo, uint32_t mfgr_hi, int64_t* sums, uint64_t* counts) {
    memset(sums, 0, sizeof(int64_t) * 175);
    memset(counts, 0, sizeof(uint64_t) * 175);
    for (uint64_t i = 0; i < n; ++i) {
        if (s_region[suppfk[i]] != reg_s) continue;
        if (c_region[custfk[i]] != reg_c) continue;
        if (p_mfgr[partfk[i]] < mfgr_lo || p_mfgr[partfk[i]] > mfgr_hi) continue;
        const uint32_t slot = (c_nation[custfk[i]] * 7 + (d_year[datefk[i]] - 1992)) % 175;
        sums[slot] += (int64_t)revenue[i] - (int64_t)supplycost[i];
        counts[slot] += 1;
    }
}

/* ================================================================ CPU baseline */
void orc_unpack_lanewise_u32(const uint32_t* in, uint32_t* out, unsigned bw) {
    const uint32_t mask = bw >= 32 ? 0xffffffffu : ((1u << bw) - 1u);
    for (unsigned l = 0; l < 32; ++l) {
        for (unsigned r = 0; r < 32; ++r) {
            if (bw == 0) {
                out[32 * r + l] = 0;
                continue;
            }
            unsigned bit = r * bw, w = bit / 32, s = bit % 32;
            uint32_t x = in[32 * w + l] >> s;
            if (s + bw > 32) x |= in[32 * (w + 1) + l] << (32 - s);
            out[32 * r + l] = x & mask;
        }
    }
}

void orc_decode_vector(const orc_col* c, uint64_t v, uint32_t* out) {
    if (c->encoding == 1) {
        uint64_t m = c->vmeta[v];
        uint32_t d[1024];
        orc_unpack_lanewise_u32(c->words + (m >> 8), d, (unsigned)(m & 0xff));
        orc_rsum_u32(d, out, c->bases + 32 * v);
    } else if (c->encoding == 2) {
        /* run index: delta-coded like a delta vector; run values after it */
        const uint64_t m = c->vmeta[v], a = c->aux[v];
        const unsigned ibw = (unsigned)(m & 0xff), vbw = (unsigned)((a >> 48) & 0xff);
        const uint32_t vbase = (uint32_t)a, n_runs = (uint32_t)((a >> 32) & 0xffff);
        const uint32_t* vals = c->words + (m >> 8) + 32u * ibw;
        const uint32_t mask = vbw >= 32 ? 0xffffffffu : ((1u << vbw) - 1u);
        uint32_t d[1024], idx[1024], run[1024];
        for (uint32_t k = 0; k < n_runs && k < 1024; ++k) { /* word-window extraction */
            const uint64_t bit = (uint64_t)k * vbw;
            const uint32_t w = (uint32_t)(bit / 32), sh = (uint32_t)(bit % 32);
            uint32_t x = vbw ? vals[w] >> sh : 0;
            if (vbw && sh + vbw > 32) x |= vals[w + 1] << (32 - sh);
            run[k] = vbase + (x & mask);
        }
        orc_unpack_lanewise_u32(c->words + (m >> 8), d, ibw);
        orc_rsum_u32(d, idx, c->bases + 32 * v);
        for (unsigned i = 0; i < 1024; ++i) out[i] = run[idx[i] < n_runs ? idx[i] : n_runs - 1u];
    } else {
        orc_unpack_lanewise_u32(c->words + v * 32 * c->bw, out, c->bw);
        if (c->base)
            for (unsigned i = 0; i < 1024; ++i) out[i] += c->base;
    }
}

int64_t orc_q11_encoded(const orc_col* od, const orc_col* q, const orc_col* d, const orc_col* p, uint64_t n_rows,
                        uint64_t v_begin, uint64_t v_end, uint32_t date_gt, uint32_t date_lt, uint32_t qty_lt,
                        uint32_t disc_lo, uint32_t disc_hi, int n_threads, uint64_t* count) {
    int64_t sum = 0;
    uint64_t cnt = 0;
    const int tr = p->transposed != 0;
#pragma omp parallel for num_threads(n_threads) reduction(+ : sum, cnt) schedule(static)
    for (int64_t vv = (int64_t)v_begin; vv < (int64_t)v_end; ++vv) {
        uint64_t v = (uint64_t)vv;
        uint32_t a[1024], b[1024], c[1024], e[1024];
        orc_decode_vector(od, v, a);
        orc_decode_vector(q, v, b);
        orc_decode_vector(d, v, c);
        orc_decode_vector(p, v, e);
        for (unsigned k = 0; k < 1024; ++k) {
            uint64_t row = v * 1024 + (tr ? (uint64_t)transpose_pos_inv(k) : k);
            if (row >= n_rows) continue;
            if (a[k] > date_gt && a[k] < date_lt && b[k] < qty_lt && c[k] >= disc_lo && c[k] <= disc_hi) {
                sum += (int64_t)e[k] * (int64_t)c[k];
                cnt += 1;
            }
        }
    }
    if (count) *count = cnt;
    return sum;
}

/* =========================================================== batched segments */
enum { ST_SEG_PARAMS = 1000, ST_SEG_VALUES = 100000, ST_SEG_RUNS = 200000 };

void orc_c4_params(uint64_t seed, uint32_t s, uint32_t lmin, uint32_t lmax, uint32_t* enc, uint32_t* bw,
                   uint64_t* len) {
    uint32_t e = lmin + orc_uniform(seed, ST_SEG_PARAMS + s, 0, lmax - lmin);
    *len = (1ULL << e) + orc_uniform(seed, ST_SEG_PARAMS + s, 1, 1ULL << e);
    *bw = 1 + orc_uniform(seed, ST_SEG_PARAMS + s, 2, 32);
    uint32_t r = orc_uniform(seed, ST_SEG_PARAMS + s, 3, 10);
    *enc = r < 4 ? 0 : (r < 7 ? 1 : 2); /* bitpack, delta, RLE */
}

void orc_c4_values(uint64_t seed, uint32_t s, uint32_t enc, uint32_t bw, uint64_t len, uint32_t* out) {
    if (enc == 0) {
        for (uint64_t i = 0; i < len; ++i) out[i] = orc_uniform(seed, ST_SEG_VALUES + s, i, 1ULL << bw);
    } else if (enc == 1) {
        /* sorted: i*m + U[0,m) with m = 2^(b-1), b = min(bw, 32 - bitwidth(len)) */
        unsigned lw = 0;
        for (uint64_t x = len; x; x >>= 1) ++lw;
        unsigned b = lw < 32 ? 32 - lw : 1;
        if (b < 1) b = 1;
        if (bw < b) b = bw;
        uint64_t m = 1ULL << (b - 1);
        for (uint64_t i = 0; i < len; ++i) out[i] = (uint32_t)(i * m) + orc_uniform(seed, ST_SEG_VALUES + s, i, m);
    } else {
        uint32_t cur = 0;
        for (uint64_t i = 0; i < len; ++i) {
            if (i == 0 || orc_uniform(seed, ST_SEG_RUNS + s, i, 8) == 0)
                cur = orc_uniform(seed, ST_SEG_VALUES + s, i, 1ULL << bw);
            out[i] = cur;
        }
    }
}

void orc_c4_counts(uint64_t seed, uint32_t n_seg, uint32_t lmin, uint32_t lmax, uint64_t* counts, uint32_t* lo,
                   uint32_t* hi, int n_threads) {
#pragma omp parallel for num_threads(n_threads) schedule(dynamic, 1)
    for (int64_t ss = 0; ss < (int64_t)n_seg; ++ss) {
        uint32_t s = (uint32_t)ss, enc, bw;
        uint64_t len;
        orc_c4_params(seed, s, lmin, lmax, &enc, &bw, &len);
        uint32_t* v = (uint32_t*)malloc(sizeof(uint32_t) * len);
        orc_c4_values(seed, s, enc, bw, len, v);
        uint32_t mn = 0xffffffffu, mx = 0;
        for (uint64_t i = 0; i < len; ++i) {
            if (v[i] < mn) mn = v[i];
            if (v[i] > mx) mx = v[i];
        }
        uint64_t range = (uint64_t)(mx - mn);
        uint32_t l = mn + (uint32_t)(range / 4), h = mn + (uint32_t)(range * 3 / 4);
        uint64_t c = 0;
        for (uint64_t i = 0; i < len; ++i) c += (v[i] >= l && v[i] <= h);
        counts[s] = c;
        lo[s] = l;
        hi[s] = h;
        free(v);
    }
}
