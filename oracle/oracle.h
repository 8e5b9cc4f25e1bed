/*
 * CPU oracle for the FastLanes decode + Crystal query hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product library (fastlanesgpu_amd/libfls.so) never
 * links, loads or calls it.
 *
 * Every function restates one piece of the reference's algorithm in plain,
 * deliberately naive C (bit-by-bit packing, closed-form permutations, scalar
 * query loops).  The citation on each declaration names the reference code it
 * follows (paths relative to the reference repository root).
 *
 * Pinning status (what independently checks each function):
 *   orc_pack_u32 / orc_unpack_u32      pinned: golden helper::rand_arr_* vectors
 *                                      packed by the reference's own pack.cpp
 *                                      (tests/golden), oracle/_ref build, bw 0/32
 *                                      special cases, round trip.
 *   orc_transpose_{i,o}_u32            pinned: oracle/_ref transpose.cpp build,
 *                                      committed 1024-entry permutation fixture.
 *   orc_unrsum_u32 / orc_rsum_u32      pinned: oracle/_ref unrsum.cpp build,
 *                                      committed unrsum fixtures, rsum∘unrsum = id.
 *   orc_delta_*                        pinned through the three above.
 *   orc_find_bw4 / orc_bit_width       pinned: closed form + brute force.
 *   orc_rle_*                          format defined by this project (the
 *                                      reference's RLE is a CPU sketch):
 *                                      round trip + brute-force run counting.
 *   orc_q11 / orc_q21 / orc_q31 / orc_q41
 *                                      pinned: textbook semantics checked against
 *                                      numpy brute force on small tables.
 *   orc_gen_* / orc_rand*              definitions of the synthetic inputs.
 *   orc_mix_* / orc_exhaustive_class   the encoders above run literally; sizes
 *                                      are their word counts (brute force).
 *   orc_vector_features                definition of the selector's input;
 *                                      closed-form cases in the tests.
 *   orc_mlp_*                          pinned: finite differences in double.
 */
#ifndef FLS_ORACLE_H
#define FLS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- bit layout */
/* fastlanes/src/pack.cpp:2970-3123 (pack_11bit_32ow), dispatch pack.cpp:29607-29709 */
void orc_pack_u32(const uint32_t* in, uint32_t* out, unsigned bw);
/* fastlanes/src/include/fls_gen/unpack/unpack.cuh:877-971 (per-lane decode), :3330-3432 */
void orc_unpack_u32(const uint32_t* in, uint32_t* out, unsigned bw);

/* fastlanes/src/transpose.cpp:4110-5135 (transpose_i u32), :5136-6161 (transpose_o u32) */
void orc_transpose_i_u32(const uint32_t* in, uint32_t* out);
void orc_transpose_o_u32(const uint32_t* in, uint32_t* out);
/* fastlanes/src/unrsum.cpp:120-256 */
void orc_unrsum_u32(const uint32_t* in, uint32_t* out);
/* fastlanes/src/include/fls_gen/rsum/rsum.cuh:3-107 */
void orc_rsum_u32(const uint32_t* in, uint32_t* out, const uint32_t* base);
/* position order of the rsum chain (rsum.cuh:8-106): returns O_j for j in [0,32) */
unsigned orc_rsum_offset(unsigned j);

/* ---------------------------------------------------------------- FOR / bw */
unsigned orc_bit_width(uint32_t x);
/* fastlanes/src/include/ssb_utils.h:222-239 (find_bw: round up to a multiple of 4) */
unsigned orc_find_bw4(const uint32_t* in, unsigned n);

/* ---------------------------------------------------------------- delta (a8)
 * Encode one 1024-value vector (original order) following
 * fastlanes/example/fastlanes_bench_delta.cu:143-155: transpose_i, bases = the
 * first 32 transposed values, unrsum, pack at the smallest bw.  Returns bw.
 * `packed` needs room for 32*32 words.                                        */
unsigned orc_delta_encode_u32(const uint32_t* in, uint32_t* packed, uint32_t* bases);
/* unpack + rsum; output is in transposed order (fastlanes_bench_delta.cu:183-191) */
void orc_delta_decode_u32(const uint32_t* packed, unsigned bw, const uint32_t* bases, uint32_t* out_t);

/* ---------------------------------------------------------------- RLE (a9)
 * Project-defined FastLanes RLE of one vector of n <= 1024 values (original
 * order; rows >= n are padding that repeats the last value):
 *   run values (FOR base + linear LSB-first bit-packing at width vbw),
 *   run index idx[i] delta-encoded exactly like orc_delta_encode_u32.
 * Decoding yields transposed order, like delta.                               */
typedef struct orc_rle_vec {
    uint32_t n_runs;
    uint32_t vbase;      /* min run value */
    uint32_t vbw;        /* width of (run value - vbase) */
    uint32_t ibw;        /* width of the run-index deltas: 0 or 1 */
    uint32_t bases[32];  /* run index at each lane's chain start */
} orc_rle_vec;
/* values_out needs ceil(1024*32/32) words, idx_out 32*32 words. */
void orc_rle_encode_u32(const uint32_t* in, unsigned n, orc_rle_vec* hdr, uint32_t* values_out,
                        uint32_t* idx_out);
void orc_rle_decode_u32(const orc_rle_vec* hdr, const uint32_t* values, const uint32_t* idx,
                        uint32_t* out_t);
uint32_t orc_linear_extract(const uint32_t* words, uint64_t k, unsigned bw);

/* ---------------------------------------------------------------- checksums */
/* sum over i of (i+1)*v[i] modulo 2^64 (order sensitive) */
uint64_t orc_checksum_u32(const uint32_t* v, uint64_t n);

/* ---------------------------------------------------------------- RNG + data */
uint64_t orc_mix64(uint64_t z);
uint64_t orc_rand64(uint64_t seed, uint64_t stream, uint64_t i);
uint32_t orc_uniform(uint64_t seed, uint64_t stream, uint64_t i, uint64_t range);
/* days since 1992-01-01 -> yyyymmdd */
uint32_t orc_yyyymmdd(uint32_t day);

/* fused filter+sum fact table (config C2 shape), n rows */
void orc_gen_c1(uint64_t seed, uint64_t n, unsigned bw, uint32_t* v);
void orc_gen_c2(uint64_t seed, uint64_t n, uint32_t* orderdate, uint32_t* quantity, uint32_t* discount,
                uint32_t* price);
/* probe+group-by tables (config C3 shape); fks are 0-based dense keys */
void orc_gen_c3_fact(uint64_t seed, uint64_t n, uint32_t* datefk, uint32_t* suppfk, uint32_t* partfk,
                     uint32_t* custfk, uint32_t* revenue);
void orc_gen_c3_dims(uint64_t seed, uint32_t n_date, uint32_t n_supp, uint32_t n_part, uint32_t n_cust,
                     uint32_t* d_year, uint32_t* s_region, uint32_t* s_nation, uint32_t* p_category,
                     uint32_t* p_brand, uint32_t* c_region, uint32_t* c_nation);

/* ---------------------------------------------------------------- queries
 * Scalar Q1.1 shape, fastlanes/src/ssb/fls_q11.cu:29-81: rows with
 * date_gt < orderdate < date_lt, quantity < qty_lt, disc_lo <= discount <= disc_hi;
 * returns sum(price*discount) as int64 and the qualifying row count.           */
int64_t orc_q11(const uint32_t* orderdate, const uint32_t* quantity, const uint32_t* discount,
                const uint32_t* price, uint64_t n, uint32_t date_gt, uint32_t date_lt, uint32_t qty_lt,
                uint32_t disc_lo, uint32_t disc_hi, uint64_t* count);
/* Scalar Q2.1 shape, fastlanes/src/ssb/fls_q21.cu:21-78,197-263: keep rows whose
 * part has category==cat and supplier has region==reg; group slot
 * (brand*7 + year-1992) % 7000; sums[7000] int64 and counts[7000].            */
void orc_q21(const uint32_t* datefk, const uint32_t* suppfk, const uint32_t* partfk, const uint32_t* revenue,
             uint64_t n, const uint32_t* d_year, const uint32_t* s_region, const uint32_t* p_category,
             const uint32_t* p_brand, uint32_t cat, uint32_t reg, int64_t* sums, uint64_t* counts);

/* Extra columns of the Q3.1 / Q4.1 shapes: lo_supplycost U[60000, 120000),
 * p_mfgr 1 + U{0..4}.                                                          */
void orc_gen_star_extra(uint64_t seed, uint64_t n, uint32_t* supplycost, uint32_t n_part, uint32_t* p_mfgr);
/* Scalar Q3.1 shape (fls_q31.cu:28-94,300-372): sums/counts[4375].           */
void orc_q31(const uint32_t* datefk, const uint32_t* custfk, const uint32_t* suppfk, const uint32_t* revenue,
             uint64_t n, const uint32_t* d_year, const uint32_t* s_region, const uint32_t* s_nation,
             const uint32_t* c_region, const uint32_t* c_nation, uint32_t reg_c, uint32_t reg_s, uint32_t year_lo,
             uint32_t year_hi, int64_t* sums, uint64_t* counts);
/* Scalar Q4.1 shape (fls_q41.cu:27-107,186-271): sums/counts[175] of
 * revenue - supplycost.                                                       */
void orc_q41(const uint32_t* datefk, const uint32_t* custfk, const uint32_t* suppfk, const uint32_t* partfk,
             const uint32_t* revenue, const uint32_t* supplycost, uint64_t n, const uint32_t* d_year,
             const uint32_t* s_region, const uint32_t* c_region, const uint32_t* c_nation, const uint32_t* p_mfgr,
             uint32_t reg_c, uint32_t reg_s, uint32_t mfgr_lo, uint32_t mfgr_hi, int64_t* sums, uint64_t* counts);

/* ---------------------------------------------------------------- batched segments
 * Definition of the batched-segments workload (the GPU builder generates the
 * same values from the same seed): parameters of segment s, its values (RLE
 * walked sequentially here), and per-segment counts of values inside the
 * middle half of the segment's [min, max].                                   */
void orc_c4_params(uint64_t seed, uint32_t s, uint32_t lmin, uint32_t lmax, uint32_t* enc, uint32_t* bw,
                   uint64_t* len);
void orc_c4_values(uint64_t seed, uint32_t s, uint32_t enc, uint32_t bw, uint64_t len, uint32_t* out);
void orc_c4_counts(uint64_t seed, uint32_t n_seg, uint32_t lmin, uint32_t lmax, uint64_t* counts, uint32_t* lo,
                   uint32_t* hi, int n_threads);

/* ---------------------------------------------------------------- CPU baseline
 * Decode + query over ENCODED columns, the CPU counterpart of the fused GPU
 * kernels.  Per-lane scalar extraction in the style of the generated decoder
 * (unpack.cuh:877-971: a word window, shift and mask per row), FastLanes rsum
 * for delta, scalar predicate/aggregate loop; OpenMP over vector ranges.
 * Pinned against the bit-by-bit functions above.                             */
typedef struct orc_col {
    uint32_t encoding;   /* 0 bitpack(+FOR base), 1 delta, 2 RLE */
    uint32_t bw;
    uint32_t base;
    uint32_t transposed;
    const uint32_t* words;
    const uint64_t* vmeta; /* (offset << 8) | bw (RLE: the run-index delta width) */
    const uint32_t* bases;
    const uint64_t* aux;   /* RLE: vbase | n_runs << 32 | vbw << 48 | ibw << 56 per vector */
} orc_col;

void orc_unpack_lanewise_u32(const uint32_t* in, uint32_t* out, unsigned bw);
void orc_decode_vector(const orc_col* c, uint64_t v, uint32_t* out);
int64_t orc_q11_encoded(const orc_col* od, const orc_col* q, const orc_col* d, const orc_col* p, uint64_t n_rows,
                        uint64_t v_begin, uint64_t v_end, uint32_t date_gt, uint32_t date_lt, uint32_t qty_lt,
                        uint32_t disc_lo, uint32_t disc_hi, int n_threads, uint64_t* count);

/* Same-run CPU baselines of the other workloads (baseline.c), each the
 * oracle's scalar decode (orc_decode_vector) followed by the scalar query of
 * the matching plain function above, OpenMP over vector (or segment) ranges:
 *   orc_c1_encoded    decode fixed-width vectors to out[] + orc_checksum_u32;
 *   orc_q21_encoded   orc_q21 over four encoded columns (dense keys
 *                     keys_min + i; keys outside the dimension do not match);
 *   orc_q31_encoded   orc_q31, orc_q41_encoded orc_q41 likewise;
 *   orc_c4_encoded    per-segment count of lo <= value <= hi over encoded
 *                     segments (the batched-segments query);
 *   orc_exhaustive_classes  orc_exhaustive_class per vector (the
 *                     try-all-encodings search the learned selector replaces). */
uint64_t orc_c1_encoded(const uint32_t* words, uint64_t n_values, unsigned bw, uint32_t* out, int n_threads);
void orc_q21_encoded(const orc_col* date, const orc_col* supp, const orc_col* part, const orc_col* rev,
                     uint64_t n_rows, uint64_t v_begin, uint64_t v_end, const uint32_t* d_year,
                     const uint32_t* s_region, const uint32_t* p_category, const uint32_t* p_brand, uint32_t n_date,
                     uint32_t n_supp, uint32_t n_part, uint32_t keys_min, uint32_t cat, uint32_t reg, int n_threads,
                     int64_t* sums, uint64_t* counts);
void orc_q31_encoded(const orc_col* date, const orc_col* cust, const orc_col* supp, const orc_col* rev,
                     uint64_t n_rows, uint64_t v_begin, uint64_t v_end, const uint32_t* d_year,
                     const uint32_t* s_region, const uint32_t* s_nation, const uint32_t* c_region,
                     const uint32_t* c_nation, uint32_t n_date, uint32_t n_supp, uint32_t n_cust, uint32_t keys_min,
                     uint32_t reg_c, uint32_t reg_s, uint32_t year_lo, uint32_t year_hi, int n_threads,
                     int64_t* sums, uint64_t* counts);
void orc_q41_encoded(const orc_col* date, const orc_col* cust, const orc_col* supp, const orc_col* part,
                     const orc_col* rev, const orc_col* cost, uint64_t n_rows, uint64_t v_begin, uint64_t v_end,
                     const uint32_t* d_year, const uint32_t* s_region, const uint32_t* c_region,
                     const uint32_t* c_nation, const uint32_t* p_mfgr, uint32_t n_date, uint32_t n_supp,
                     uint32_t n_cust, uint32_t n_part, uint32_t keys_min, uint32_t reg_c, uint32_t reg_s,
                     uint32_t mfgr_lo, uint32_t mfgr_hi, int n_threads, int64_t* sums, uint64_t* counts);
void orc_c4_encoded(const orc_col* segs, const uint64_t* n_values, const uint32_t* lo, const uint32_t* hi,
                    uint32_t n_seg, int n_threads, uint64_t* counts);
void orc_exhaustive_classes(const uint32_t* in, uint64_t n_vec, uint8_t* cls, int n_threads);

/* ---------------------------------------------------------------- selector
 * (oracle/selector.c; no reference counterpart.)  Mixed-vector encodings:
 * 0 bit-pack, 1 FOR, 2 FastLanes delta, 3 FastLanes RLE, each over one
 * 1024-value vector in original order, payload in transposed row order.
 * orc_mix_encode_u32 returns the payload words (words: up to 1088).
 * Class = 10*encoding + width bucket (0 | 1-2 | 3-4 | 5-6 | 7-8 | 9-12 |
 * 13-16 | 17-20 | 21-24 | 25-32).                                             */
unsigned orc_width_bucket(unsigned w);
unsigned orc_mix_encode_u32(const uint32_t* in, unsigned enc, uint32_t* words, unsigned* width, uint64_t* aux);
void orc_mix_sizes(const uint32_t* in, uint32_t* sizes4, unsigned* widths4);
unsigned orc_exhaustive_class(const uint32_t* in);
void orc_vector_features(const uint32_t* in, float* f);
void orc_gen_train_vector(uint64_t seed, uint64_t v, const uint32_t* zipf_cdf, uint32_t* out);
/* MLP 16->64->64->40 in double: parameters as fls_mlp_* (7848 values).
 * round_bf16 != 0 rounds inputs, weights, hidden activations and the
 * back-propagated dZ/dH2/dH1 to BF16 where the GPU kernels do.
 * orc_mlp_grad returns the summed loss and writes d(mean loss)/dw when
 * inv_batch = 1/n.                                                           */
unsigned orc_mlp_param_count(void);
double orc_bf16(double x);
void orc_mlp_forward(const double* w, const float* feat, uint64_t n, int round_bf16, double* logits);
double orc_mlp_grad(const double* w, const float* feat, const uint8_t* label, uint64_t n, double inv_batch,
                    int round_bf16, double* grad);

#ifdef __cplusplus
}
#endif
/* Tile-based competitor formats (tile.c): GPU-FOR, GPU-DFOR, GPU-RFOR.  Output
 * buffers hold orc_tile_max_words(n) words; starts hold ceil(n/128) + 1 (FOR,
 * DFOR: 4 per 512-value tile + 1) or ceil(n/512) + 1 (RFOR) entries.         */
uint64_t orc_tile_max_words(uint64_t n);
uint64_t orc_tile_for_encode(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts);
void orc_tile_for_decode(const uint32_t* data, const uint32_t* starts, uint64_t n, uint32_t* out);
uint64_t orc_tile_dfor_encode(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts);
void orc_tile_dfor_decode(const uint32_t* data, const uint32_t* starts, uint64_t n, uint32_t* out);
void orc_tile_rfor_encode(const uint32_t* in, uint64_t n, uint32_t* values, uint32_t* val_starts,
                          uint32_t* run_lengths, uint32_t* rl_starts, uint64_t* n_val_words, uint64_t* n_rl_words);
void orc_tile_rfor_decode(const uint32_t* values, const uint32_t* val_starts, const uint32_t* run_lengths,
                          const uint32_t* rl_starts, uint64_t n, uint32_t* out);

#endif
