/*
 * fls.h -- C ABI of the MI355X (gfx950) FastLanes decode + query library.
 *
 * Every entry point returns an int status (FLS_OK = 0, negative on error) and
 * never exits the process.  Device pointers are caller-owned and must stay
 * valid until the work queued on `stream` completes; `stream` is a hipStream_t
 * passed as void* (NULL = the null stream).  Kernels never allocate, except
 * where a function says it owns a workspace.  Calls on one stream are ordered;
 * calls on different streams that write the same output race.
 *
 * Data unit: the FastLanes vector of 1024 uint32 values, 32 lanes x 32 rows,
 * value 32*r + l is row r of lane l; lane l's rows form an LSB-first bit
 * stream whose word w sits at packed[32*w + l]  (fastlanes/src/pack.cpp:2970,
 * fastlanes/src/include/fls_gen/unpack/unpack.cuh:877-971).  A vector packed
 * at bit width bw occupies 32*bw words; vectors of a fixed-width column are
 * contiguous at v*32*bw.
 *
 * Each declaration cites the reference interface it replaces (paths relative
 * to the reference repository root).
 */
#ifndef FLS_FLS_H
#define FLS_FLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
enum {
    FLS_OK = 0,
    FLS_ERR_BW = -1,        /* bit width outside 0..32 */
    FLS_ERR_SIZE = -2,      /* count or buffer size out of range */
    FLS_ERR_NULL = -3,      /* required pointer is NULL */
    FLS_ERR_HIP = -4,       /* a HIP runtime call failed (fls_last_hip_error) */
    FLS_ERR_ENCODING = -5,  /* unknown or unsupported encoding */
    FLS_ERR_ARG = -6,       /* other invalid argument */
    FLS_ERR_CAPACITY = -7,  /* output buffer too small */
    FLS_ERR_NO_DEVICE = -8, /* no usable gfx950 device */
    FLS_ERR_FORMAT = -9,    /* column image or metadata: bad magic, version, sizes, checksum or vmeta */
    FLS_ERR_SCRATCH = -10   /* the work needs more than the library's static scratch array holds */
};

#define FLS_VECTOR_SIZE 1024u
#define FLS_LANES 32u

/* Human-readable text for a status code. */
const char* fls_status_string(int status);
/* Last HIP error code seen by the library on this thread (0 if none). */
int fls_last_hip_error(void);
/* Library version, e.g. "fls-mi355x 0.1". */
const char* fls_version(void);
/* Plain streaming copy of n_bytes (multiple of 16, 16-byte aligned): the
 * calibration kernel for bandwidth and HBM-counter measurements.  No
 * counterpart in the reference.                                              */
int fls_copy_bytes(const void* d_src, void* d_dst, uint64_t n_bytes, void* stream);
/* Read-only streaming pass: *d_sum = sum of the 32-bit words mod 2^64.  The
 * achievable read bandwidth the decode kernels are compared with (16-byte
 * aligned source, n_bytes a multiple of 16).  No reference counterpart.     */
int fls_read_bytes(const void* d_src, uint64_t n_bytes, uint64_t* d_sum, void* stream);

/* ========================================================= host encoders (CPU)
 * One 1024-value vector.  Replaces
 * generated::pack::fallback::scalar::pack(const uint32_t*, uint32_t*, uint8_t)
 * (fastlanes/src/include/fls_gen/pack/pack.hpp:8-23; bodies pack.cpp:29607).
 * Values are masked to bw bits; writes 32*bw words.                          */
int fls_pack_u32(const uint32_t* in, uint32_t* out, uint32_t bw);
/* generated::transpose::fallback::scalar::transpose_i (transpose.hpp:8-20,
 * transpose.cpp:4110).                                                        */
int fls_transpose_i_u32(const uint32_t* in, uint32_t* out);
/* Inverse of fls_transpose_i_u32.  The reference's transpose_o
 * (transpose.cpp:5136) repeats transpose_i's permutation; this one is the
 * untranspose its name promises.                                              */
int fls_transpose_o_u32(const uint32_t* in, uint32_t* out);
/* generated::unrsum::fallback::scalar::unrsum (unrsum.hpp:8, unrsum.cpp:120). */
int fls_unrsum_u32(const uint32_t* in, uint32_t* out);

/* Column-level host encoders.  A column of n values becomes ceil(n/1024)
 * vectors; the tail of the last vector is padded (0 for bit-packing, the last
 * value for delta/RLE).
 * FLS_ENC_FLAG_TRANSPOSE applies transpose_i to each vector before packing,
 * which puts a bit-packed column in the same row order as a delta- or
 * RLE-decoded one (FastLanes' unified transposed layout).                     */
#define FLS_ENC_FLAG_TRANSPOSE 1u
uint64_t fls_packed_words(uint64_t n, uint32_t bw);
int fls_pack_column_u32(const uint32_t* in, uint64_t n, uint32_t bw, uint32_t flags, uint32_t* out);

/* Per-vector metadata word: bits 0..7 bit width, bits 8..63 word offset of the
 * vector's packed data (cf. the per-vector {bw, offset} tables of
 * fastlanes/src/ssb/fls_q21_bitpacked_opt_v4.cu:293-331).                     */
typedef uint64_t fls_vmeta;
#define FLS_VMETA(offset, bw) (((uint64_t)(offset) << 8) | (uint64_t)(bw))
#define FLS_VMETA_BW(m) ((uint32_t)((m)&0xffu))
#define FLS_VMETA_OFFSET(m) ((uint64_t)(m) >> 8)

/* FastLanes delta of a column (fastlanes/example/fastlanes_bench_delta.cu:143-155:
 * transpose_i, bases = first 32 transposed values, unrsum, pack at the
 * smallest bw of each vector).  Writes vmeta[n_vec], bases[32*n_vec] and the
 * packed words (capacity out_cap words); *out_words gets the words used.     */
int fls_delta_encode_column_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint64_t out_cap,
                                fls_vmeta* vmeta, uint32_t* bases, uint64_t* out_words);

/* FastLanes RLE of one vector (project format; the reference's RLE is a CPU
 * sketch, fastlanes/src/include/fls_gen/rle/rle.hpp:8-43): the vector's runs
 * become
 *   - a run index per row (0-based run number), FastLanes-delta coded exactly
 *     like a delta column: 32 bases (index at each lane's chain start) and
 *     per-row deltas of width ibw (0 or 1) packed at the vector's offset;
 *   - the run values minus vbase, linear LSB-first bit-packing at vbw bits,
 *     right after the 32*ibw index words.
 * vmeta holds (offset, ibw); fls_rle_meta the rest.  Decoding yields
 * transposed row order, like delta.                                           */
typedef struct fls_rle_meta {
    uint32_t vbase;  /* smallest run value */
    uint16_t n_runs; /* runs in the vector (<= 1024) */
    uint8_t vbw;     /* width of run value - vbase */
    uint8_t ibw;     /* width of the run-index deltas (0 or 1) */
} fls_rle_meta;

/* ============================================================ GPU encoders
 * Device buffers in, device buffers out, on `stream`.  Variable-width
 * encodings run as plan (per-vector widths, bases, sizes and the offsets'
 * prefix sum into vmeta; the total word count into *d_total_words) then pack
 * (into a buffer of that many words).  d_scratch needs
 * fls_encode_scratch_bytes(n) bytes.                                          */
uint64_t fls_encode_scratch_bytes(uint64_t n);
/* pack (value - base) at bw for n values; FLS_ENC_FLAG_TRANSPOSE as above     */
int fls_pack_column_device(const uint32_t* d_in, uint64_t n, uint32_t bw, uint32_t base, uint32_t flags,
                           uint32_t* d_out, void* stream);
int fls_delta_plan_device(const uint32_t* d_in, uint64_t n, fls_vmeta* d_vmeta, uint32_t* d_bases,
                          uint64_t* d_total_words, void* d_scratch, void* stream);
int fls_delta_pack_device(const uint32_t* d_in, uint64_t n, const fls_vmeta* d_vmeta, uint32_t* d_out, void* stream);
int fls_rle_plan_device(const uint32_t* d_in, uint64_t n, fls_vmeta* d_vmeta, fls_rle_meta* d_rmeta,
                        uint32_t* d_bases, uint64_t* d_total_words, void* d_scratch, void* stream);
int fls_rle_pack_device(const uint32_t* d_in, uint64_t n, const fls_vmeta* d_vmeta, const fls_rle_meta* d_rmeta,
                        uint32_t* d_out, void* stream);

/* ========================================================== GPU decode (gfx950)
 * Grid-level decode of n_vec fixed-width vectors into normal order
 * out[v*1024 + 32*r + l].  Replaces __global__ unpack_global(in, out, bw)
 * launched <<<n_vec, 32>>> (unpack.cuh:3436-3442, unpack_fused.cuh:3444-3453).
 * A 16-byte aligned d_out takes 16-byte stores; any 4-byte aligned d_out is
 * correct.                                                                    */
int fls_unpack_u32(const uint32_t* d_in, uint32_t* d_out, uint64_t n_vec, uint32_t bw, void* stream);
/* Same decode that also produces the order-sensitive checksum
 * sum_i (i+1)*value_i mod 2^64 over the first n_values values into
 * d_checksum[0] (overwritten), in one launch.  d_checksum points to
 * FLS_CHECKSUM_WORDS uint64 words: words 1.. hold the in-launch reduction's
 * state, which must be zero before the first call on the buffer (allocate it
 * zeroed once) and which every completed call leaves zero again.  Calls that
 * overlap in time need buffers of their own.  d_out may be NULL to decode
 * without materialising; when given, it receives whole vectors and must hold
 * ceil(n_values/1024)*1024 words (the padding rows of the last vector are
 * written too).                                                              */
#define FLS_CHECKSUM_WORDS 1056u
int fls_unpack_u32_checksum(const uint32_t* d_in, uint32_t* d_out, uint64_t n_values, uint32_t bw,
                            uint64_t* d_checksum, void* stream);
/* FastLanes delta decode: unpack at each vector's bw + d_rsum_32 with the
 * vector's 32 bases (rsum.cuh:3-107, delta_shared_memory.cu:41-45).  Output is
 * in transposed order (as fastlanes_bench_delta.cu:183-191 checks it).        */
int fls_delta_decode_u32(const uint32_t* d_words, const fls_vmeta* d_vmeta, const uint32_t* d_bases,
                         uint64_t n_vec, uint32_t* d_out, void* stream);
/* FastLanes RLE decode (format at fls_rle_meta), transposed order.  No
 * reference kernel: the reference's RLE exists only as a CPU sketch
 * (fls_gen/rle/rle.hpp:8-43) and an unused bit-unpack (hardcoded_16.cuh:1210). */
int fls_rle_decode_u32(const uint32_t* d_words, const fls_vmeta* d_vmeta, const fls_rle_meta* d_rmeta,
                       const uint32_t* d_bases, uint64_t n_vec, uint32_t* d_out, void* stream);

/* ================================================================ columns
 * A compressed column on the device.  All columns read by one query must use
 * the same row order inside each vector: either normal order, or FastLanes'
 * transposed order (transposed = 1), which delta and RLE always produce.     */
enum {
    FLS_ENC_BITPACK = 0, /* fixed bw for every vector, + frame of reference `base` */
    FLS_ENC_DELTA = 1,   /* FastLanes delta, per-vector (offset,bw) + 32 bases     */
    FLS_ENC_RLE = 2,     /* FastLanes RLE (run values + delta-coded run index)     */
    FLS_ENC_MIXED = 3    /* each vector its own encoding (learned/exhaustive pick) */
};

/* Per-vector encodings of a MIXED column.  vmeta[v] = offset << 8 |
 * enc << 6 | width; aux[v] = FOR base, or for RLE vbase | n_runs << 32 |
 * ibw << 48.  Payload at `offset` (words), every vector in transposed order:
 *   BITPACK: 32*width words;  FOR: 32*width words of value - base;
 *   DELTA: 32 bases, then 32*width words of deltas;
 *   RLE: 32 run-index bases, 32*ibw index-delta words, then the run values -
 *        vbase linear-packed at `width` bits.
 * A vector's encoded size is 4*payload words + 16 bytes (vmeta + aux).       */
enum { FLS_MIX_BITPACK = 0, FLS_MIX_FOR = 1, FLS_MIX_DELTA = 2, FLS_MIX_RLE = 3 };

typedef struct fls_column {
    uint32_t encoding;       /* FLS_ENC_* */
    uint32_t bw;             /* BITPACK: width of every vector */
    uint32_t base;           /* BITPACK: value = decoded + base (FOR) */
    uint32_t transposed;     /* rows in FastLanes transposed order */
    uint64_t n_words;        /* packed words in `words` (for sizes/bounds) */
    const uint32_t* words;   /* packed words */
    const fls_vmeta* vmeta;  /* DELTA / RLE / MIXED: per-vector metadata; else NULL */
    const uint32_t* bases;   /* DELTA / RLE: 32 bases per vector; else NULL */
    const uint64_t* aux;     /* RLE: fls_rle_meta per vector; MIXED: aux per vector */
} fls_column;

/* ==================================================== fused decode + query
 * Q1.1 shape, fls_q11.cu:29-81 (QueryKernel_v1) and its driver
 * query<BT,IPT>() (fls_q11.cu:137-171): rows with
 *   date_gt < orderdate < date_lt  AND  quantity < qty_lt  AND
 *   disc_lo <= discount <= disc_hi
 * contribute price*discount (64-bit product) to *d_revenue; *d_count (may be
 * NULL) gets the number of qualifying rows.  Both are overwritten unless flags
 * has FLS_Q_ACCUMULATE.  One kernel launch per call: the blocks' sums are
 * reduced inside it through one of 64 reduction slots of a static per-device
 * array, taken round-robin, so up to 64 calls may run at the same time on one
 * device.  orderdate may be DELTA or BITPACK; the others BITPACK.           */
typedef struct fls_q11_params {
    uint32_t date_gt, date_lt, qty_lt, disc_lo, disc_hi;
} fls_q11_params;

/* Skip loading (not only decoding) a wave's later columns once none of its
 * rows can qualify -- the predicated-load variant (Crystal-opt BlockPredLoad,
 * crystal-opt/load.cuh:25-97; fls_q11.cu:83-135).  Without it every byte of
 * every column is read (full scan).                                          */
#define FLS_Q_PRED_LOAD 1u
/* Add into *d_revenue / *d_count instead of overwriting them: the caller
 * zeroes them first, as query<>() does with cudaMemset (fls_q11.cu:146-148). */
#define FLS_Q_ACCUMULATE 2u

int fls_query_q11(const fls_column* orderdate, const fls_column* quantity, const fls_column* discount,
                  const fls_column* price, uint64_t n_rows, const fls_q11_params* params, uint32_t flags,
                  int64_t* d_revenue, uint64_t* d_count, void* stream);

/* ============================================ hash probe + group-by (Q2.1)
 * Shape of fls_q21.cu:21-78: keep fact rows whose supplier passes one
 * dimension filter and whose part passes another, look up the part's brand
 * and the date's year, and sum revenue per group slot
 *   slot = (brand * n_years + (year - year_base)) % n_slots    (fls_q21.cu:71)
 * into d_sums[n_slots] (int64) with d_counts[n_slots] qualifying rows per
 * slot; a group exists where its count is non-zero.  Both tables are
 * overwritten unless flags has FLS_Q_ACCUMULATE.  The compact probe and the
 * bit-filtered reference-table probe (n_slots <= 8192) add each block's groups
 * to one of 16 partial tables and a second launch sums those into the output;
 * the partial tables are one of 8 sets of a static per-device array, taken
 * round-robin, so up to 8 such calls may run at the same time on one device.
 * The other paths zero the output first and add to it.                      */
typedef struct fls_q21_params {
    uint32_t category;  /* part filter: p_category == category */
    uint32_t region;    /* supplier filter: s_region == region */
    uint32_t year_base; /* 1992 */
    uint32_t n_years;   /* 7 */
    uint32_t n_slots;   /* 7000 */
} fls_q21_params;

/* Reference perfect hash tables (crystal/join.cuh:6-275).  PHT_1 holds the key
 * in slot HASH(key) = (key - min) % len, 0 meaning empty; PHT_2 holds
 * [key, value] pairs.  ht_s: supplier PHT_1; ht_p: part PHT_2 -> brand; ht_d:
 * date PHT_2 -> year.  occ_s / occ_p: NULL, or the occupancy bitmaps of ht_s
 * and ht_p (fls_pht_occupancy_u32); with both given, the probe tests the bits
 * (staged in LDS) and reads the tables only for rows that pass both.       */
typedef struct fls_q21_pht {
    const uint32_t* ht_s;
    uint32_t s_len, s_min;
    const uint32_t* ht_p;
    uint32_t p_len, p_min;
    const uint32_t* ht_d;
    uint32_t d_len, d_min;
    const uint32_t* occ_s;
    const uint32_t* occ_p;
} fls_q21_pht;

/* Occupancy bitmap of a reference hash table: bit i of d_bits[i / 32] is set
 * when slot i is not empty (PHT_1: ht[i] != 0; pair != 0, PHT_2: the 64-bit
 * slot [ht[2i], ht[2i+1]] != 0), the test the reference's probe makes per
 * row (crystal/join.cuh:8-149).  d_bits holds ceil(ht_len / 32) words.  No
 * reference counterpart.                                                     */
int fls_pht_occupancy_u32(const uint32_t* d_ht, uint32_t ht_len, uint32_t pair, uint32_t* d_bits, void* stream);

/* build_hashtable_s/_p/_d (fls_q21.cu:197-263): rows with
 * d_filter[i] == filter_eq (every row when d_filter is NULL) enter the table
 * as PHT_1 (d_vals NULL) or PHT_2 (value d_vals[i]).  d_ht must be zeroed by
 * the caller (2*ht_len words for PHT_2), as runQuery does with cudaMemset
 * (fls_q21.cu:294-297).  Keys must be non-zero.                              */
int fls_build_pht_u32(const uint32_t* d_filter, uint32_t filter_eq, const uint32_t* d_keys, const uint32_t* d_vals,
                      uint32_t n, uint32_t* d_ht, uint32_t ht_len, uint32_t keys_min, void* stream);
/* probe_v1 (fls_q21.cu:21-78) over reference hash tables: drop-in path.  With
 * the occupancy bitmaps in *ht and columns at the SSB widths (date 12,
 * supplier 16, part 20, revenue 24 bits), the bit-filtered probe; otherwise
 * one table slot read per row and table through the Crystal primitives.  The
 * same rows qualify either way.                                             */
int fls_query_q21_pht(const fls_column* datekey, const fls_column* suppkey, const fls_column* partkey,
                      const fls_column* revenue, uint64_t n_rows, const fls_q21_pht* ht, const fls_q21_params* params,
                      int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream);

/* Dense-key dimension columns: row i holds key keys_min + i; each holds at
 * least one row (FLS_ERR_ARG otherwise).                                      */
typedef struct fls_q21_dims {
    const uint32_t* s_region;
    uint32_t n_supp;
    const uint32_t* p_category;
    const uint32_t* p_brand;
    uint32_t n_part;
    const uint32_t* d_year;
    uint32_t n_date;
    uint32_t keys_min;
} fls_q21_dims;

/* MI355X-native path: fls_q21_build reduces the dimensions into a caller-owned
 * workspace of fls_q21_workspace_bytes() bytes (bitmaps of qualifying
 * suppliers and parts, 16-bit brand per part, 16-bit year per date);
 * fls_query_q21 probes it with the bitmaps and the year table staged in LDS.
 * Keys outside [keys_min, keys_min + n) do not match.                        */
uint64_t fls_q21_workspace_bytes(uint32_t n_supp, uint32_t n_part, uint32_t n_date);
int fls_q21_build(const fls_q21_dims* dims, const fls_q21_params* params, void* d_workspace, uint64_t ws_bytes,
                  void* stream);
int fls_query_q21(const fls_column* datekey, const fls_column* suppkey, const fls_column* partkey,
                  const fls_column* revenue, uint64_t n_rows, const fls_q21_dims* dims, const void* d_workspace,
                  const fls_q21_params* params, int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream);

/* Allocation-free fls_q21_build + fls_query_q21 in one call, with the same
 * arguments minus the workspace and the same results.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The compact dimension tables are built into a static device array of
 * FLS_Q21_STATIC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends; dimensions whose fls_q21_workspace_bytes() exceeds it
 * return FLS_ERR_SCRATCH before any work is queued.  The call allocates
 * nothing and clears what must start at zero itself, but every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams, other
 * host threads) race.                                                        */
#define FLS_Q21_STATIC_SCRATCH_BYTES (4u << 20)
int fls_query_q21_noalloc(const fls_column* datekey, const fls_column* suppkey, const fls_column* partkey,
                          const fls_column* revenue, uint64_t n_rows, const fls_q21_dims* dims,
                          const fls_q21_params* params, int64_t* d_sums, uint64_t* d_counts, uint32_t flags,
                          void* stream);

/* ============================================ star joins (Q3.1, Q4.1)
 * Dense-key dimension columns for the three- and four-way star joins: row i of
 * each dimension holds key keys_min + i, and each used dimension holds at
 * least one row (FLS_ERR_ARG otherwise).  Nations must be below 255 (rows with
 * a larger nation do not match).  Columns a shape does not use may be NULL:
 * p_mfgr (with n_part 0) for Q3.1, s_nation for Q4.1.                        */
typedef struct fls_star_dims {
    const uint32_t* s_region;
    const uint32_t* s_nation;
    uint32_t n_supp;
    const uint32_t* c_region;
    const uint32_t* c_nation;
    uint32_t n_cust;
    const uint32_t* p_mfgr;
    uint32_t n_part;
    const uint32_t* d_year;
    uint32_t n_date;
    uint32_t keys_min;
} fls_star_dims;

/* Shape of fls_q31.cu:28-94 (probe_v1) with builds fls_q31.cu:300-372: rows
 * whose customer has c_region == c_region, whose supplier has s_region ==
 * s_region and whose date's year is in [year_lo, year_hi] add revenue to
 *   slot = (s_nation * n_nations * n_years + c_nation * n_years
 *           + (year - year_base)) % n_slots                 (fls_q31.cu:84-85)
 * The reference's constants: n_nations 25, n_years 7, year_base 1992,
 * n_slots 4375.                                                              */
typedef struct fls_q31_params {
    uint32_t c_region, s_region;
    uint32_t year_lo, year_hi;
    uint32_t year_base, n_years, n_nations;
    uint32_t n_slots;
} fls_q31_params;

/* Shape of fls_q41.cu:27-107 (probe_v1) with builds fls_q41.cu:186-271: rows
 * whose supplier has s_region == s_region, whose customer has c_region ==
 * c_region and whose part has mfgr_lo <= p_mfgr <= mfgr_hi add
 * revenue - supplycost (signed) to
 *   slot = (c_nation * n_years + (year - year_base)) % n_slots  (fls_q41.cu:97)
 * Dates whose year is outside [year_base, year_base + 254] do not match (every
 * SSB date is 1992-1998).  The reference's constants: n_years 7, year_base
 * 1992, n_slots 175.                                                          */
typedef struct fls_q41_params {
    uint32_t c_region, s_region;
    uint32_t mfgr_lo, mfgr_hi;
    uint32_t year_base, n_years;
    uint32_t n_slots;
} fls_q41_params;

/* MI355X-native builds: one byte per supplier, customer and date (nation or
 * year offset, 0xFF for a filtered row), one bit per part and per qualifying
 * supplier (Q4.1) and one bit per qualifying customer (Q3.1), in a
 * caller-owned workspace of fls_star_workspace_bytes() bytes.  The probes
 * stage the small tables (supplier codes or bits, part or customer bits,
 * date codes) in LDS, read the customer codes through L2, and group in
 * per-block LDS tables flushed once per block.
 * d_sums / d_counts hold n_slots (<= 8192) entries each, overwritten unless
 * flags has FLS_Q_ACCUMULATE.  Keys outside [keys_min, keys_min + n) do not
 * match.  All fact columns are BITPACK with one row order.                   */
uint64_t fls_star_workspace_bytes(const fls_star_dims* dims);
int fls_q31_build(const fls_star_dims* dims, const fls_q31_params* params, void* d_workspace, uint64_t ws_bytes,
                  void* stream);
int fls_query_q31(const fls_column* datekey, const fls_column* custkey, const fls_column* suppkey,
                  const fls_column* revenue, uint64_t n_rows, const fls_star_dims* dims, const void* d_workspace,
                  const fls_q31_params* params, int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream);
int fls_q41_build(const fls_star_dims* dims, const fls_q41_params* params, void* d_workspace, uint64_t ws_bytes,
                  void* stream);
int fls_query_q41(const fls_column* datekey, const fls_column* custkey, const fls_column* suppkey,
                  const fls_column* partkey, const fls_column* revenue, const fls_column* supplycost,
                  uint64_t n_rows, const fls_star_dims* dims, const void* d_workspace, const fls_q41_params* params,
                  int64_t* d_sums, uint64_t* d_counts, uint32_t flags, void* stream);

/* ============================================ untranspose and materialize
 * transpose_i (inverse == 0) or transpose_o (inverse != 0) of n_vec whole
 * vectors, d_in != d_out: the unified transposed layout of
 * fastlanes/src/transpose.cpp:4110-5135 and its true inverse (the reference's
 * transpose_o at transpose.cpp:5136-6161 repeats transpose_i's permutation).
 * d_in and d_out must be 16-byte aligned (FLS_ERR_ARG otherwise).            */
int fls_transpose_u32(const uint32_t* d_in, uint64_t n_vec, uint32_t inverse, uint32_t* d_out, void* stream);

/* Decode a column of any encoding into plain values for a downstream operator
 * (a decode followed by Crystal's BlockStore, crystal-opt/store.cuh:57-69):
 *   FLS_ORDER_STORED  the column's own row order (transposed for DELTA, RLE,
 *                     MIXED and transposed BITPACK); whole vectors, so d_out
 *                     holds ceil(n_rows / 1024) * 1024 values;
 *   FLS_ORDER_ROWS    original row order (transposed columns are untransposed
 *                     on the fly); exactly n_rows values.
 * A 16-byte aligned d_out takes 16-byte stores; any 4-byte aligned d_out is
 * correct.                                                                   */
#define FLS_ORDER_STORED 0u
#define FLS_ORDER_ROWS 1u
int fls_decode_column_u32(const fls_column* col, uint64_t n_rows, uint32_t order, uint32_t* d_out, void* stream);

/* ============================== tile-based competitor formats (comparison)
 * GPU-FOR, GPU-DFOR and GPU-RFOR of the tile-based compression baseline the
 * reference carries (tile_based/, Shanbhag et al.), word-compatible with its
 * encoders: bin_pack (tile_based_bench_bitpack.cu:6-83), deltaBinPack
 * (tile_based_bench_delta.cu:8-112) and rleBinPack (rlebinpack.cpp:9-146).
 * A 4-word header {block size, 4, total count, first value}, then blocks;
 * starts[] holds every block's word offset and the end: ceil(n/128) + 1
 * entries for FOR, 4 ceil(n/512) + 1 for DFOR (4 blocks behind each 512-value
 * tile's first value), ceil(n/512) + 1 for RFOR (one run-value and one
 * run-length block per tile).  Every block has one width; width 32 is one
 * value per word (the reference loops shift by 32 there, undefined in C++).
 * Inputs are padded to whole blocks with the last value.  Encoders run on the
 * host into buffers of fls_tile_max_words(n) words.                         */
uint64_t fls_tile_max_words(uint64_t n);
int fls_tile_for_encode_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts, uint64_t* n_words);
int fls_tile_dfor_encode_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* starts, uint64_t* n_words);
int fls_tile_rfor_encode_u32(const uint32_t* in, uint64_t n, uint32_t* values, uint32_t* val_starts,
                             uint32_t* run_lengths, uint32_t* rl_starts, uint64_t* n_val_words, uint64_t* n_rl_words);
/* Device decoders (LoadBinPack / LoadDBinPack / LoadRBinPack,
 * tile_based/src/include/binpack_kernel.cuh:20-84, deltabinpack_kernel.cuh:
 * 39-103, rlebinpack_kernel.cuh:20-146): n values to d_out and/or their sum to
 * *d_sum (zeroed first); at least one of the two non-NULL.                   */
int fls_tile_for_decode_u32(const uint32_t* d_data, const uint32_t* d_starts, uint64_t n, uint32_t* d_out,
                            uint64_t* d_sum, void* stream);
int fls_tile_dfor_decode_u32(const uint32_t* d_data, const uint32_t* d_starts, uint64_t n, uint32_t* d_out,
                             uint64_t* d_sum, void* stream);
int fls_tile_rfor_decode_u32(const uint32_t* d_values, const uint32_t* d_val_starts, const uint32_t* d_run_lengths,
                             const uint32_t* d_rl_starts, uint64_t n, uint32_t* d_out, uint64_t* d_sum,
                             void* stream);

/* Encoded bytes a full scan of a column reads for n_rows (packed words, plus
 * per-vector metadata and bases for DELTA).                                   */
uint64_t fls_column_bytes(const fls_column* col, uint64_t n_rows);

/* ================================================== serialized columns
 * A self-describing image of one compressed column, for files and the wire
 * (no reference counterpart; the reference keeps encoded columns in memory
 * only, with per-vector metadata {bw, base, offset} as in
 * fls_q21_bitpacked_opt_v4.cu:293-331).  Layout, little-endian:
 *   fls_column_image header (128 bytes, magic "FLSCOL01", version 1),
 *   then the sections words (u32 x n_words), vmeta (u64 x n_vmeta),
 *   bases (u32 x n_bases), aux (u64 x n_aux), each starting at a multiple of
 *   64 bytes from the image start; checksum = FNV-1a 64 over the sections.
 * n_vmeta / n_bases / n_aux follow from the encoding and n_rows:
 *   BITPACK 0 / 0 / 0, DELTA n_vec / 32 n_vec / 0, RLE n_vec / 32 n_vec /
 *   n_vec (fls_rle_meta), MIXED n_vec / 0 / n_vec.
 * Serialization reads host memory (copy device columns to the host first);
 * deserialization validates the header, every size, the checksum and every
 * vector's metadata (widths 0..32, RLE run counts 1..1024, each vector's
 * payload inside the words), and returns a column whose pointers are views
 * into the image (no copy).  Serialization rejects a column whose metadata
 * fails the same checks (FLS_ERR_FORMAT).                                   */
typedef struct fls_column_image {
    char magic[8];
    uint32_t version, encoding, bw, base, transposed, reserved;
    uint64_t n_rows, n_words, n_vmeta, n_bases, n_aux;
    uint64_t off_words, off_vmeta, off_bases, off_aux, total_bytes, checksum;
    uint8_t pad[8];
} fls_column_image;

uint64_t fls_column_image_bytes(const fls_column* col, uint64_t n_rows);
int fls_column_serialize(const fls_column* col, uint64_t n_rows, void* out, uint64_t out_bytes);
int fls_column_deserialize(const void* image, uint64_t image_bytes, fls_column* col, uint64_t* n_rows);

/* ===================================================== encoding selection
 * Classes: 4 encodings (FLS_MIX_*) x 10 width buckets
 * (0 | 1-2 | 3-4 | 5-6 | 7-8 | 9-12 | 13-16 | 17-20 | 21-24 | 25-32);
 * class = 10*encoding + bucket(width of that encoding).  No reference
 * counterpart (the reference fixes each column's encoding by hand; its
 * nearest helper is find_bw, ssb_utils.h:222-239).
 * fls_exhaustive_select: for every vector the exact encoded size of all four
 * encodings (d_sizes4[4*v+e], may be NULL) and the smallest class d_class[v].
 * fls_vector_features: 16 float features per vector from a 1/8 sample: the
 * 32-row (128-byte) lines 0, 10, 20 and 31 of the vector.                  */
int fls_exhaustive_select(const uint32_t* d_in, uint64_t n, uint8_t* d_class, uint32_t* d_sizes4, void* stream);
int fls_vector_features(const uint32_t* d_in, uint64_t n, float* d_feat, void* stream);
/* Mixed column from per-vector classes (only each vector's chosen encoding is
 * analysed): plan (vmeta, aux, total words; scratch as fls_encode_scratch_bytes)
 * then pack into at most fls_mixed_max_words(n) words.  d_class is read and,
 * for fallback vectors, rewritten: when d_conf is not NULL, a vector with
 * d_conf[v] < conf_min is planned by exhaustive search instead, and
 * *d_fallbacks (may be NULL) is incremented for it.                          */
uint64_t fls_mixed_max_words(uint64_t n);
int fls_mixed_plan_device(const uint32_t* d_in, uint64_t n, uint8_t* d_class, const float* d_conf, float conf_min,
                          fls_vmeta* d_vmeta, uint64_t* d_aux, uint64_t* d_total_words, uint32_t* d_fallbacks,
                          void* d_scratch, void* stream);
int fls_mixed_pack_device(const uint32_t* d_in, uint64_t n, const fls_vmeta* d_vmeta, const uint64_t* d_aux,
                          uint32_t* d_out, void* stream);
/* The same mixed column in one pass over the input: plan, offsets (a
 * decoupled look-back over the blocks, in the scratch of
 * fls_encode_scratch_bytes(n), cleared by a small fill launch first) and pack
 * in one launch.  Arguments and results as plan + pack above; d_out receives
 * the words.  Calls sharing one scratch must not overlap in time.  At 120M
 * rows plan + pack is the faster of the two on MI355X.                     */
int fls_mixed_encode_device(const uint32_t* d_in, uint64_t n, uint8_t* d_class, const float* d_conf, float conf_min,
                            fls_vmeta* d_vmeta, uint64_t* d_aux, uint32_t* d_out, uint64_t* d_total_words,
                            uint32_t* d_fallbacks, void* d_scratch, void* stream);
/* Try-all-encodings search: every vector is encoded with each of the four
 * encodings (four full plan + pack passes into candidate columns held in
 * d_work, fls_trial_workspace_bytes(n) bytes), then the smallest candidate of
 * each vector is copied into the MIXED column (d_out: fls_mixed_max_words(n)
 * words).  The baseline the learned selector is measured against; its
 * output equals the exhaustive-class column byte for byte.                  */
uint64_t fls_trial_workspace_bytes(uint64_t n);
int fls_mixed_encode_trial(const uint32_t* d_in, uint64_t n, uint8_t* d_class, fls_vmeta* d_vmeta, uint64_t* d_aux,
                           uint32_t* d_out, uint64_t* d_total_words, void* d_work, void* stream);
/* decode a MIXED column into transposed order (out[v*1024 + 32r + l])        */
int fls_mixed_decode_u32(const uint32_t* d_words, const fls_vmeta* d_vmeta, const uint64_t* d_aux, uint64_t n_vec,
                         uint32_t* d_out, void* stream);

/* ======================================================= learned selector
 * MLP 16 -> 64 -> 64 -> 40 (ReLU, softmax) on BF16 MFMA with FP32
 * accumulation.  Parameters (fls_mlp_param_count() = 7848 floats):
 * W1[16][64], b1[64], W2[64][64], b2[64], W3[64][40], b3[40], W[in][out].
 * d_wb is the BF16 copy the kernels read (fls_mlp_to_bf16 / fls_mlp_adam
 * refresh it).  No reference counterpart.
 *   forward: per feature row, the encoding e with the largest softmax mass
 *            (classes 10e..10e+9), the class = its most likely width bucket,
 *            and conf = that mass; d_conf / d_logits may be NULL.  d_feat
 *            (n rows of 16 floats) must be 16-byte aligned (FLS_ERR_ARG).
 *   grad: adds the gradient of the mean cross-entropy (inv_batch = 1/batch)
 *         over n rows into d_grad, the summed loss into *d_loss (may be NULL);
 *         d_work holds fls_mlp_grad_workspace_bytes(n) bytes.  Per-wave
 *         partial sums are added in a fixed order: the result is
 *         bit-reproducible.
 *   adam: one Adam step (step >= 1) on the FP32 weights; refreshes d_wb and
 *         zeroes d_grad.
 *   mfma_selftest: D = A * B for one 16x16x16 BF16 tile (row-major) through
 *         the kernels' own operand path, to pin the matrix-core lane layout;
 *         mfma32_selftest likewise for A 16x32, B 32x16 through the
 *         16x16x32 instruction's path.                                     */
uint32_t fls_mlp_param_count(void);
int fls_mlp_to_bf16(const float* d_w, uint16_t* d_wb, void* stream);
int fls_mlp_forward(const uint16_t* d_wb, const float* d_feat, uint64_t n, uint8_t* d_class, float* d_conf,
                    float* d_logits, void* stream);
uint64_t fls_mlp_grad_workspace_bytes(uint64_t n);
int fls_mlp_grad(const uint16_t* d_wb, const float* d_feat, const uint8_t* d_label, uint64_t n, float inv_batch,
                 void* d_work, float* d_grad, float* d_loss, void* stream);
int fls_mlp_adam(float* d_w, float* d_m, float* d_v, float* d_grad, uint16_t* d_wb, uint32_t step, float lr, float beta1,
                 float beta2, float eps, void* stream);
int fls_mlp_mfma_selftest(const uint16_t* d_a, const uint16_t* d_b, float* d_d, void* stream);
int fls_mlp_mfma32_selftest(const uint16_t* d_a, const uint16_t* d_b, float* d_d, void* stream);
/* Training vectors: n_vec x 1024 values from the four families of the
 * learned selector's workload (uniform, Zipf s=1.1 via d_zipf_cdf[4096], the
 * CDF of ranks 1..4096 scaled to 2^32, sorted + Gaussian noise sigma 16, runs
 * of geometric mean 32); vector v's family and parameters come from the seed. */
int fls_gen_train_vectors(uint64_t seed, uint64_t first_vector, uint64_t n_vec, const uint32_t* d_zipf_cdf,
                          uint32_t* d_out, void* stream);

/* ============================================ batched segments (no reference)
 * Many compressed segments of uneven length, width and encoding decoded and
 * range-filtered (lo <= value <= hi) in ONE persistent launch, counting the
 * qualifying values per segment into d_counts[n_segments] (overwritten).
 * The reference launches one kernel per query column set; this has no
 * counterpart there.                                                         */
typedef struct fls_batch fls_batch; /* opaque */

typedef struct fls_segment {
    fls_column col;             /* BITPACK, DELTA, RLE or MIXED (device buffers) */
    const fls_rle_meta* rmeta;  /* RLE only; NULL: taken from col.aux */
    uint64_t n_values;
    uint32_t lo, hi;            /* range predicate */
} fls_segment;

typedef struct fls_segment_info {
    uint32_t encoding, bw;
    uint64_t n_values;
    uint32_t lo, hi;
    uint64_t n_words;
} fls_segment_info;

typedef struct fls_batch_info {
    uint32_t n_segments;
    uint64_t n_values, n_vectors, encoded_bytes;
    uint32_t n_bitpack, n_delta, n_rle;
} fls_batch_info;

enum { FLS_SCHED_STATIC = 0, FLS_SCHED_DYNAMIC = 1 };
/* launch shape: vectors per tile (a tile never spans segments), work-items
 * per block (256 or 512), tiles handed out round-robin or from a work queue.
 * class_tiles: 0, or a tile size per segment class c = 2*encoding + (more
 * than 8 encoded bits per value): nibble c (bits 4c..4c+3) = log2 of that
 * class's tile size, 0 meaning tile_vectors.                                */
typedef struct fls_tune {
    uint32_t tile_vectors, block_threads, schedule, class_tiles;
} fls_tune;

/* Segments the caller already holds on the device (pointers are borrowed). */
int fls_batch_create(const fls_segment* segs, uint32_t n_segments, fls_batch** out);
/* The batched-segments workload, generated and encoded on the GPU: segment s
 * has length 2^e + U[0,2^e) with octave e uniform in [log2_min, log2_max)
 * (log-uniform by octave), bw U{1..32}, encoding 40% BITPACK (values
 * U[0,2^bw)), 30% DELTA (sorted: value_i = i*m + U[0,m), m = 2^(b-1),
 * b = min(bw, 32 - bitwidth(len))), 30% RLE (runs starting with probability
 * 1/8 per row, i.e. geometric lengths of mean 8, run values U[0,2^bw)); the
 * predicate is the middle half of the segment's [min, max]:
 * lo = min + (max-min)/4, hi = min + 3(max-min)/4.                           */
int fls_batch_create_c4(uint64_t seed, uint32_t n_segments, uint32_t log2_min_len, uint32_t log2_max_len,
                        fls_batch** out, void* stream);
void fls_c4_segment_params(uint64_t seed, uint32_t s, uint32_t log2_min_len, uint32_t log2_max_len,
                           uint32_t* encoding, uint32_t* bw, uint64_t* n_values);
int fls_gen_c4_segment(uint64_t seed, uint32_t s, uint32_t log2_min_len, uint32_t log2_max_len, uint32_t* d_out,
                       void* stream);
int fls_batch_get_info(const fls_batch* b, fls_batch_info* info);
int fls_batch_segment(const fls_batch* b, uint32_t s, fls_segment_info* info);
/* Segment s as fls_batch_create takes it: device pointers borrowed from the
 * batch (valid until fls_batch_destroy), e.g. to copy a segment out.        */
int fls_batch_segment_column(const fls_batch* b, uint32_t s, fls_segment* seg);
/* one persistent launch (after two tiny zero-fill kernels) */
int fls_batch_run(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream);
/* baseline: one launch per segment */
int fls_batch_run_per_segment(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream);
/* fls_batch_run captured once in a hipGraph (re-captured when tune or
 * d_counts change) and replayed on `stream`                                  */
int fls_batch_graph_launch(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream);
/* time every (tile, block, schedule) candidate on this batch, then, at the
 * best block and schedule, every tile size per segment class on that class's
 * segments alone; keep the per-class tiling when the whole batch runs faster
 * with it;
 * results are cached in cache_path (NULL: $FLS_AUTOTUNE_CACHE, else
 * ~/.cache/fls_mi355x/autotune.txt) under a key of device + batch shape and
 * reused unless force != 0.  *best_ms = -1 when the cache answered.          */
int fls_batch_autotune(fls_batch* b, const char* cache_path, int force, uint64_t* d_counts, fls_tune* best,
                       double* best_ms, void* stream);
void fls_batch_destroy(fls_batch* b);

/* ====================================================== synthetic workloads
 * Deterministic generators computed on the GPU (no host data moves).  Every
 * value is a pure function of (seed, column, row).  They have no counterpart
 * in the reference, whose drivers read dbgen files (crystal_ssb_utils.h:108).
 *   c1: n values uniform in [0, 2^bw).
 *   c2: fact table of n rows: orderdate sorted and uniform over 1992-01-01 ..
 *       1998-12-31 (row i on day floor(i*2557/n)), quantity U{1..50},
 *       discount U{0..10}, price U[90000, 10^7).
 *   c3: fact foreign keys as 0-based dense keys, date U[0,2556),
 *       supplier U[0,40000), part U[0,800000), customer U[0,600000) (may be
 *       NULL), revenue U[0,10^7); dimensions: year of each date row, supplier
 *       and customer nation U{1..25} with region (nation - 1) / 5 + 1, part
 *       category U{1..25} with brand 40 (category - 1) + U{1..40} -- SSB's
 *       nation -> region and category -> brand hierarchies.
 *   star extra (Q3.1 / Q4.1 columns the c3 table lacks): supplycost of each
 *       fact row U[60000, 120000), manufacturer of each part U{1..5}.       */
int fls_gen_c1(uint64_t seed, uint64_t n, uint32_t bw, uint32_t* d_out, void* stream);
int fls_gen_c2(uint64_t seed, uint64_t n, uint32_t* d_orderdate, uint32_t* d_quantity, uint32_t* d_discount,
               uint32_t* d_price, void* stream);
int fls_gen_c3_fact(uint64_t seed, uint64_t n, uint32_t* d_datefk, uint32_t* d_suppfk, uint32_t* d_partfk,
                    uint32_t* d_custfk, uint32_t* d_revenue, void* stream);
int fls_gen_c3_dims(uint64_t seed, uint32_t n_date, uint32_t n_supp, uint32_t n_part, uint32_t n_cust,
                    uint32_t* d_year, uint32_t* d_s_region, uint32_t* d_s_nation, uint32_t* d_p_category,
                    uint32_t* d_p_brand, uint32_t* d_c_region, uint32_t* d_c_nation, void* stream);
int fls_gen_star_extra(uint64_t seed, uint64_t n, uint32_t* d_supplycost, uint32_t n_part, uint32_t* d_p_mfgr,
                       void* stream);

#ifdef __cplusplus
}
#endif
#endif /* FLS_FLS_H */
