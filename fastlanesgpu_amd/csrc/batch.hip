// Batched decode + range filter over many compressed segments of uneven
// length, bit width and encoding, in one persistent launch.
//
// Segments are cut into tiles of `tile_vectors` FastLanes vectors; a tile
// never spans two segments, so every block-wide branch on encoding and width
// is uniform.  A persistent grid (every block resident) walks the tile list
// either round-robin (static) or through an atomic work queue (dynamic).  Per
// tile the block counts the rows with lo <= value <= hi and adds the count to
// the segment's counter with one 64-bit atomic.  The launch (plus the kernels
// that zero the counters and the queue) can be captured once in a hipGraph
// and replayed; the baseline is one launch per segment.  An autotuner times
// tile sizes x schedule x block size on the batch and caches the winner on
// disk.
//
// No reference counterpart: every reference launch is one kernel per query
// over one column set (fls_q11.cu:137-171, the query driver).
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <string>
#include <vector>

#include "fls/device/unpack.hpp"
#include "fls/device/wave.hpp"
#include "fls_columns.hpp"
#include "fls_encode.hpp"
#include "fls_rng.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

struct SegDev {
    uint32_t enc, bw, transposed, pad;  // transposed: rows in FastLanes transposed order
    uint64_t n_values;
    const uint32_t* words;
    const fls_vmeta* vmeta;
    const uint32_t* bases;
    const fls_rle_meta* rmeta;
    const uint64_t* aux;  // MIXED: per-vector aux words
    uint32_t lo, hi;
};

struct TileDev {
    uint32_t seg, v0, nv, pad;
};

// ---------------------------------------------------------------- kernel
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void batch_kernel(const SegDev* __restrict segs, const TileDev* __restrict tiles,
                                                      uint32_t n_tiles, uint64_t* __restrict counts,
                                                      uint32_t* __restrict queue, uint32_t dynamic) {
    constexpr int VPB = BLOCK / 32;
    __shared__ uint32_t stage[VPB][1024];
    __shared__ uint64_t s_part[BLOCK / 64];
    __shared__ uint32_t s_tile;
    uint32_t iter = 0;
    for (;;) {
        uint32_t t;
        if (dynamic) {
            if (threadIdx.x == 0) s_tile = atomicAdd(queue, 1u);
            __syncthreads();
            t = s_tile;
        } else {
            t = blockIdx.x + iter * gridDim.x;
            ++iter;
        }
        if (t >= n_tiles) break;  // uniform over the block
        const TileDev tile = tiles[t];
        const SegDev sg = segs[tile.seg];
        const ColDev col{sg.enc, sg.bw, 0u, sg.transposed, sg.words, sg.vmeta, sg.bases, sg.aux};
        const fls_rle_meta* rmeta = sg.rmeta;
        const uint32_t span = sg.hi - sg.lo;
        uint64_t cnt = 0;
        for (uint32_t vb = 0; vb < tile.nv; vb += VPB) {
            const uint32_t local = vb + (threadIdx.x >> 5);
            const bool live = local < tile.nv;
            const uint64_t v = (uint64_t)tile.v0 + local;
            const uint32_t lo = sg.lo;
            const uint64_t first = v * 1024u;
            uint32_t c = 0;
            if (first + 1024u <= sg.n_values) {
                // a full vector: count in-range rows directly (compare + add with carry)
                auto count = [&](int, uint32_t x) { c += (uint32_t)((x - lo) <= span); };
                column_walk(col, rmeta, v, stage[threadIdx.x >> 5], live, count);
            } else {
                // the segment's partial last vector: row mask, then the rows that exist
                uint32_t m = 0;
                auto mark = [&](int r, uint32_t x) { m |= (uint32_t)((x - lo) <= span) << r; };
                column_walk(col, rmeta, v, stage[threadIdx.x >> 5], live, mark);
                if (live)
                    c = (uint32_t)__builtin_popcount(
                        m & live_row_mask((uint32_t)(sg.n_values - first), col.transposed != 0));
            }
            cnt += c;
        }
        cnt = dev::wave_sum(cnt);
        if (dev::is_wave_leader()) s_part[threadIdx.x >> 6] = cnt;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t total = 0;
#pragma unroll
            for (int w = 0; w < BLOCK / 64; ++w) total += s_part[w];
            if (total) atomicAdd((unsigned long long*)&counts[tile.seg], (unsigned long long)total);
        }
        __syncthreads();  // s_part / s_tile are rewritten next iteration
    }
}

__global__ void zero_u32_kernel(uint32_t* __restrict p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// ------------------------------------------------------------ generation
struct SegParams {
    uint32_t enc, bw;
    uint64_t len;
};

SegParams c4_params(uint64_t seed, uint32_t s, uint32_t lmin, uint32_t lmax) {
    const uint64_t key = rng::stream_key(seed, rng::kSegParams + s);
    const uint32_t e = lmin + rng::uniform(key, 0, lmax - lmin);
    SegParams p;
    p.len = (1ull << e) + rng::uniform(key, 1, 1ull << e);
    p.bw = 1u + rng::uniform(key, 2, 32);
    const uint32_t r = rng::uniform(key, 3, 10);
    p.enc = r < 4 ? FLS_ENC_BITPACK : r < 7 ? FLS_ENC_DELTA : FLS_ENC_RLE;
    return p;
}

__host__ __device__ inline uint32_t bit_width64(uint64_t x) {
    uint32_t w = 0;
    while (x) {
        ++w;
        x >>= 1;
    }
    return w;
}

// Values of segment s (see fls.h: fls_batch_create_c4).
__global__ void gen_segment_kernel(uint64_t seed, uint32_t s, uint32_t enc, uint32_t bw, uint64_t len,
                                   uint32_t* __restrict out) {
    const uint64_t kv = rng::stream_key(seed, rng::kSegValues + s);
    const uint64_t kr = rng::stream_key(seed, rng::kSegRuns + s);
    uint32_t b = 32u - bit_width64(len);
    b = b < 1u ? 1u : b;
    b = bw < b ? bw : b;
    const uint64_t m = 1ull << (b - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t x;
        if (enc == FLS_ENC_BITPACK) {
            x = rng::uniform(kv, i, 1ull << bw);
        } else if (enc == FLS_ENC_DELTA) {
            x = (uint32_t)(i * m) + rng::uniform(kv, i, m);
        } else {
            uint64_t j = i;
            while (j > 0 && rng::uniform(kr, j, 8) != 0) --j;  // start of row i's run
            x = rng::uniform(kv, j, 1ull << bw);
        }
        out[i] = x;
    }
}

__global__ void minmax_kernel(const uint32_t* __restrict in, uint64_t n, uint32_t* __restrict mm) {
    uint32_t lo = 0xffffffffu, hi = 0u;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = in[i];
        lo = min(lo, x);
        hi = max(hi, x);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (uint32_t)__shfl_xor(lo, off, 64));
        hi = max(hi, (uint32_t)__shfl_xor(hi, off, 64));
    }
    // one atomic pair per block (blockDim.x == 256)
    __shared__ uint32_t s_lo[4], s_hi[4];
    if (dev::is_wave_leader()) {
        s_lo[threadIdx.x >> 6] = lo;
        s_hi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            lo = min(lo, s_lo[w]);
            hi = max(hi, s_hi[w]);
        }
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

__global__ void mm_init_kernel(uint32_t* __restrict mm, uint32_t n_seg) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_seg; i += gridDim.x * blockDim.x) {
        mm[2 * i] = 0xffffffffu;
        mm[2 * i + 1] = 0u;
    }
}

uint32_t grid_for(uint64_t n) {
    const uint64_t g = ceil_div(n, 256);
    return (uint32_t)(g < 4096 ? (g ? g : 1) : 4096);
}

}  // namespace
}  // namespace fls

// ------------------------------------------------------------------- batch
struct fls_batch {
    int device = 0;
    std::vector<fls::SegDev> segs;
    std::vector<fls_segment_info> info;
    fls::SegDev* d_segs = nullptr;
    // owned pools (C4 builder); null for caller-described segments
    uint32_t* d_words = nullptr;
    fls_vmeta* d_vmeta = nullptr;
    uint32_t* d_bases = nullptr;
    fls_rle_meta* d_rmeta = nullptr;
    // tiles for the current tile size
    uint32_t tile_vectors = 0;
    fls::TileDev* d_tiles = nullptr;
    uint32_t n_tiles = 0;
    std::vector<uint32_t> seg_tile_off, seg_tile_cnt;
    uint32_t* d_queue = nullptr;
    // graph
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    fls_tune graph_tune{};
    uint64_t* graph_counts = nullptr;
    uint64_t n_values = 0, n_vectors = 0, encoded_bytes = 0;
    uint32_t count_by_enc[3] = {0, 0, 0};
};

namespace fls {
namespace {

int set_tiles(fls_batch* b, uint32_t tile_vectors) {
    if (tile_vectors == 0) return FLS_ERR_ARG;
    if (b->d_tiles && b->tile_vectors == tile_vectors) return FLS_OK;
    std::vector<TileDev> tiles;
    b->seg_tile_off.assign(b->segs.size(), 0);
    b->seg_tile_cnt.assign(b->segs.size(), 0);
    for (uint32_t s = 0; s < b->segs.size(); ++s) {
        const uint64_t nv = ceil_div(b->segs[s].n_values, 1024);
        b->seg_tile_off[s] = (uint32_t)tiles.size();
        for (uint64_t v0 = 0; v0 < nv; v0 += tile_vectors)
            tiles.push_back(TileDev{s, (uint32_t)v0, (uint32_t)(nv - v0 < tile_vectors ? nv - v0 : tile_vectors), 0});
        b->seg_tile_cnt[s] = (uint32_t)tiles.size() - b->seg_tile_off[s];
    }
    if (b->d_tiles) FLS_HIP_CHECK(hipFree(b->d_tiles));
    b->d_tiles = nullptr;
    FLS_HIP_CHECK(hipMalloc((void**)&b->d_tiles, sizeof(TileDev) * (tiles.size() ? tiles.size() : 1)));
    if (!tiles.empty())
        FLS_HIP_CHECK(hipMemcpy(b->d_tiles, tiles.data(), sizeof(TileDev) * tiles.size(), hipMemcpyHostToDevice));
    b->n_tiles = (uint32_t)tiles.size();
    b->tile_vectors = tile_vectors;
    return FLS_OK;
}

int upload_segs(fls_batch* b) {
    if (b->d_segs) FLS_HIP_CHECK(hipFree(b->d_segs));
    b->d_segs = nullptr;
    FLS_HIP_CHECK(hipMalloc((void**)&b->d_segs, sizeof(SegDev) * (b->segs.size() ? b->segs.size() : 1)));
    if (!b->segs.empty())
        FLS_HIP_CHECK(hipMemcpy(b->d_segs, b->segs.data(), sizeof(SegDev) * b->segs.size(), hipMemcpyHostToDevice));
    if (!b->d_queue) FLS_HIP_CHECK(hipMalloc((void**)&b->d_queue, 256));
    b->n_values = 0;
    b->n_vectors = 0;
    for (const auto& s : b->segs) {
        b->n_values += s.n_values;
        b->n_vectors += ceil_div(s.n_values, 1024);
    }
    return FLS_OK;
}

template <int BLOCK>
int launch_batch(fls_batch* b, uint64_t* d_counts, bool dynamic, uint32_t tile_off, uint32_t n_tiles, uint32_t grid,
                 hipStream_t s) {
    if (n_tiles == 0) return FLS_OK;
    hipLaunchKernelGGL(batch_kernel<BLOCK>, dim3(grid), dim3(BLOCK), 0, s, b->d_segs, b->d_tiles + tile_off, n_tiles,
                       d_counts, b->d_queue, (uint32_t)dynamic);
    return check_launch();
}

int check_tune(const fls_tune* t) {
    if (!t) return FLS_ERR_NULL;
    if (t->tile_vectors == 0 || t->tile_vectors > (1u << 20)) return FLS_ERR_ARG;
    if (t->block_threads != 256 && t->block_threads != 512) return FLS_ERR_ARG;
    if (t->schedule > FLS_SCHED_DYNAMIC) return FLS_ERR_ARG;
    return FLS_OK;
}

// zero counters + queue, then the persistent launch
int run_batch(fls_batch* b, const fls_tune* t, uint64_t* d_counts, hipStream_t s) {
    int st;
    if ((st = check_tune(t)) != FLS_OK) return st;
    if ((st = set_tiles(b, t->tile_vectors)) != FLS_OK) return st;
    if ((st = zero_u64(d_counts, b->segs.size(), s)) != FLS_OK) return st;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, b->d_queue, 1u);
    const bool dynamic = t->schedule == FLS_SCHED_DYNAMIC;
    uint32_t grid = 0;
    if (t->block_threads == 512) {
        if ((st = resident_grid_of((const void*)batch_kernel<512>, 512, 0, b->n_tiles, &grid)) != FLS_OK) return st;
        return launch_batch<512>(b, d_counts, dynamic, 0, b->n_tiles, grid, s);
    }
    if ((st = resident_grid_of((const void*)batch_kernel<256>, 256, 0, b->n_tiles, &grid)) != FLS_OK) return st;
    return launch_batch<256>(b, d_counts, dynamic, 0, b->n_tiles, grid, s);
}

// ---------------------------------------------------------------- autotune
std::string default_cache_path() {
    const char* env = getenv("FLS_AUTOTUNE_CACHE");
    if (env && *env) return env;
    const char* home = getenv("HOME");
    return std::string(home && *home ? home : "/tmp") + "/.cache/fls_mi355x/autotune.txt";
}

void mkdirs_for(const std::string& path) {
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') mkdir(path.substr(0, i).c_str(), 0755);
}

std::string batch_key(const fls_batch* b) {
    hipDeviceProp_t prop;
    char name[64] = "gfx";
    if (hipGetDeviceProperties(&prop, b->device) == hipSuccess) snprintf(name, sizeof name, "%s", prop.gcnArchName);
    for (char* c = name; *c; ++c)
        if (*c == ':' || *c == ' ') *c = '_';
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the segment shape
    for (const auto& s : b->segs) {
        const uint64_t parts[3] = {s.enc, s.bw, s.n_values};
        for (uint64_t p : parts) {
            h ^= p;
            h *= 1099511628211ull;
        }
    }
    uint32_t cus = 0;
    device_cus(&cus);
    char buf[160];
    snprintf(buf, sizeof buf, "%s-cu%u-seg%zu-vec%llu-%016llx", name, cus, b->segs.size(),
             (unsigned long long)b->n_vectors, (unsigned long long)h);
    return buf;
}

bool cache_lookup(const std::string& path, const std::string& key, fls_tune* t) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    char k[200];
    unsigned tile, block, sched;
    double ms;
    bool found = false;
    while (fscanf(f, "%199s %u %u %u %lf", k, &tile, &block, &sched, &ms) == 5) {
        if (key == k) {
            *t = fls_tune{tile, block, sched, 0};
            found = true;  // keep the last entry for the key
        }
    }
    fclose(f);
    return found && check_tune(t) == FLS_OK;
}

void cache_store(const std::string& path, const std::string& key, const fls_tune& t, double ms) {
    mkdirs_for(path);
    FILE* f = fopen(path.c_str(), "a");
    if (!f) return;
    fprintf(f, "%s %u %u %u %.6f\n", key.c_str(), t.tile_vectors, t.block_threads, t.schedule, ms);
    fclose(f);
}

}  // namespace
}  // namespace fls

extern "C" {

int fls_batch_create(const fls_segment* segs, uint32_t n_segments, fls_batch** out) {
    using namespace fls;
    FLS_RETURN_IF(!out || (n_segments && !segs), FLS_ERR_NULL);
    fls_batch* b = new fls_batch();
    FLS_HIP_CHECK(hipGetDevice(&b->device));
    for (uint32_t s = 0; s < n_segments; ++s) {
        const fls_segment& g = segs[s];
        const fls_column& c = g.col;
        ColDev tmp;
        // RLE metadata comes from the segment's rmeta or, when that is NULL, the column's aux
        const fls_rle_meta* rmeta = g.rmeta ? g.rmeta : reinterpret_cast<const fls_rle_meta*>(c.aux);
        fls_column cc = c;
        if (c.encoding == FLS_ENC_RLE) cc.aux = reinterpret_cast<const uint64_t*>(rmeta);
        const int st = to_dev_column(cc, g.n_values, &tmp, true);
        if (st) {
            delete b;
            return st;
        }
        const uint32_t tr = c.encoding == FLS_ENC_BITPACK ? (c.transposed ? 1u : 0u) : 1u;
        b->segs.push_back(
            SegDev{c.encoding, c.bw, tr, 0u, g.n_values, c.words, c.vmeta, c.bases, rmeta, c.aux, g.lo, g.hi});
        b->info.push_back(fls_segment_info{c.encoding, c.bw, g.n_values, g.lo, g.hi, c.n_words});
        if (c.encoding <= FLS_ENC_RLE) b->count_by_enc[c.encoding]++;
        b->encoded_bytes += 4ull * c.n_words;
    }
    int st = upload_segs(b);
    if (st) {
        fls_batch_destroy(b);
        return st;
    }
    *out = b;
    return FLS_OK;
}

void fls_c4_segment_params(uint64_t seed, uint32_t s, uint32_t log2_min_len, uint32_t log2_max_len,
                           uint32_t* encoding, uint32_t* bw, uint64_t* n_values) {
    const fls::SegParams p = fls::c4_params(seed, s, log2_min_len, log2_max_len);
    if (encoding) *encoding = p.enc;
    if (bw) *bw = p.bw;
    if (n_values) *n_values = p.len;
}

int fls_gen_c4_segment(uint64_t seed, uint32_t s, uint32_t log2_min_len, uint32_t log2_max_len, uint32_t* d_out,
                       void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!d_out, FLS_ERR_NULL);
    FLS_RETURN_IF(log2_min_len >= log2_max_len || log2_max_len > 30, FLS_ERR_ARG);
    const SegParams p = c4_params(seed, s, log2_min_len, log2_max_len);
    hipLaunchKernelGGL(gen_segment_kernel, dim3(grid_for(p.len)), dim3(256), 0, as_stream(stream), seed, s, p.enc,
                       p.bw, p.len, d_out);
    return check_launch();
}

int fls_batch_create_c4(uint64_t seed, uint32_t n_segments, uint32_t log2_min_len, uint32_t log2_max_len,
                        fls_batch** out, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!out, FLS_ERR_NULL);
    FLS_RETURN_IF(n_segments == 0 || log2_min_len >= log2_max_len || log2_max_len > 30, FLS_ERR_ARG);
    hipStream_t s = as_stream(stream);
    std::vector<SegParams> P(n_segments);
    uint64_t max_len = 0, meta_vec = 0;
    std::vector<uint64_t> meta_off(n_segments, 0);
    for (uint32_t i = 0; i < n_segments; ++i) {
        P[i] = c4_params(seed, i, log2_min_len, log2_max_len);
        max_len = P[i].len > max_len ? P[i].len : max_len;
        meta_off[i] = meta_vec;
        if (P[i].enc != FLS_ENC_BITPACK) meta_vec += ceil_div(P[i].len, 1024);
    }
    fls_batch* b = new fls_batch();
    int st = FLS_OK;
    uint32_t* d_raw = nullptr;
    void* d_scratch = nullptr;
    uint64_t* d_totals = nullptr;
    uint32_t* d_mm = nullptr;
    std::vector<uint64_t> totals(n_segments, 0);
    std::vector<uint32_t> mm(2ull * n_segments, 0);
    std::vector<uint64_t> word_off(n_segments, 0);
    uint64_t words_total = 0;
#define FLS_B(expr)                                  \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) {                      \
            st = record_hip_error(_e);               \
            goto fail;                               \
        }                                            \
    } while (0)
#define FLS_S(expr)          \
    do {                     \
        st = (expr);         \
        if (st) goto fail;   \
    } while (0)
    FLS_B(hipGetDevice(&b->device));
    FLS_B(hipMalloc((void**)&d_raw, 4ull * max_len));
    FLS_B(hipMalloc(&d_scratch, fls_encode_scratch_bytes(max_len)));
    FLS_B(hipMalloc((void**)&d_totals, 8ull * n_segments));
    FLS_B(hipMalloc((void**)&d_mm, 8ull * n_segments));
    FLS_B(hipMalloc((void**)&b->d_vmeta, 8ull * (meta_vec ? meta_vec : 1)));
    FLS_B(hipMalloc((void**)&b->d_bases, 128ull * (meta_vec ? meta_vec : 1)));
    FLS_B(hipMalloc((void**)&b->d_rmeta, 8ull * (meta_vec ? meta_vec : 1)));
    FLS_S(zero_u64(d_totals, n_segments, s));
    hipLaunchKernelGGL(mm_init_kernel, dim3(64), dim3(256), 0, s, d_mm, n_segments);
    // pass 1: generate, min/max, plan variable-width encodings
    for (uint32_t i = 0; i < n_segments; ++i) {
        hipLaunchKernelGGL(gen_segment_kernel, dim3(grid_for(P[i].len)), dim3(256), 0, s, seed, i, P[i].enc, P[i].bw,
                           P[i].len, d_raw);
        hipLaunchKernelGGL(minmax_kernel, dim3(grid_for(P[i].len)), dim3(256), 0, s, d_raw, P[i].len, d_mm + 2 * i);
        if (P[i].enc == FLS_ENC_DELTA)
            FLS_S(fls_delta_plan_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], b->d_bases + 32 * meta_off[i],
                                        d_totals + i, d_scratch, s));
        else if (P[i].enc == FLS_ENC_RLE)
            FLS_S(fls_rle_plan_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], b->d_rmeta + meta_off[i],
                                      b->d_bases + 32 * meta_off[i], d_totals + i, d_scratch, s));
        FLS_S(check_launch());
    }
    FLS_B(hipMemcpyAsync(totals.data(), d_totals, 8ull * n_segments, hipMemcpyDeviceToHost, s));
    FLS_B(hipMemcpyAsync(mm.data(), d_mm, 8ull * n_segments, hipMemcpyDeviceToHost, s));
    FLS_B(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < n_segments; ++i) {
        word_off[i] = words_total;
        words_total += P[i].enc == FLS_ENC_BITPACK ? ceil_div(P[i].len, 1024) * 32u * P[i].bw : totals[i];
    }
    FLS_B(hipMalloc((void**)&b->d_words, 4ull * (words_total ? words_total : 1)));
    // pass 2: generate again, pack
    for (uint32_t i = 0; i < n_segments; ++i) {
        hipLaunchKernelGGL(gen_segment_kernel, dim3(grid_for(P[i].len)), dim3(256), 0, s, seed, i, P[i].enc, P[i].bw,
                           P[i].len, d_raw);
        uint32_t* w = b->d_words + word_off[i];
        if (P[i].enc == FLS_ENC_BITPACK) FLS_S(fls_pack_column_device(d_raw, P[i].len, P[i].bw, 0, 0, w, s));
        else if (P[i].enc == FLS_ENC_DELTA) FLS_S(fls_delta_pack_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], w, s));
        else FLS_S(fls_rle_pack_device(d_raw, P[i].len, b->d_vmeta + meta_off[i], b->d_rmeta + meta_off[i], w, s));
        const uint32_t mn = mm[2 * i], mx = mm[2 * i + 1];
        const uint64_t range = (uint64_t)(mx - mn);
        const uint32_t lo = mn + (uint32_t)(range / 4), hi = mn + (uint32_t)(range * 3 / 4);
        const bool meta = P[i].enc != FLS_ENC_BITPACK;
        b->segs.push_back(SegDev{P[i].enc, P[i].bw, P[i].enc == FLS_ENC_BITPACK ? 0u : 1u, 0u, P[i].len, w,
                                 meta ? b->d_vmeta + meta_off[i] : nullptr,
                                 meta ? b->d_bases + 32 * meta_off[i] : nullptr,
                                 P[i].enc == FLS_ENC_RLE ? b->d_rmeta + meta_off[i] : nullptr, nullptr, lo, hi});
        const uint64_t seg_words = P[i].enc == FLS_ENC_BITPACK ? ceil_div(P[i].len, 1024) * 32u * P[i].bw : totals[i];
        b->info.push_back(fls_segment_info{P[i].enc, P[i].bw, P[i].len, lo, hi, seg_words});
        const uint64_t nv = ceil_div(P[i].len, 1024);
        b->encoded_bytes += 4ull * seg_words + (meta ? nv * (8u + 128u) : 0u) + (P[i].enc == FLS_ENC_RLE ? nv * 8u : 0u);
        b->count_by_enc[P[i].enc]++;
    }
    FLS_B(hipStreamSynchronize(s));
    FLS_S(upload_segs(b));
    (void)hipFree(d_raw);
    (void)hipFree(d_scratch);
    (void)hipFree(d_totals);
    (void)hipFree(d_mm);
    *out = b;
    return FLS_OK;
fail:
    if (d_raw) (void)hipFree(d_raw);
    if (d_scratch) (void)hipFree(d_scratch);
    if (d_totals) (void)hipFree(d_totals);
    if (d_mm) (void)hipFree(d_mm);
    fls_batch_destroy(b);
    return st;
#undef FLS_B
#undef FLS_S
}

int fls_batch_get_info(const fls_batch* b, fls_batch_info* info) {
    if (!b || !info) return FLS_ERR_NULL;
    info->n_segments = (uint32_t)b->segs.size();
    info->n_values = b->n_values;
    info->n_vectors = b->n_vectors;
    info->encoded_bytes = b->encoded_bytes;
    info->n_bitpack = b->count_by_enc[0];
    info->n_delta = b->count_by_enc[1];
    info->n_rle = b->count_by_enc[2];
    return FLS_OK;
}

int fls_batch_segment(const fls_batch* b, uint32_t s, fls_segment_info* info) {
    if (!b || !info) return FLS_ERR_NULL;
    if (s >= b->info.size()) return FLS_ERR_ARG;
    *info = b->info[s];
    return FLS_OK;
}

int fls_batch_segment_column(const fls_batch* b, uint32_t s, fls_segment* seg) {
    if (!b || !seg) return FLS_ERR_NULL;
    if (s >= b->segs.size()) return FLS_ERR_ARG;
    const fls::SegDev& d = b->segs[s];
    fls_segment g{};
    g.col.encoding = d.enc;
    g.col.bw = d.bw;
    g.col.base = 0;
    g.col.transposed = d.transposed;
    g.col.n_words = b->info[s].n_words;
    g.col.words = d.words;
    g.col.vmeta = d.vmeta;
    g.col.bases = d.bases;
    g.col.aux = d.enc == FLS_ENC_RLE ? reinterpret_cast<const uint64_t*>(d.rmeta) : d.aux;
    g.rmeta = d.rmeta;
    g.n_values = d.n_values;
    g.lo = d.lo;
    g.hi = d.hi;
    *seg = g;
    return FLS_OK;
}

int fls_batch_run(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream) {
    if (!b || !d_counts) return FLS_ERR_NULL;
    return fls::run_batch(b, tune, d_counts, fls::as_stream(stream));
}

int fls_batch_run_per_segment(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream) {
    using namespace fls;
    if (!b || !d_counts) return FLS_ERR_NULL;
    int st;
    if ((st = check_tune(tune)) != FLS_OK) return st;
    if ((st = set_tiles(b, tune->tile_vectors)) != FLS_OK) return st;
    hipStream_t s = as_stream(stream);
    if ((st = zero_u64(d_counts, b->segs.size(), s)) != FLS_OK) return st;
    for (uint32_t i = 0; i < b->segs.size(); ++i) {
        // one launch per segment: one block per tile of the segment
        const uint32_t nt = b->seg_tile_cnt[i];
        st = tune->block_threads == 512 ? launch_batch<512>(b, d_counts, false, b->seg_tile_off[i], nt, nt, s)
                                        : launch_batch<256>(b, d_counts, false, b->seg_tile_off[i], nt, nt, s);
        if (st) return st;
    }
    return FLS_OK;
}

int fls_batch_graph_launch(fls_batch* b, const fls_tune* tune, uint64_t* d_counts, void* stream) {
    using namespace fls;
    if (!b || !d_counts) return FLS_ERR_NULL;
    int st;
    if ((st = check_tune(tune)) != FLS_OK) return st;
    const bool same = b->exec && b->graph_counts == d_counts && b->graph_tune.tile_vectors == tune->tile_vectors &&
                      b->graph_tune.block_threads == tune->block_threads && b->graph_tune.schedule == tune->schedule;
    if (!same) {
        if (b->exec) FLS_HIP_CHECK(hipGraphExecDestroy(b->exec));
        if (b->graph) FLS_HIP_CHECK(hipGraphDestroy(b->graph));
        b->exec = nullptr;
        b->graph = nullptr;
        if ((st = set_tiles(b, tune->tile_vectors)) != FLS_OK) return st;  // uploads before capture
        if (!b->cap_stream) FLS_HIP_CHECK(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        FLS_HIP_CHECK(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        st = run_batch(b, tune, d_counts, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (st) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return record_hip_error(e);
        b->graph = g;
        FLS_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->graph_counts = d_counts;
        b->graph_tune = *tune;
    }
    FLS_HIP_CHECK(hipGraphLaunch(b->exec, as_stream(stream)));
    return FLS_OK;
}

int fls_batch_autotune(fls_batch* b, const char* cache_path, int force, uint64_t* d_counts, fls_tune* best,
                       double* best_ms, void* stream) {
    using namespace fls;
    if (!b || !best || !d_counts) return FLS_ERR_NULL;
    const std::string path = cache_path && *cache_path ? cache_path : default_cache_path();
    const std::string key = batch_key(b);
    if (!force && cache_lookup(path, key, best)) {
        if (best_ms) *best_ms = -1.0;
        return FLS_OK;
    }
    hipStream_t s = as_stream(stream);
    hipEvent_t e0, e1;
    FLS_HIP_CHECK(hipEventCreate(&e0));
    FLS_HIP_CHECK(hipEventCreate(&e1));
    const uint32_t tiles[] = {8, 16, 32, 64, 128, 256, 512};
    const uint32_t blocks[] = {256, 512};
    double best_t = 1e30;
    int st = FLS_OK;
    for (uint32_t tv : tiles) {
        for (uint32_t bt : blocks) {
            for (uint32_t sched = 0; sched <= FLS_SCHED_DYNAMIC; ++sched) {
                fls_tune t{tv, bt, sched, 0};
                if ((st = run_batch(b, &t, d_counts, s)) != FLS_OK) break;  // warm-up (uploads tiles)
                double tmin = 1e30;
                for (int rep = 0; rep < 3 && st == FLS_OK; ++rep) {
                    (void)hipEventRecord(e0, s);
                    st = run_batch(b, &t, d_counts, s);
                    (void)hipEventRecord(e1, s);
                    (void)hipEventSynchronize(e1);
                    float ms = 0;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    tmin = ms < tmin ? ms : tmin;
                }
                if (st) break;
                if (tmin < best_t) {
                    best_t = tmin;
                    *best = t;
                }
            }
            if (st) break;
        }
        if (st) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    cache_store(path, key, *best, best_t);
    if (best_ms) *best_ms = best_t;
    return FLS_OK;
}

void fls_batch_destroy(fls_batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_segs) (void)hipFree(b->d_segs);
    if (b->d_queue) (void)hipFree(b->d_queue);
    if (b->d_words) (void)hipFree(b->d_words);
    if (b->d_vmeta) (void)hipFree(b->d_vmeta);
    if (b->d_bases) (void)hipFree(b->d_bases);
    if (b->d_rmeta) (void)hipFree(b->d_rmeta);
    delete b;
}

}  // extern "C"
