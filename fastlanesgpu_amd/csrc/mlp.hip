// The learned encoding selector: an MLP 16 -> 64 -> 64 -> 40 (ReLU hidden
// layers, softmax over the 40 classes), BF16 operands on the CDNA4 matrix
// cores with FP32 accumulation -- v_mfma_f32_16x16x32_bf16 for the products
// whose inner dimension is 32 or more, v_mfma_f32_16x16x16_bf16 for the
// 16-deep ones (the 16 input features, the 16 samples of a tile) -- trained
// on the GPU by its own forward/backward kernels and an FP32 Adam step.
//
// Parameters (FP32 master copy, and a BF16 copy the kernels read), in this
// order: W1[16][64], b1[64], W2[64][64], b2[64], W3[64][40], b3[40]
// (W[in][out]); 7,848 values.
//
// A wave computes 16 samples at a time (one 16-row MFMA tile).  Training:
// weights staged once per block in LDS as BF16, activations through per-wave
// LDS tiles, so every product -- including the transposed ones of the
// backward pass -- reads its operands with plain strides.  Inference keeps
// activations in registers between layers (see mlp_forward_kernel).
//
// MFMA operand layouts used here (lane L of the wave):
//   16x16x16: A[L % 16][4*(L/16) + i], B[4*(L/16) + i][L % 16], i < 4
//   16x16x32: A[L % 16][8*(L/16) + i], B[8*(L/16) + i][L % 16], i < 8
//   D (16 x 16, both): D[4*(L/16) + i][L % 16], i < 4
// fls_mlp_mfma_selftest checks both against a host matmul.
#include <math.h>

#include "fls/device/wave.hpp"
#include "fls_runtime.hpp"

namespace fls {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kIn = 16, kH = 64, kOut = 40, kOutPad = 48;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kOut, kParams = oB3 + kOut;
static_assert(kParams == 7848, "parameter count");

// round-to-nearest-even FP32 -> BF16 (bit pattern), NaN kept quiet
__host__ __device__ inline uint16_t to_bf16(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t h) {
    const uint32_t u = (uint32_t)h << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

__device__ __forceinline__ int lane64() { return (int)(threadIdx.x & 63u); }

// element (r, c) at p[r*rs + c*cs]
struct View {
    const uint16_t* p;
    int rs, cs;
    __device__ uint16_t at(int r, int c) const { return p[r * rs + c * cs]; }
};

__device__ __forceinline__ bf16x4 frag_a(const View& A, int m0, int k0) {
    const int L = lane64(), m = m0 + (L & 15), k = k0 + 4 * (L >> 4);
    return bf16x4{(short)A.at(m, k), (short)A.at(m, k + 1), (short)A.at(m, k + 2), (short)A.at(m, k + 3)};
}
__device__ __forceinline__ bf16x4 frag_b(const View& B, int k0, int n0) {
    const int L = lane64(), n = n0 + (L & 15), k = k0 + 4 * (L >> 4);
    return bf16x4{(short)B.at(k, n), (short)B.at(k + 1, n), (short)B.at(k + 2, n), (short)B.at(k + 3, n)};
}
__device__ __forceinline__ s16x8 frag_a8(const View& A, int m0, int k0) {
    const int L = lane64(), m = m0 + (L & 15), k = k0 + 8 * (L >> 4);
    s16x8 f;
#pragma unroll
    for (int i = 0; i < 8; ++i) f[i] = (short)A.at(m, k + i);
    return f;
}
__device__ __forceinline__ s16x8 frag_b8(const View& B, int k0, int n0) {
    const int L = lane64(), n = n0 + (L & 15), k = k0 + 8 * (L >> 4);
    s16x8 f;
#pragma unroll
    for (int i = 0; i < 8; ++i) f[i] = (short)B.at(k + i, n);
    return f;
}
__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma32(s16x8 a, s16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}
// C tile (m0, n0) of A[M x K] * B[K x N], K a multiple of 16: 32-deep steps
// when K is a multiple of 32, else 16-deep ones.  One product never mixes the
// two shapes on an accumulator: a 16x16x32 step followed by a 16x16x16 one
// (K = 48) gave wrong sums in the gradient test on gfx950.
__device__ __forceinline__ f32x4 tile_mm(const View& A, const View& B, int m0, int n0, int K) {
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if (K % 32 == 0) {
        for (int k0 = 0; k0 < K; k0 += 32) acc = mfma32(frag_a8(A, m0, k0), frag_b8(B, k0, n0), acc);
    } else {
        for (int k0 = 0; k0 < K; k0 += 16) acc = mfma(frag_a(A, m0, k0), frag_b(B, k0, n0), acc);
    }
    return acc;
}
// row of item i of the D tile held by this lane, and its column
__device__ __forceinline__ int d_row(int i) { return 4 * (lane64() >> 4) + i; }
__device__ __forceinline__ int d_col() { return lane64() & 15; }

// Per-block LDS: BF16 weights (transposed copies for the backward products
// are not needed: strides do it) + per-wave activation tiles.
struct WaveTiles {
    uint16_t x[16 * kIn];        // X tile, row-major [16][16]
    uint16_t h1[16 * kH];        // relu(X W1 + b1)
    uint16_t h2[16 * kH];        // relu(H1 W2 + b2)
    uint16_t dz[16 * kOutPad];   // dL/dZ (bf16), padded columns zero
    uint16_t dh2[16 * kH];       // dL/dH2 (masked)
    uint16_t dh1[16 * kH];       // dL/dH1 (masked)
    float z[16 * kOutPad];       // logits
};
struct MlpLds {
    uint16_t w1[kIn * kH], w2[kH * kH], w3[kH * kOutPad];
    float b1[kH], b2[kH], b3[kOutPad];
    WaveTiles t[1];  // the gradient kernel runs one wave per block
};

__device__ void stage_weights(const uint16_t* __restrict wb, MlpLds& s) {
    for (int i = threadIdx.x; i < kIn * kH; i += blockDim.x) s.w1[i] = wb[oW1 + i];
    for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) s.w2[i] = wb[oW2 + i];
    for (int i = threadIdx.x; i < kH * kOutPad; i += blockDim.x) {
        const int r = i / kOutPad, c = i % kOutPad;
        s.w3[i] = c < kOut ? wb[oW3 + r * kOut + c] : (uint16_t)0;
    }
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        s.b1[i] = from_bf16(wb[oB1 + i]);
        s.b2[i] = from_bf16(wb[oB2 + i]);
    }
    for (int i = threadIdx.x; i < kOutPad; i += blockDim.x) s.b3[i] = i < kOut ? from_bf16(wb[oB3 + i]) : 0.f;
}

// forward of one 16-sample tile (rows >= valid are zero inputs)
__device__ void forward_tile(const float* __restrict feat, uint64_t s0, uint64_t n, MlpLds& s, WaveTiles& t) {
    const int L = lane64();
    for (int i = L; i < 16 * kIn; i += 64) {
        const uint64_t row = s0 + i / kIn;
        t.x[i] = to_bf16(row < n ? feat[row * kIn + (i % kIn)] : 0.f);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const View X{t.x, kIn, 1}, W1{s.w1, kH, 1}, W2{s.w2, kH, 1}, W3{s.w3, kOutPad, 1};
    for (int n0 = 0; n0 < kH; n0 += 16) {
        const f32x4 acc = tile_mm(X, W1, 0, n0, kIn);
        for (int i = 0; i < 4; ++i) {
            const float h = acc[i] + s.b1[n0 + d_col()];
            t.h1[d_row(i) * kH + n0 + d_col()] = to_bf16(h > 0.f ? h : 0.f);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const View H1{t.h1, kH, 1};
    for (int n0 = 0; n0 < kH; n0 += 16) {
        const f32x4 acc = tile_mm(H1, W2, 0, n0, kH);
        for (int i = 0; i < 4; ++i) {
            const float h = acc[i] + s.b2[n0 + d_col()];
            t.h2[d_row(i) * kH + n0 + d_col()] = to_bf16(h > 0.f ? h : 0.f);
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const View H2{t.h2, kH, 1};
    for (int n0 = 0; n0 < kOutPad; n0 += 16) {
        const f32x4 acc = tile_mm(H2, W3, 0, n0, kH);
        for (int i = 0; i < 4; ++i) t.z[d_row(i) * kOutPad + n0 + d_col()] = acc[i] + s.b3[n0 + d_col()];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------- inference
// The forward holds the weights and the activations in registers and runs
// the transposed products D = W^T X^T: a wave's 16 samples are the columns of
// every product, the hidden units its rows.  A D tile leaves lane L holding
// rows 4(L/16) + i (i < 4) -- units 16t + 4(L/16) + i of tile t -- of sample
// L % 16, and the next layer's 32-deep B fragment wants 8 k of that sample.
// The k order of a product is free, so k slot 32c + 8(L/16) + j of chunk c is
// taken to be unit 16(2c + j/4) + 4(L/16) + j%4: exactly the lane's own
// values of tiles 2c and 2c + 1, converted to BF16 in place.  The A fragments
// (W^T rows in that k order) are built once per block; the biases initialise
// the accumulators.  The 40 classes sit in 48 output rows so that
// row 16t + r is encoding r % 4, width bucket 3(r / 4) + t (buckets 10 and 11:
// zero weights, -inf bias): a lane then holds three buckets of every encoding
// and the softmax reduces within the lane and over the 4 lanes of a sample.
__device__ __forceinline__ uint32_t unit_of_slot(uint32_t c, uint32_t q, uint32_t j) {
    return 16u * (2u * c + (j >> 2)) + 4u * q + (j & 3u);
}
__device__ __forceinline__ int class_of_row(int t, int r) {
    const int w = 3 * (r >> 2) + t;
    return w < 10 ? 10 * (r & 3) + w : -1;
}

typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
// two FP32 values to a BF16 pair, round to nearest even (v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint32_t cvt2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2v));
}
// the B fragment of chunk c from the ReLU of D tiles 2c and 2c + 1
__device__ __forceinline__ s16x8 relu_frag(const f32x4& lo, const f32x4& hi) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 u = {cvt2(fmaxf(lo[0], 0.f), fmaxf(lo[1], 0.f)), cvt2(fmaxf(lo[2], 0.f), fmaxf(lo[3], 0.f)),
                     cvt2(fmaxf(hi[0], 0.f), fmaxf(hi[1], 0.f)), cvt2(fmaxf(hi[2], 0.f), fmaxf(hi[3], 0.f))};
    return __builtin_bit_cast(s16x8, u);
}

// A-fragment images, one entry per (tile, chunk, lane): a wave reads the
// second layer's into registers once and the first and last layers' from LDS
// per tile (VGPRs for all three spill at four waves per SIMD).
struct alignas(16) InferLds {
    bf16x4 f1[4][64];     // A[16t + m][4q + i] = W1[4q + i][16t + m]
    s16x8 f2[4][2][64];   // A[16t + m][slot (c, q, j)] = W2[unit][16t + m]
    s16x8 f3[3][2][64];   // A[16t + m][slot (c, q, j)] = W3[unit][class of row 16t + m]
    float b1[kH], b2[kH], b3[kOutPad];  // b3 by output row, -inf on the padding rows
};

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void mlp_forward_kernel(
    const uint16_t* __restrict wb, const float* __restrict feat, uint64_t n, uint8_t* __restrict cls,
    float* __restrict conf, float* __restrict logits) {
    __shared__ InferLds s;
    for (int e = threadIdx.x; e < 4 * 64; e += blockDim.x) {
        const int t = e >> 6, lq = (e >> 4) & 3, lm = e & 15;
#pragma unroll
        for (int i = 0; i < 4; ++i) s.f1[t][e & 63][i] = (short)wb[oW1 + (4 * lq + i) * kH + 16 * t + lm];
    }
    for (int e = threadIdx.x; e < 8 * 64; e += blockDim.x) {
        const int t = e >> 7, c = (e >> 6) & 1, lq = (e >> 4) & 3, lm = e & 15;
#pragma unroll
        for (int j = 0; j < 8; ++j) s.f2[t][c][e & 63][j] = (short)wb[oW2 + unit_of_slot(c, lq, j) * kH + 16 * t + lm];
    }
    for (int e = threadIdx.x; e < 6 * 64; e += blockDim.x) {
        const int t = e >> 7, c = (e >> 6) & 1, lq = (e >> 4) & 3, lm = e & 15;
        const int cl = class_of_row(t, lm);
#pragma unroll
        for (int j = 0; j < 8; ++j)
            s.f3[t][c][e & 63][j] = cl >= 0 ? (short)wb[oW3 + unit_of_slot(c, lq, j) * kOut + cl] : (short)0;
    }
    for (int i = threadIdx.x; i < kH; i += blockDim.x) {
        s.b1[i] = from_bf16(wb[oB1 + i]);
        s.b2[i] = from_bf16(wb[oB2 + i]);
    }
    for (int i = threadIdx.x; i < kOutPad; i += blockDim.x) {
        const int c = class_of_row(i >> 4, i & 15);
        s.b3[i] = c >= 0 ? from_bf16(wb[oB3 + c]) : -INFINITY;
    }
    __syncthreads();
    const int L = lane64(), q = L >> 4, m = L & 15;
    s16x8 a2[4][2];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int c = 0; c < 2; ++c) a2[t][c] = s.f2[t][c][L];
    const uint64_t n_tiles = ceil_div(n, 16);
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    for (uint64_t tile = (uint64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < n_tiles; tile += waves) {
        const uint64_t row = tile * 16 + m;
        // an index the compiler cannot prove loop-invariant: the first and
        // last layers' fragments are read from LDS per tile, not hoisted
        // into registers (which spill)
        int Lx = L;
        asm volatile("" : "+v"(Lx));
        f32x4 f = {0.f, 0.f, 0.f, 0.f};
        if (row < n) f = *reinterpret_cast<const f32x4*>(feat + row * kIn + 4 * q);
        typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
        const bf16x4 x = __builtin_bit_cast(bf16x4, u32x2{cvt2(f[0], f[1]), cvt2(f[2], f[3])});
        f32x4 d[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) d[t] = mfma(s.f1[t][Lx], x, *reinterpret_cast<const f32x4*>(s.b1 + 16 * t + 4 * q));
        const s16x8 h0 = relu_frag(d[0], d[1]), h1 = relu_frag(d[2], d[3]);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            d[t] = mfma32(a2[t][1], h1, mfma32(a2[t][0], h0, *reinterpret_cast<const f32x4*>(s.b2 + 16 * t + 4 * q)));
        const s16x8 g0 = relu_frag(d[0], d[1]), g1 = relu_frag(d[2], d[3]);
        f32x4 z[3];
#pragma unroll
        for (int t = 0; t < 3; ++t)
            z[t] = mfma32(s.f3[t][1][Lx], g1,
                          mfma32(s.f3[t][0][Lx], g0, *reinterpret_cast<const f32x4*>(s.b3 + 16 * t + 4 * q)));
        // z[t][i]: encoding i, bucket 3q + t.  Per encoding the largest logit
        // and its bucket (the lowest bucket on a tie), over the lane's three
        // and then over the sample's four lanes
        float gm[4];
        int gb[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            gm[i] = z[0][i];
            gb[i] = 3 * q;
#pragma unroll
            for (int t = 1; t < 3; ++t)
                if (z[t][i] > gm[i]) {
                    gm[i] = z[t][i];
                    gb[i] = 3 * q + t;
                }
        }
#pragma unroll
        for (int off = 16; off <= 32; off <<= 1)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float om = __shfl_xor(gm[i], off, 64);
                const int ob = __shfl_xor(gb[i], off, 64);
                if (om > gm[i] || (om == gm[i] && ob < gb[i])) {
                    gm[i] = om;
                    gb[i] = ob;
                }
            }
        const float rmax = fmaxf(fmaxf(gm[0], gm[1]), fmaxf(gm[2], gm[3]));
        float mass[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            mass[i] = __expf(z[0][i] - rmax) + __expf(z[1][i] - rmax) + __expf(z[2][i] - rmax);
            mass[i] += __shfl_xor(mass[i], 16, 64);
            mass[i] += __shfl_xor(mass[i], 32, 64);
        }
        // the encoding with the largest mass wins (the first on a tie); its
        // most likely bucket is the class, its share of the mass the confidence
        int be = 0, bucket = gb[0];
        float mb = mass[0];
#pragma unroll
        for (int i = 1; i < 4; ++i)
            if (mass[i] > mb) {
                mb = mass[i];
                be = i;
                bucket = gb[i];
            }
        if (row < n) {
            if (q == 0) {
                cls[row] = (uint8_t)(10 * be + bucket);
                if (conf) conf[row] = mb / (mass[0] + mass[1] + mass[2] + mass[3]);
            }
            if (logits)
#pragma unroll
                for (int t = 0; t < 3; ++t)
                    if (3 * q + t < 10)
#pragma unroll
                        for (int i = 0; i < 4; ++i) logits[row * kOut + 10 * i + 3 * q + t] = z[t][i];
        }
    }
}

// -------------------------------------------------------------- gradients
// Softmax cross-entropy, scaled by inv_batch.  Every wave accumulates the
// gradients of its tiles in registers and writes them, with its summed loss,
// to its own row of `part` (kPart floats); grad_reduce_kernel then adds the
// rows in wave order, so a step's gradient is bit-reproducible.  One wave per
// block and one 16-sample tile per wave up to kGradMaxWaves waves: a 4,096-
// sample batch is 256 blocks, one per CU (the kernel's chain of dependent
// products and LDS round trips is latency-bound, so spreading the tiles over
// the CUs is what shortens it).
constexpr int kPart = kParams + 8;  // [kParams] = loss, rest padding
constexpr int kGradMaxWaves = 1024;

__global__ __launch_bounds__(64) void mlp_grad_kernel(const uint16_t* __restrict wb, const float* __restrict feat,
                                                       const uint8_t* __restrict label, uint64_t n, float inv_batch,
                                                       float* __restrict part) {
    __shared__ MlpLds s;
    stage_weights(wb, s);
    __syncthreads();
    WaveTiles& t = s.t[0];
    const int L = lane64();
    // register accumulators of this wave's weight gradients (D-tile layout)
    f32x4 gw1[4], gw2[16], gw3[12];
    for (int i = 0; i < 4; ++i) gw1[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 16; ++i) gw2[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int i = 0; i < 12; ++i) gw3[i] = f32x4{0.f, 0.f, 0.f, 0.f};
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, lsum = 0.f;  // lane L owns column L of each bias
    const uint64_t n_tiles = ceil_div(n, 16);
    for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
        const uint64_t s0 = tile * 16;
        forward_tile(feat, s0, n, s, t);
        // softmax + dZ = (p - onehot) * inv_batch, rows past n contribute nothing
        if (L < 16) {
            const uint64_t row = s0 + L;
            float* z = t.z + L * kOutPad;
            uint16_t* dz = t.dz + L * kOutPad;
            if (row < n) {
                float mx = z[0];
                for (int c = 1; c < kOut; ++c) mx = fmaxf(mx, z[c]);
                float sum = 0.f;
                for (int c = 0; c < kOut; ++c) sum += expf(z[c] - mx);
                const int y = label[row];
                lsum += logf(sum) + mx - z[y];
                for (int c = 0; c < kOut; ++c) {
                    const float p = expf(z[c] - mx) / sum;
                    dz[c] = to_bf16((p - (c == y ? 1.f : 0.f)) * inv_batch);
                }
            } else {
                for (int c = 0; c < kOut; ++c) dz[c] = 0;
            }
            for (int c = kOut; c < kOutPad; ++c) dz[c] = 0;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const View DZ{t.dz, kOutPad, 1};
        const View H2T{t.h2, 1, kH}, H1T{t.h1, 1, kH}, XT{t.x, 1, kIn};
        // dW3 += H2^T dZ   (64 x 48, K = 16 samples)
        for (int mi = 0; mi < 4; ++mi)
            for (int ni = 0; ni < 3; ++ni)
                gw3[mi * 3 + ni] = mfma(frag_a(H2T, 16 * mi, 0), frag_b(DZ, 0, 16 * ni), gw3[mi * 3 + ni]);
        // dH2 = dZ W3^T, masked by H2 > 0
        const View W3T{s.w3, 1, kOutPad};
        for (int n0 = 0; n0 < kH; n0 += 16) {
            const f32x4 acc = tile_mm(DZ, W3T, 0, n0, kOutPad);
            for (int i = 0; i < 4; ++i) {
                const int r = d_row(i), c = n0 + d_col();
                t.dh2[r * kH + c] = to_bf16(from_bf16(t.h2[r * kH + c]) > 0.f ? acc[i] : 0.f);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const View DH2{t.dh2, kH, 1};
        // dW2 += H1^T dH2
        for (int mi = 0; mi < 4; ++mi)
            for (int ni = 0; ni < 4; ++ni)
                gw2[mi * 4 + ni] = mfma(frag_a(H1T, 16 * mi, 0), frag_b(DH2, 0, 16 * ni), gw2[mi * 4 + ni]);
        // dH1 = dH2 W2^T, masked by H1 > 0
        const View W2T{s.w2, 1, kH};
        for (int n0 = 0; n0 < kH; n0 += 16) {
            const f32x4 acc = tile_mm(DH2, W2T, 0, n0, kH);
            for (int i = 0; i < 4; ++i) {
                const int r = d_row(i), c = n0 + d_col();
                t.dh1[r * kH + c] = to_bf16(from_bf16(t.h1[r * kH + c]) > 0.f ? acc[i] : 0.f);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const View DH1{t.dh1, kH, 1};
        // dW1 += X^T dH1  (16 x 64)
        for (int ni = 0; ni < 4; ++ni) gw1[ni] = mfma(frag_a(XT, 0, 0), frag_b(DH1, 0, 16 * ni), gw1[ni]);
        // bias gradients: column sums over the tile's 16 rows
        for (int r = 0; r < 16; ++r) {
            gb1 += from_bf16(t.dh1[r * kH + L]);
            gb2 += from_bf16(t.dh2[r * kH + L]);
            if (L < kOutPad) gb3 += from_bf16(t.dz[r * kOutPad + L]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // this wave's row of partial sums: every parameter written exactly once
    float* p = part + (uint64_t)blockIdx.x * kPart;
    for (int mi = 0; mi < 4; ++mi)
        for (int ni = 0; ni < 3; ++ni)
            for (int i = 0; i < 4; ++i) {
                const int r = 16 * mi + d_row(i), c = 16 * ni + d_col();
                if (c < kOut) p[oW3 + r * kOut + c] = gw3[mi * 3 + ni][i];
            }
    for (int mi = 0; mi < 4; ++mi)
        for (int ni = 0; ni < 4; ++ni)
            for (int i = 0; i < 4; ++i) p[oW2 + (16 * mi + d_row(i)) * kH + 16 * ni + d_col()] = gw2[mi * 4 + ni][i];
    for (int ni = 0; ni < 4; ++ni)
        for (int i = 0; i < 4; ++i) p[oW1 + d_row(i) * kH + 16 * ni + d_col()] = gw1[ni][i];
    p[oB1 + L] = gb1;
    p[oB2 + L] = gb2;
    if (L < kOut) p[oB3 + L] = gb3;
    const float wl = dev::wave_sum(lsum);
    if (L == 0) p[kParams] = wl;
}

// grad[i] += sum over waves of part[w][i]; *loss likewise (i = kParams).  A
// block serves 16 parameters with 16 slices each: slice s adds waves
// s, s+16, s+32, ... in order, then slice sums are added in slice order --
// a fixed order, so the result is bit-reproducible.
__global__ __launch_bounds__(256) void grad_reduce_kernel(const float* __restrict part, uint32_t n_waves,
                                                          float* __restrict grad, float* __restrict loss) {
    __shared__ float red[16][17];
    const int col = threadIdx.x & 15, sl = threadIdx.x >> 4;
    const int i = blockIdx.x * 16 + col;
    float acc = 0.f;
    if (i <= kParams)
        for (uint32_t w = sl; w < n_waves; w += 16) acc += part[(uint64_t)w * kPart + i];
    red[sl][col] = acc;
    __syncthreads();
    if (sl == 0 && i <= kParams) {
        float t = 0.f;
        for (int k = 0; k < 16; ++k) t += red[k][col];
        if (i < kParams) grad[i] += t;
        else if (loss) *loss += t;
    }
}

uint32_t grad_blocks(uint64_t n) {
    const uint64_t tiles = ceil_div(n, 16);
    return (uint32_t)(tiles < 1 ? 1 : tiles > kGradMaxWaves ? kGradMaxWaves : tiles);
}

// Adam on the FP32 master weights; refreshes the BF16 copy and zeroes grad
__global__ void mlp_adam_kernel(float* __restrict w, float* __restrict m, float* __restrict v, float* __restrict grad,
                                uint16_t* __restrict wb, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kParams; i += gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        wb[i] = to_bf16(wi);
        grad[i] = 0.f;
    }
}

__global__ void to_bf16_kernel(const float* __restrict w, uint16_t* __restrict wb, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) wb[i] = to_bf16(w[i]);
}

// one 16x16x16 tile through the same fragment path, for the layout check
__global__ void mfma_selftest_kernel(const uint16_t* __restrict a, const uint16_t* __restrict b, float* __restrict d) {
    const View A{a, 16, 1}, B{b, 16, 1};
    const f32x4 acc = tile_mm(A, B, 0, 0, 16);
    for (int i = 0; i < 4; ++i) d[d_row(i) * 16 + d_col()] = acc[i];
}

// one 16x16x32 tile (A 16 x 32, B 32 x 16, row-major) through the K = 32 path
__global__ void mfma32_selftest_kernel(const uint16_t* __restrict a, const uint16_t* __restrict b,
                                       float* __restrict d) {
    const View A{a, 32, 1}, B{b, 16, 1};
    const f32x4 acc = tile_mm(A, B, 0, 0, 32);
    for (int i = 0; i < 4; ++i) d[d_row(i) * 16 + d_col()] = acc[i];
}

uint32_t grid_of(uint64_t n) {
    const uint64_t tiles = ceil_div(n, 16);
    const uint64_t g = ceil_div(tiles, 4);
    return (uint32_t)(g < 1024 ? (g ? g : 1) : 1024);
}

}  // namespace
}  // namespace fls

extern "C" {

uint32_t fls_mlp_param_count(void) { return fls::kParams; }

int fls_mlp_to_bf16(const float* d_w, uint16_t* d_wb, void* stream) {
    if (!d_w || !d_wb) return FLS_ERR_NULL;
    hipLaunchKernelGGL(fls::to_bf16_kernel, dim3(32), dim3(256), 0, fls::as_stream(stream), d_w, d_wb, fls::kParams);
    return fls::check_launch();
}

int fls_mlp_forward(const uint16_t* d_wb, const float* d_feat, uint64_t n, uint8_t* d_class, float* d_conf,
                    float* d_logits, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_wb || !d_feat || !d_class, FLS_ERR_NULL);
    FLS_RETURN_IF(reinterpret_cast<uintptr_t>(d_feat) % 16u, FLS_ERR_ARG);  // a lane reads 4 features at once
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid_of(n)), dim3(256), 0, as_stream(stream), d_wb, d_feat, n, d_class,
                       d_conf, d_logits);
    return check_launch();
}

uint64_t fls_mlp_grad_workspace_bytes(uint64_t n) {
    return (uint64_t)fls::grad_blocks(n) * fls::kPart * sizeof(float);
}

int fls_mlp_grad(const uint16_t* d_wb, const float* d_feat, const uint8_t* d_label, uint64_t n, float inv_batch,
                 void* d_work, float* d_grad, float* d_loss, void* stream) {
    using namespace fls;
    if (n == 0) return FLS_OK;
    FLS_RETURN_IF(!d_wb || !d_feat || !d_label || !d_work || !d_grad, FLS_ERR_NULL);
    const uint32_t blocks = grad_blocks(n);
    float* part = static_cast<float*>(d_work);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(blocks), dim3(64), 0, as_stream(stream), d_wb, d_feat, d_label, n,
                       inv_batch, part);
    if (const int rc = check_launch()) return rc;
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((uint32_t)ceil_div(kParams + 1, 16)), dim3(256), 0,
                       as_stream(stream), part, blocks, d_grad, d_loss);
    return check_launch();
}

int fls_mlp_adam(float* d_w, float* d_m, float* d_v, float* d_grad, uint16_t* d_wb, uint32_t step, float lr, float beta1,
                 float beta2, float eps, void* stream) {
    using namespace fls;
    FLS_RETURN_IF(!d_w || !d_m || !d_v || !d_grad || !d_wb, FLS_ERR_NULL);
    FLS_RETURN_IF(step == 0, FLS_ERR_ARG);
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(32), dim3(256), 0, as_stream(stream), d_w, d_m, d_v, d_grad, d_wb, lr,
                       beta1, beta2, eps, c1, c2);
    return check_launch();
}

int fls_mlp_mfma_selftest(const uint16_t* d_a, const uint16_t* d_b, float* d_d, void* stream) {
    if (!d_a || !d_b || !d_d) return FLS_ERR_NULL;
    hipLaunchKernelGGL(fls::mfma_selftest_kernel, dim3(1), dim3(64), 0, fls::as_stream(stream), d_a, d_b, d_d);
    return fls::check_launch();
}

int fls_mlp_mfma32_selftest(const uint16_t* d_a, const uint16_t* d_b, float* d_d, void* stream) {
    if (!d_a || !d_b || !d_d) return FLS_ERR_NULL;
    hipLaunchKernelGGL(fls::mfma32_selftest_kernel, dim3(1), dim3(64), 0, fls::as_stream(stream), d_a, d_b, d_d);
    return fls::check_launch();
}

}  // extern "C"
