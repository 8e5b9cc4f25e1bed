/*
 * CPU oracle, learned-selector part.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The reference has no encoding selector; its nearest analogue is the
 * per-column width search find_bw (fastlanes/src/include/ssb_utils.h:222-239).
 * Everything here defines this project's per-vector mixed encoding and its
 * selector, built from the pinned pieces of oracle.c:
 *   orc_mix_encode_u32   a vector encoded with one of the four encodings by
 *                        literally running the pinned encoders (pack,
 *                        transpose_i, delta = transpose_i + unrsum + pack,
 *                        RLE); its word count IS the size.
 *   orc_mix_sizes / orc_exhaustive_class
 *                        brute force: encode four times, keep the smallest.
 *   orc_vector_features  the 16 selector features, plain loops.
 *   orc_gen_train_vector the training workload definition.
 *   orc_mlp_*            the MLP in double; with round_bf16 set, operands are
 *                        rounded to BF16 where the GPU kernels round them.
 *                        Pinned by finite differences of the exact (unrounded)
 *                        loss.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define ST_TRAIN_FAMILY 300001ULL
#define ST_TRAIN_VALUES 300002ULL
#define ST_TRAIN_NOISE 300003ULL
#define ST_TRAIN_RUNS 300004ULL

unsigned orc_width_bucket(unsigned w) {
    static const unsigned top[10] = {0, 2, 4, 6, 8, 12, 16, 20, 24, 32};
    for (unsigned b = 0; b < 10; ++b)
        if (w <= top[b]) return b;
    return 9;
}

unsigned orc_mix_encode_u32(const uint32_t* in, unsigned enc, uint32_t* words, unsigned* width, uint64_t* aux) {
    uint32_t t[1024];
    *aux = 0;
    if (enc == 0 || enc == 1) {
        uint32_t mn = 0xffffffffu, mx = 0;
        for (unsigned i = 0; i < 1024; ++i) {
            if (in[i] < mn) mn = in[i];
            if (in[i] > mx) mx = in[i];
        }
        const uint32_t base = enc == 1 ? mn : 0u;
        uint32_t shifted[1024];
        for (unsigned i = 0; i < 1024; ++i) shifted[i] = in[i] - base;
        orc_transpose_i_u32(shifted, t);
        const unsigned w = orc_bit_width(mx - base);
        memset(words, 0, sizeof(uint32_t) * 32 * w);
        orc_pack_u32(t, words, w);
        *width = w;
        if (enc == 1) *aux = mn;
        return 32 * w;
    }
    if (enc == 2) {
        const unsigned w = orc_delta_encode_u32(in, words + 32, words);
        *width = w;
        return 32 + 32 * w;
    }
    orc_rle_vec h;
    uint32_t vals[1024], idx[1024];
    orc_rle_encode_u32(in, 1024, &h, vals, idx);
    memcpy(words, h.bases, sizeof(uint32_t) * 32);
    memcpy(words + 32, idx, sizeof(uint32_t) * 32 * h.ibw);
    const unsigned nv = (h.n_runs * h.vbw + 31) / 32;
    memcpy(words + 32 + 32 * h.ibw, vals, sizeof(uint32_t) * nv);
    *width = h.vbw;
    *aux = (uint64_t)h.vbase | ((uint64_t)h.n_runs << 32) | ((uint64_t)h.ibw << 48);
    return 32 + 32 * h.ibw + nv;
}

void orc_mix_sizes(const uint32_t* in, uint32_t* sizes4, unsigned* widths4) {
    uint32_t words[1088];
    for (unsigned e = 0; e < 4; ++e) {
        unsigned w;
        uint64_t aux;
        sizes4[e] = 4 * orc_mix_encode_u32(in, e, words, &w, &aux) + 16;
        if (widths4) widths4[e] = w;
    }
}

unsigned orc_exhaustive_class(const uint32_t* in) {
    uint32_t s[4];
    unsigned w[4], best = 0;
    orc_mix_sizes(in, s, w);
    for (unsigned e = 1; e < 4; ++e)
        if (s[e] < s[best]) best = e;
    return 10 * best + orc_width_bucket(w[best]);
}

/* ---------------------------------------------------------------- features
 * Sample: 32 groups of 4 consecutive rows, group g = rows 32*L(g/8) + 4*(g%8)
 * + 0..3 with L(k) = k*31/3 = 0, 10, 20, 31: four whole 32-row lines, the
 * vector's first and last among them.                                       */
static unsigned sample_row(unsigned g, unsigned j) { return 32 * ((g / 8) * 31 / 3) + 4 * (g % 8) + j; }
void orc_vector_features(const uint32_t* in, float* f) {
    uint32_t orv = 0, mn = 0xffffffffu, mx = 0, dor = 0, zero = 0, nondec = 0, absmax = 0, dsum = 0, big = 0;
    uint32_t eq = 0, fmin = 0xffffffffu, fmax = 0, wsum = 0, at_min = 0;
    for (unsigned blk = 0; blk < 32; ++blk) {
        const uint32_t* s = in + sample_row(blk, 0);
        for (unsigned j = 0; j < 4; ++j) {
            orv |= s[j];
            if (s[j] < mn) mn = s[j];
            if (s[j] > mx) mx = s[j];
            wsum += orc_bit_width(s[j]);
            if (j) {
                const uint32_t d = s[j] - s[j - 1];
                const uint32_t a = (int32_t)d < 0 ? 0u - d : d;
                dor |= d;
                zero += d == 0;
                nondec += s[j] >= s[j - 1];
                if (a > absmax) absmax = a;
                dsum += a < 0xffffu ? a : 0xffffu;
                big += a >= 256u;
            }
        }
        eq += s[0] == s[1] && s[1] == s[2] && s[2] == s[3];
        if (s[0] < fmin) fmin = s[0];
        if (s[0] > fmax) fmax = s[0];
    }
    for (unsigned blk = 0; blk < 32; ++blk)
        for (unsigned j = 0; j < 4; ++j) at_min += in[sample_row(blk, j)] == mn;
    f[0] = orc_bit_width(orv) / 32.f;
    f[1] = orc_bit_width(mx - mn) / 32.f;
    f[2] = orc_bit_width(dor) / 32.f;
    f[3] = zero / 96.f;
    f[4] = nondec / 96.f;
    f[5] = orc_bit_width(absmax) / 32.f;
    f[6] = at_min / 128.f;
    f[7] = orc_bit_width(dsum / 96u) / 32.f;
    f[8] = orc_bit_width(fmax - fmin) / 32.f;
    f[9] = eq / 32.f;
    f[10] = orc_bit_width(mn) / 32.f;
    f[11] = big / 96.f;
    f[12] = wsum / (128.f * 32.f);
    f[13] = orc_bit_width(mx) / 32.f;
    f[14] = (float)(orc_bit_width(mx - mn) > orc_bit_width(dor));
    f[15] = 1.0f;
}

/* ---------------------------------------------------------------- training data
 * Vector v of the training workload (families: 0 uniform, 1 Zipf s=1.1 over
 * ranks 1..4096 through the inverse CDF zipf_cdf, 2 sorted + Gaussian noise
 * sigma 16 (Irwin-Hall sum of 12 uniforms), 3 runs of geometric length with
 * mean 32).                                                                  */
void orc_gen_train_vector(uint64_t seed, uint64_t v, const uint32_t* zipf_cdf, uint32_t* out) {
    const uint32_t fam = orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v, 4);
    for (uint64_t i = 0; i < 1024; ++i) {
        const uint64_t g = v * 1024 + i;
        uint32_t x;
        if (fam == 0) {
            const uint32_t bw = 1 + orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 1, 32);
            const uint32_t base =
                orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 2, 2) ? orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 3, 1ULL << 31) : 0;
            x = base + orc_uniform(seed, ST_TRAIN_VALUES, g, 1ULL << bw);
        } else if (fam == 1) {
            const uint32_t u = orc_uniform(seed, ST_TRAIN_VALUES, g, 1ULL << 32);
            uint32_t r = 0;
            while (r < 4095 && zipf_cdf[r] <= u) ++r;
            const uint32_t stride = 1u << orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 1, 12);
            x = orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 2, 1u << 20) + (r + 1) * stride;
        } else if (fam == 2) {
            const uint32_t step = orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 1, 256);
            const int64_t start = orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 2, 1u << 24);
            int64_t sum = 0;
            for (int k = 0; k < 12; ++k) sum += orc_uniform(seed, ST_TRAIN_NOISE, 12 * g + k, 1u << 16);
            const int64_t noise = (16 * (sum - 6 * 65536)) / 65536;
            const int64_t val = start + (int64_t)i * step + noise;
            x = (uint32_t)(val < 0 ? 0 : val);
        } else {
            const uint32_t bw = 1 + orc_uniform(seed, ST_TRAIN_FAMILY, 8 * v + 1, 32);
            uint64_t j = i;
            while (j > 0 && orc_uniform(seed, ST_TRAIN_RUNS, v * 1024 + j, 32) != 0) --j;
            x = orc_uniform(seed, ST_TRAIN_VALUES, v * 1024 + j, 1ULL << bw);
        }
        out[i] = x;
    }
}

/* ---------------------------------------------------------------- MLP
 * 16 -> 64 -> 64 -> 40, ReLU, softmax cross-entropy; parameters W1[16][64],
 * b1[64], W2[64][64], b2[64], W3[64][40], b3[40] (W[in][out]).              */
#define NI 16
#define NH 64
#define NO 40
enum { OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH * NO,
       NPARAM = OB3 + NO };

/* round-to-nearest-even to BF16, returned as a double */
double orc_bf16(double x) {
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

static double rnd(double x, int on) { return on ? orc_bf16(x) : x; }

/* one sample: activations and logits */
static void fwd1(const double* w, const float* feat, int rb, double* x, double* h1, double* h2, double* z) {
    for (int i = 0; i < NI; ++i) x[i] = rnd(feat[i], rb);
    for (int j = 0; j < NH; ++j) {
        double a = 0;
        for (int i = 0; i < NI; ++i) a += x[i] * rnd(w[OW1 + i * NH + j], rb);
        a += rnd(w[OB1 + j], rb);
        h1[j] = rnd(a > 0 ? a : 0, rb);
    }
    for (int j = 0; j < NH; ++j) {
        double a = 0;
        for (int i = 0; i < NH; ++i) a += h1[i] * rnd(w[OW2 + i * NH + j], rb);
        a += rnd(w[OB2 + j], rb);
        h2[j] = rnd(a > 0 ? a : 0, rb);
    }
    for (int c = 0; c < NO; ++c) {
        double a = 0;
        for (int i = 0; i < NH; ++i) a += h2[i] * rnd(w[OW3 + i * NO + c], rb);
        z[c] = a + rnd(w[OB3 + c], rb);
    }
}

unsigned orc_mlp_param_count(void) { return NPARAM; }

void orc_mlp_forward(const double* w, const float* feat, uint64_t n, int round_bf16, double* logits) {
    double x[NI], h1[NH], h2[NH];
    for (uint64_t s = 0; s < n; ++s) fwd1(w, feat + s * NI, round_bf16, x, h1, h2, logits + s * NO);
}

double orc_mlp_grad(const double* w, const float* feat, const uint8_t* label, uint64_t n, double inv_batch,
                    int round_bf16, double* grad) {
    double x[NI], h1[NH], h2[NH], z[NO], dz[NO], dh2[NH], dh1[NH], loss = 0;
    const int rb = round_bf16;
    memset(grad, 0, sizeof(double) * NPARAM);
    for (uint64_t s = 0; s < n; ++s) {
        fwd1(w, feat + s * NI, rb, x, h1, h2, z);
        double mx = z[0], sum = 0;
        for (int c = 1; c < NO; ++c)
            if (z[c] > mx) mx = z[c];
        for (int c = 0; c < NO; ++c) sum += exp(z[c] - mx);
        const int y = label[s];
        loss += log(sum) + mx - z[y];
        for (int c = 0; c < NO; ++c) dz[c] = rnd((exp(z[c] - mx) / sum - (c == y)) * inv_batch, rb);
        for (int i = 0; i < NH; ++i)
            for (int c = 0; c < NO; ++c) grad[OW3 + i * NO + c] += h2[i] * dz[c];
        for (int c = 0; c < NO; ++c) grad[OB3 + c] += dz[c];
        for (int i = 0; i < NH; ++i) {
            double a = 0;
            for (int c = 0; c < NO; ++c) a += dz[c] * rnd(w[OW3 + i * NO + c], rb);
            dh2[i] = rnd(h2[i] > 0 ? a : 0, rb);
        }
        for (int i = 0; i < NH; ++i)
            for (int j = 0; j < NH; ++j) grad[OW2 + i * NH + j] += h1[i] * dh2[j];
        for (int j = 0; j < NH; ++j) grad[OB2 + j] += dh2[j];
        for (int i = 0; i < NH; ++i) {
            double a = 0;
            for (int j = 0; j < NH; ++j) a += dh2[j] * rnd(w[OW2 + i * NH + j], rb);
            dh1[i] = rnd(h1[i] > 0 ? a : 0, rb);
        }
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NH; ++j) grad[OW1 + i * NH + j] += x[i] * dh1[j];
        for (int j = 0; j < NH; ++j) grad[OB1 + j] += dh1[j];
    }
    return loss;
}
