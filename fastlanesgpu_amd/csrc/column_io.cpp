// Serialized compressed columns (fls_column_image): a fixed header followed
// by 64-byte-aligned sections, FNV-1a 64 checksum over the sections.  Host
// code; the image is plain bytes, so files, sockets and mmap all work.
#include <stdint.h>
#include <string.h>

#include "fls/fls.h"

namespace {

constexpr char kMagic[8] = {'F', 'L', 'S', 'C', 'O', 'L', '0', '1'};
constexpr uint32_t kVersion = 1;
constexpr uint64_t kMaxRows = uint64_t(1) << 56;
static_assert(sizeof(fls_column_image) == 128, "image header is 128 bytes");

uint64_t align64(uint64_t x) { return (x + 63u) & ~uint64_t(63); }

struct Counts {
    uint64_t vmeta, bases, aux;
};

// per-vector arrays an encoding carries, or false for an unknown encoding
bool counts_of(uint32_t encoding, uint64_t n_rows, Counts* c) {
    const uint64_t n_vec = (n_rows + 1023u) / 1024u;
    switch (encoding) {
        case FLS_ENC_BITPACK: *c = {0, 0, 0}; return true;
        case FLS_ENC_DELTA: *c = {n_vec, 32 * n_vec, 0}; return true;
        case FLS_ENC_RLE: *c = {n_vec, 32 * n_vec, n_vec}; return true;
        case FLS_ENC_MIXED: *c = {n_vec, 0, n_vec}; return true;
        default: return false;
    }
}

void layout(const Counts& c, uint64_t n_words, fls_column_image* h) {
    uint64_t o = sizeof(fls_column_image);
    h->off_words = o;
    o = align64(o + 4 * n_words);
    h->off_vmeta = o;
    o = align64(o + 8 * c.vmeta);
    h->off_bases = o;
    o = align64(o + 4 * c.bases);
    h->off_aux = o;
    o = align64(o + 8 * c.aux);
    h->total_bytes = o;
}

uint64_t fnv1a(const uint8_t* p, uint64_t n, uint64_t h) {
    for (uint64_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}

uint64_t checksum(const uint8_t* image, const fls_column_image& h) {
    uint64_t x = 1469598103934665603ull;
    x = fnv1a(image + h.off_words, 4 * h.n_words, x);
    x = fnv1a(image + h.off_vmeta, 8 * h.n_vmeta, x);
    x = fnv1a(image + h.off_bases, 4 * h.n_bases, x);
    x = fnv1a(image + h.off_aux, 8 * h.n_aux, x);
    return x;
}

// the words a column of this shape must hold at least (fixed width: all vectors)
bool words_ok(const fls_column& c, uint64_t n_rows) {
    if (c.encoding == FLS_ENC_BITPACK) return c.bw <= 32 && c.n_words >= (n_rows + 1023u) / 1024u * 32u * c.bw;
    return true;
}

// Every vector's metadata describes a payload inside the column's words, at
// widths the decoders take (0..32), with an RLE run count of 1..1024: the
// kernels read exactly the words this describes, so an image that passes
// never makes a decode or query kernel read outside the words array.
bool vectors_ok(const fls_column& c, uint64_t n_rows) {
    if (c.encoding == FLS_ENC_BITPACK) return true;
    const uint64_t n_vec = (n_rows + 1023u) / 1024u;
    for (uint64_t v = 0; v < n_vec; ++v) {
        const fls_vmeta m = c.vmeta[v];
        const uint64_t off = FLS_VMETA_OFFSET(m);
        uint64_t need = 0;
        if (c.encoding == FLS_ENC_DELTA) {
            const uint32_t bw = FLS_VMETA_BW(m);
            if (bw > 32) return false;
            need = 32ull * bw;
        } else if (c.encoding == FLS_ENC_RLE) {
            const uint32_t ibw = FLS_VMETA_BW(m);
            fls_rle_meta rm;
            memcpy(&rm, &c.aux[v], sizeof rm);
            if (ibw > 32 || rm.vbw > 32 || rm.n_runs == 0 || rm.n_runs > 1024) return false;
            need = 32ull * ibw + ((uint64_t)rm.n_runs * rm.vbw + 31u) / 32u;
        } else {  // MIXED: vmeta = offset << 8 | enc << 6 | width
            const uint32_t enc = (uint32_t)(m >> 6) & 3u, w = (uint32_t)m & 63u;
            if (w > 32) return false;
            const uint64_t ax = c.aux[v];
            switch (enc) {
                case FLS_MIX_BITPACK:
                case FLS_MIX_FOR: need = 32ull * w; break;
                case FLS_MIX_DELTA: need = 32ull + 32ull * w; break;
                default: {
                    const uint64_t nr = (ax >> 32) & 0xffffu, ibw = ax >> 48;
                    if (ibw > 1 || nr == 0 || nr > 1024) return false;
                    need = 32ull + 32ull * ibw + (nr * w + 31u) / 32u;
                }
            }
        }
        if (off > c.n_words || need > c.n_words - off) return false;
    }
    return true;
}

}  // namespace

extern "C" {

uint64_t fls_column_image_bytes(const fls_column* col, uint64_t n_rows) {
    Counts c;
    if (!col || !counts_of(col->encoding, n_rows, &c)) return 0;
    fls_column_image h{};
    layout(c, col->n_words, &h);
    return h.total_bytes;
}

int fls_column_serialize(const fls_column* col, uint64_t n_rows, void* out, uint64_t out_bytes) {
    if (!col || !out) return FLS_ERR_NULL;
    Counts c;
    if (n_rows > kMaxRows) return FLS_ERR_SIZE;
    if (!counts_of(col->encoding, n_rows, &c)) return FLS_ERR_ENCODING;
    if (!words_ok(*col, n_rows)) return FLS_ERR_SIZE;
    if ((col->n_words && !col->words) || (c.vmeta && !col->vmeta) || (c.bases && !col->bases) ||
        (c.aux && !col->aux))
        return FLS_ERR_NULL;
    if (!vectors_ok(*col, n_rows)) return FLS_ERR_FORMAT;
    fls_column_image h{};
    memcpy(h.magic, kMagic, 8);
    h.version = kVersion;
    h.encoding = col->encoding;
    h.bw = col->bw;
    h.base = col->base;
    h.transposed = col->transposed ? 1u : 0u;
    h.n_rows = n_rows;
    h.n_words = col->n_words;
    h.n_vmeta = c.vmeta;
    h.n_bases = c.bases;
    h.n_aux = c.aux;
    layout(c, col->n_words, &h);
    if (out_bytes < h.total_bytes) return FLS_ERR_CAPACITY;
    uint8_t* o = static_cast<uint8_t*>(out);
    memset(o, 0, h.total_bytes);  // padding between sections is zero
    if (h.n_words) memcpy(o + h.off_words, col->words, 4 * h.n_words);
    if (h.n_vmeta) memcpy(o + h.off_vmeta, col->vmeta, 8 * h.n_vmeta);
    if (h.n_bases) memcpy(o + h.off_bases, col->bases, 4 * h.n_bases);
    if (h.n_aux) memcpy(o + h.off_aux, col->aux, 8 * h.n_aux);
    h.checksum = checksum(o, h);
    memcpy(o, &h, sizeof h);
    return FLS_OK;
}

int fls_column_deserialize(const void* image, uint64_t image_bytes, fls_column* col, uint64_t* n_rows) {
    if (!image || !col || !n_rows) return FLS_ERR_NULL;
    if (image_bytes < sizeof(fls_column_image)) return FLS_ERR_FORMAT;
    fls_column_image h;
    memcpy(&h, image, sizeof h);
    if (memcmp(h.magic, kMagic, 8) != 0 || h.version != kVersion) return FLS_ERR_FORMAT;
    if (h.n_rows > kMaxRows) return FLS_ERR_FORMAT;  // keeps the vector-count arithmetic exact
    Counts c;
    if (!counts_of(h.encoding, h.n_rows, &c)) return FLS_ERR_FORMAT;
    if (c.vmeta != h.n_vmeta || c.bases != h.n_bases || c.aux != h.n_aux) return FLS_ERR_FORMAT;
    if (h.n_words > (image_bytes / 4)) return FLS_ERR_FORMAT;  // before the layout arithmetic
    fls_column_image expect = h;
    layout(c, h.n_words, &expect);
    if (expect.off_words != h.off_words || expect.off_vmeta != h.off_vmeta || expect.off_bases != h.off_bases ||
        expect.off_aux != h.off_aux || expect.total_bytes != h.total_bytes || h.total_bytes > image_bytes)
        return FLS_ERR_FORMAT;
    const uint8_t* p = static_cast<const uint8_t*>(image);
    if (checksum(p, h) != h.checksum) return FLS_ERR_FORMAT;
    fls_column out{};
    out.encoding = h.encoding;
    out.bw = h.bw;
    out.base = h.base;
    out.transposed = h.transposed;
    out.n_words = h.n_words;
    out.words = reinterpret_cast<const uint32_t*>(p + h.off_words);
    out.vmeta = h.n_vmeta ? reinterpret_cast<const fls_vmeta*>(p + h.off_vmeta) : nullptr;
    out.bases = h.n_bases ? reinterpret_cast<const uint32_t*>(p + h.off_bases) : nullptr;
    out.aux = h.n_aux ? reinterpret_cast<const uint64_t*>(p + h.off_aux) : nullptr;
    if (!words_ok(out, h.n_rows) || !vectors_ok(out, h.n_rows)) return FLS_ERR_FORMAT;
    *col = out;
    *n_rows = h.n_rows;
    return FLS_OK;
}

}  // extern "C"
