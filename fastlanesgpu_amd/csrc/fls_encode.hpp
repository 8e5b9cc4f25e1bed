// Internal helpers shared by the GPU encoders.  Internal to libfls.so.
#pragma once

#include "fls_runtime.hpp"

namespace fls {

// 32-bit words of scratch the plan/scan passes need for n_vec vectors.
uint64_t scratch_words(uint64_t n_vec);

// vmeta[i] = (exclusive prefix of words) << 8 | bws[i]; *d_total = sum(words).
int scan_offsets(const uint32_t* words, const uint32_t* bws, uint64_t n_vec, uint64_t* block_sums, fls_vmeta* vmeta,
                 uint64_t* d_total, hipStream_t s);

}  // namespace fls
