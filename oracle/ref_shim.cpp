// C-ABI shim over the reference's own host encoders, compiled together with
// the reference sources (fastlanes/src/pack.cpp, transpose.cpp, unrsum.cpp)
// into oracle/_ref/libflsref.so.  TEST INFRASTRUCTURE ONLY: the tests use it
// to validate the C oracle and to produce golden fixtures.
#include "fls_gen/pack/pack.hpp"
#include "fls_gen/transpose/transpose.hpp"
#include "fls_gen/unrsum/unrsum.hpp"

extern "C" {
void ref_pack_u32(const uint32_t* in, uint32_t* out, unsigned bw) {
	generated::pack::fallback::scalar::pack(in, out, static_cast<uint8_t>(bw));
}
void ref_transpose_i_u32(const uint32_t* in, uint32_t* out) {
	generated::transpose::fallback::scalar::transpose_i(in, out);
}
void ref_transpose_o_u32(const uint32_t* in, uint32_t* out) {
	generated::transpose::fallback::scalar::transpose_o(in, out);
}
void ref_unrsum_u32(const uint32_t* in, uint32_t* out) {
	generated::unrsum::fallback::scalar::unrsum(in, out);
}
}
