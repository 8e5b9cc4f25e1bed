"""The oracle's CPU baseline (decode + query over encoded columns) is pinned
against the bit-by-bit oracle functions before it is used as a baseline."""
import numpy as np

import fastlanesgpu_amd as fls
import fls_testlib as T

RNG = np.random.default_rng(5)


def test_lanewise_unpack_matches_bitwise():
    o = T.oracle()
    o.orc_unpack_lanewise_u32.argtypes = [T.U32P, T.U32P, T.ctypes.c_uint]
    for bw in range(33):
        x = RNG.integers(0, 2**bw, 1024, dtype=np.uint64).astype(np.uint32)
        packed = np.zeros(1024, np.uint32)
        packed[: 32 * bw] = T.o_pack(x, bw)
        out = np.zeros(1024, np.uint32)
        o.orc_unpack_lanewise_u32(packed, out, bw)
        np.testing.assert_array_equal(out, x)


def test_q11_encoded_matches_raw_evaluation():
    for n in (1, 700, 1024, 20 * 1024 + 333):
        cols = T.o_gen_c2(13, n)
        enc = fls.encode_c2_table(*cols)
        assert T.o_q11_encoded(enc, n, threads=4) == T.o_q11(cols, n)
    # a selective window, so the sum is non-trivial at small n
    n = 5000
    cols = T.o_gen_c2(13, n)
    enc = fls.encode_c2_table(*cols)
    params = (0, 2**31, 40, 0, 10)
    got = T.o_q11_encoded(enc, n, threads=2, params=params)
    assert got == T.o_q11(cols, n, params) and got[1] > 0
