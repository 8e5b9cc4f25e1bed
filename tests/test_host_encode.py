"""Product host encoders (libfls.so, CPU side) against the pinned oracle."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

RNG = np.random.default_rng(11)


@pytest.mark.parametrize("variant", ["normal", "fused"])
def test_pack_golden(variant):
    g = T.golden()
    for bw in range(33):
        np.testing.assert_array_equal(fls.pack(g[f"rand_{variant}"][bw], bw), g[f"packed_{variant}"][bw][: 32 * bw])


def test_pack_random_masks_like_oracle():
    for bw in range(33):
        x = RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)
        np.testing.assert_array_equal(fls.pack(x, bw), T.o_pack(x, bw), err_msg=f"bw={bw}")


def test_transpose_unrsum():
    x = RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)
    np.testing.assert_array_equal(fls.transpose_i(x), T.o_perm("orc_transpose_i_u32", x))
    np.testing.assert_array_equal(fls.transpose_o(x), T.o_perm("orc_transpose_o_u32", x))
    np.testing.assert_array_equal(fls.unrsum(x), T.o_perm("orc_unrsum_u32", x))
    g = T.golden()
    for k in range(g["unrsum_in"].shape[0]):
        np.testing.assert_array_equal(fls.unrsum(g["unrsum_in"][k]), g["unrsum_out"][k])


@pytest.mark.parametrize("n", [1, 1000, 1024, 5000, 3 * 1024])
@pytest.mark.parametrize("transpose", [False, True])
def test_pack_column(n, transpose):
    bw = 13
    x = RNG.integers(0, 2**bw, n).astype(np.uint32)
    words = fls.pack_column(x, bw, transpose=transpose)
    n_vec = (n + 1023) // 1024
    assert words.size == n_vec * 32 * bw
    padded = np.zeros(n_vec * 1024, np.uint32)
    padded[:n] = x
    for v in range(n_vec):
        vec = padded[v * 1024:(v + 1) * 1024]
        if transpose:
            vec = T.o_perm("orc_transpose_i_u32", vec)
        np.testing.assert_array_equal(words[v * 32 * bw:(v + 1) * 32 * bw], T.o_pack(vec, bw))


@pytest.mark.parametrize("n", [1, 1024, 2500, 8 * 1024])
def test_delta_encode_column(n):
    x = np.sort(RNG.integers(0, 2**30, n)).astype(np.uint32)
    words, vmeta, bases = fls.delta_encode_column(x)
    n_vec = (n + 1023) // 1024
    padded = np.full(n_vec * 1024, x[-1], np.uint32)
    padded[:n] = x
    off = 0
    for v in range(n_vec):
        packed, b, bw = T.o_delta_encode(padded[v * 1024:(v + 1) * 1024])
        assert int(vmeta[v]) & 0xFF == bw and int(vmeta[v]) >> 8 == off
        np.testing.assert_array_equal(bases[32 * v:32 * v + 32], b)
        np.testing.assert_array_equal(words[off:off + 32 * bw], packed)
        off += 32 * bw
    assert words.size == off


def test_status_codes():
    lib = fls.load()
    assert lib.fls_pack_u32(None, None, 33) == -1
    assert lib.fls_unpack_u32(None, None, 1, 40, None) == -1
    assert lib.fls_unpack_u32(None, None, 1, 5, None) == -3
    assert lib.fls_status_string(-1) == b"bit width outside 0..32"
    with pytest.raises(fls.FlsError):
        fls.pack(np.zeros(1024, np.uint32), 33)
