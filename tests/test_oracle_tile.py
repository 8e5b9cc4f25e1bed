"""Pin the oracle's restatement of the tile-based competitor formats
(GPU-FOR, GPU-DFOR, GPU-RFOR; reference tile_based/): GPU-RFOR word for word
against the reference's own rleBinPack built into oracle/_ref, a GPU-FOR block
laid out by hand, and round trips at every width for the three decoders."""
import ctypes

import numpy as np
import pytest

import fls_testlib as T

RNG = np.random.default_rng(31)
TILE_SO = T.ROOT / "oracle" / "_ref" / "libtileref.so"


def widths_data(n, bw, seed):
    r = np.random.default_rng(seed)
    hi = 1 << bw if bw < 32 else 1 << 32
    return (r.integers(0, hi, n, dtype=np.uint64) + (0 if bw == 32 else 1000)).astype(np.uint32)


def runs_data(n, seed, mean=8):
    r = np.random.default_rng(seed)
    lens = r.geometric(1.0 / mean, n)
    vals = r.integers(0, 1 << 17, n)
    return np.repeat(vals, lens)[:n].astype(np.uint32)


def test_for_block_by_hand():
    # 128 values 100 + (i % 8): min 100, width 3 everywhere, four miniblocks of
    # three words each, value i at bit 3 * (i % 32) of miniblock i // 32
    x = (100 + np.arange(128) % 8).astype(np.uint32)
    words, starts = T.o_tile_encode("for", x)
    assert list(words[:4]) == [128, 4, 128, 100]
    assert list(starts) == [4, 4 + 2 + 12]
    assert words[4] == 100 and words[5] == 0x03030303
    want = np.zeros(12, np.uint64)
    for i in range(128):
        m, k = divmod(i, 32)
        bit = 3 * k
        want[3 * m + bit // 32] |= np.uint64((i % 8) << (bit % 32)) & np.uint64(0xFFFFFFFF)
        if bit % 32 > 29:  # straddles into the next word
            want[3 * m + bit // 32 + 1] |= np.uint64((i % 8) >> (32 - bit % 32))
    np.testing.assert_array_equal(words[6:18], want.astype(np.uint32))
    np.testing.assert_array_equal(T.o_tile_decode("for", words, starts, 128), x)


def test_for_width_zero_takes_one_word_per_miniblock():
    # the reference loop steps one word after every miniblock, also at width 0
    x = np.full(256, 7, np.uint32)
    words, starts = T.o_tile_encode("for", x)
    assert list(starts) == [4, 10, 16] and words[5] == 0
    np.testing.assert_array_equal(T.o_tile_decode("for", words, starts, 256), x)


@pytest.mark.parametrize("bw", list(range(0, 33)))
def test_for_round_trip_every_width(bw):
    for n in (1, 127, 128, 129, 1000):
        x = widths_data(n, bw, bw * 7 + n)
        words, starts = T.o_tile_encode("for", x)
        np.testing.assert_array_equal(T.o_tile_decode("for", words, starts, n), x)


@pytest.mark.parametrize("kind", ["sorted", "ramp7", "random", "descending", "runs"])
def test_dfor_round_trip(kind):
    n = 5 * 512 + 77
    x = {"sorted": np.sort(RNG.integers(0, 10**9, n)), "ramp7": 7 * np.arange(n),
         "random": RNG.integers(0, 2**32, n, dtype=np.uint64), "descending": 10**9 - 3 * np.arange(n),
         "runs": runs_data(n, 3)}[kind].astype(np.uint32)
    words, starts = T.o_tile_encode("dfor", x)
    assert words[starts[0] - 1] == x[0]  # each tile leads with its first value
    np.testing.assert_array_equal(T.o_tile_decode("dfor", words, starts, n), x)


def test_dfor_ramp_is_width_three():
    # the reference delta bench's column i * 7: every delta is 7 (width 3),
    # the first delta of a tile is 0
    x = (7 * np.arange(1024)).astype(np.uint32)
    words, starts = T.o_tile_encode("dfor", x)
    for b in range(8):
        assert words[starts[b]] == 0 and words[starts[b] + 1] == 0x03030303
    assert starts[4] - starts[3] == 2 + 12 + 1  # block 3 of tile 0, then tile 1's first value


@pytest.mark.parametrize("kind", ["runs", "random", "constant", "two"])
def test_rfor_round_trip(kind):
    n = 7 * 512 + 300
    x = {"runs": runs_data(n, 5), "random": RNG.integers(0, 2**31, n), "constant": np.full(n, 9),
         "two": np.arange(n) // 256}[kind].astype(np.uint32)
    vals, vs, rls, rs = T.o_rfor_encode(x)
    np.testing.assert_array_equal(T.o_rfor_decode(vals, vs, rls, rs, n), x)


def ref_rle():
    if not TILE_SO.exists():
        pytest.skip("oracle/_ref/libtileref.so not built (reference checkout absent)")
    lib = ctypes.CDLL(str(TILE_SO))
    u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C")
    lib.ref_rle_bin_pack.argtypes = [u32p, ctypes.c_uint32, u32p, u32p, u32p, u32p, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint64)]
    return lib


@pytest.mark.parametrize("kind", ["runs", "runs_long", "random17", "constant"])
def test_rfor_matches_reference_encoder_word_for_word(kind):
    lib = ref_rle()
    n = 16 * 512  # the reference encodes whole 512-value tiles
    x = {"runs": runs_data(n, 11), "runs_long": runs_data(n, 12, mean=200),
         "random17": RNG.integers(0, 1 << 17, n), "constant": np.full(n, 123456)}[kind].astype(np.uint32)
    cap = 4 * n + 64
    v, r = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    vo, ro = np.zeros(n // 512 + 1, np.uint32), np.zeros(n // 512 + 1, np.uint32)
    nv, nr = ctypes.c_uint64(0), ctypes.c_uint64(0)
    lib.ref_rle_bin_pack(x.copy(), n, v, r, vo, ro, ctypes.byref(nv), ctypes.byref(nr))
    vals, vs, rls, rs = T.o_rfor_encode(x)
    assert (nv.value, nr.value) == (vals.size, rls.size)
    np.testing.assert_array_equal(vs, vo)
    np.testing.assert_array_equal(rs, ro)
    np.testing.assert_array_equal(vals, v[:nv.value])
    np.testing.assert_array_equal(rls, r[:nr.value])
