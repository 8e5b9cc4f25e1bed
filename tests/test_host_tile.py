"""The library's host encoders of the tile-based formats write exactly the
oracle's words (which the oracle tests pin to the reference encoders)."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

RNG = np.random.default_rng(5)


def inputs():
    yield "uniform11", RNG.integers(0, 1 << 11, 9000).astype(np.uint32)
    yield "uniform32", RNG.integers(0, 2**32, 3000, dtype=np.uint64).astype(np.uint32)
    yield "sorted", np.sort(RNG.integers(0, 10**9, 7777)).astype(np.uint32)
    yield "ramp7", (7 * np.arange(4096)).astype(np.uint32)
    yield "descending", (10**9 - 5 * np.arange(2000)).astype(np.uint32)
    yield "runs", np.repeat(RNG.integers(0, 1 << 16, 2000), RNG.geometric(1 / 9, 2000))[:10_000].astype(np.uint32)
    yield "constant", np.full(1500, 42, np.uint32)
    yield "one", np.array([5], np.uint32)
    for bw in (0, 1, 5, 17, 31, 32):
        hi = 1 << 32 if bw == 32 else 1 << bw
        yield f"w{bw}", RNG.integers(0, hi, 640, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("kind", ["for", "dfor"])
def test_for_dfor_words_equal_oracle(kind):
    for name, x in inputs():
        col = fls.tile_encode(x, kind)
        words, starts = T.o_tile_encode(kind, x)
        np.testing.assert_array_equal(col.words, words, err_msg=name)
        np.testing.assert_array_equal(col.starts, starts, err_msg=name)


def test_rfor_words_equal_oracle():
    for name, x in inputs():
        col = fls.tile_encode(x, "rfor")
        vals, vs, rls, rs = T.o_rfor_encode(x)
        for a, b in ((col.words, vals), (col.starts, vs), (col.rl_words, rls), (col.rl_starts, rs)):
            np.testing.assert_array_equal(a, b, err_msg=name)


def test_empty():
    for kind in ("for", "dfor", "rfor"):
        col = fls.tile_encode(np.zeros(0, np.uint32), kind)
        assert list(col.words[:4]) == [512 if kind == "rfor" else 128, 4, 0, 0]
