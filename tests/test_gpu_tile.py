"""GPU parity of the tile-based competitor decoders (GPU-FOR, GPU-DFOR,
GPU-RFOR): decoded values equal the encoded column (and the oracle's decode),
sums equal numpy's, with and without the values written."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(71)


def cases():
    yield "uniform11", RNG.integers(0, 1 << 11, 100_000 + 33).astype(np.uint32)
    yield "uniform32", RNG.integers(0, 2**32, 20_000, dtype=np.uint64).astype(np.uint32)
    yield "sorted", np.sort(RNG.integers(0, 10**9, 77_777)).astype(np.uint32)
    yield "descending", (10**9 - 5 * np.arange(5_000)).astype(np.uint32)
    yield "runs", np.repeat(RNG.integers(0, 1 << 16, 20_000), RNG.geometric(1 / 9, 20_000))[:100_000].astype(np.uint32)
    yield "long_runs", np.repeat(RNG.integers(0, 9, 300), 700)[:200_000].astype(np.uint32)
    yield "constant", np.full(3000, 42, np.uint32)
    yield "one", np.array([5], np.uint32)
    yield "tile", np.arange(512, dtype=np.uint32)
    for bw in (0, 1, 7, 16, 25, 31, 32):
        hi = 1 << 32 if bw == 32 else 1 << bw
        yield f"w{bw}", RNG.integers(0, hi, 4096 + 100, dtype=np.uint64).astype(np.uint32)


@pytest.mark.parametrize("kind", ["for", "dfor", "rfor"])
def test_decode_equals_column_and_oracle(kind):
    for name, x in cases():
        col = fls.tile_encode(x, kind)
        dev = col.to_device()
        vals, s = fls.tile_decode(dev)
        got = T.from_dev(vals)
        np.testing.assert_array_equal(got, x, err_msg=f"{kind} {name}")
        if kind == "rfor":
            want = T.o_rfor_decode(col.words, col.starts, col.rl_words, col.rl_starts, x.size)
        else:
            want = T.o_tile_decode(kind, col.words, col.starts, x.size)
        np.testing.assert_array_equal(got, want, err_msg=f"{kind} {name} oracle")
        total = int(x.astype(np.uint64).sum())
        assert int(T.from_dev(s, np.uint64)[0]) == total, (kind, name)
        _, s2 = fls.tile_decode(dev, values=False)
        assert int(T.from_dev(s2, np.uint64)[0]) == total, (kind, name, "sum only")


def test_large_for_column():
    torch = T.torch_cuda()
    n = (1 << 26) + 123
    x = fls.gen_c1(9, n, 11)
    col = fls.tile_encode(x.cpu().numpy().view(np.uint32), "for").to_device()
    vals, s = fls.tile_decode(col)
    assert bool(torch.equal(vals, x))
    assert int(T.from_dev(s, np.uint64)[0]) == int(x.to(torch.int64).sum().item())


def test_errors():
    col = fls.tile_encode(np.arange(1000, dtype=np.uint32), "for").to_device()
    with pytest.raises(fls.FlsError):
        fls.call("fls_tile_for_decode_u32", fls.dptr(col.words), fls.dptr(col.starts), col.n, None, None, None)
