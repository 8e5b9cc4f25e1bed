"""GPU encoders (bit-pack, FastLanes delta, FastLanes RLE) against the host
encoders / the oracle, word for word, plus decode round trips."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(99)


def u32(t):
    return T.from_dev(t, np.uint32)


@pytest.mark.parametrize("transposed", [False, True])
def test_pack_column_device_matches_host(transposed):
    for bw in range(33):
        n = 5 * 1024 + 333
        x = RNG.integers(0, 2**bw, n, dtype=np.uint64).astype(np.uint32)
        want = fls.pack_column(x, bw, transpose=transposed)
        got = u32(fls.pack_column_device(T.to_dev(x), n, bw, transposed=transposed))[: want.size]
        np.testing.assert_array_equal(got, want, err_msg=f"bw={bw}")
    # frame of reference
    x = RNG.integers(1000, 1000 + 2**9, 3000).astype(np.uint32)
    want = fls.pack_column((x - 1000).astype(np.uint32), 9)
    np.testing.assert_array_equal(u32(fls.pack_column_device(T.to_dev(x), x.size, 9, base=1000))[: want.size], want)


@pytest.mark.parametrize("kind", ["sorted", "random", "dates", "constant", "small"])
def test_delta_device_matches_host(kind):
    n = 9 * 1024 + 100
    if kind == "sorted":
        x = np.sort(RNG.integers(0, 2**31, n)).astype(np.uint32)
    elif kind == "random":
        x = RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    elif kind == "dates":
        x = T.o_gen_c2(3, n)[0]
    elif kind == "constant":
        x = np.full(n, 7, np.uint32)
    else:
        n = 5
        x = np.array([3, 9, 9, 10, 20], np.uint32)
    words, vmeta, bases = fls.delta_encode_column(x)
    dw, dv, db = fls.delta_encode_device(T.to_dev(x), n)
    np.testing.assert_array_equal(T.from_dev(dv, np.uint64)[: vmeta.size], vmeta)
    np.testing.assert_array_equal(u32(db)[: bases.size], bases)
    np.testing.assert_array_equal(u32(dw)[: words.size], words)


def rle_input(n, mean_run, bw):
    starts = RNG.random(n) < 1.0 / mean_run
    starts[0] = True
    vals = RNG.integers(0, 2**bw, n, dtype=np.uint64).astype(np.uint32)
    return vals[np.flatnonzero(starts)][np.cumsum(starts) - 1].astype(np.uint32)


@pytest.mark.parametrize("case", [(4 * 1024, 8, 13), (3 * 1024 + 77, 8, 32), (2048, 1, 5), (1024, 5000, 20),
                                  (7, 2, 3), (5 * 1024, 32, 0)])
def test_rle_device_matches_oracle(case):
    n, mean_run, bw = case
    x = rle_input(n, mean_run, bw) if bw else np.zeros(n, np.uint32)
    words, vmeta, rmeta, bases = fls.rle_encode_device(T.to_dev(x), n)
    words, vmeta = u32(words), T.from_dev(vmeta, np.uint64)
    rmeta, bases = T.from_dev(rmeta, np.uint64), u32(bases)
    n_vec = (n + 1023) // 1024
    for v in range(n_vec):
        seg = x[v * 1024:(v + 1) * 1024]
        hdr = T.OrcRle()
        ovals = np.zeros(1024, np.uint32)
        oidx = np.zeros(1024, np.uint32)
        T.oracle().orc_rle_encode_u32(np.ascontiguousarray(seg), seg.size, hdr, ovals, oidx)
        rm = int(rmeta[v])
        vbase, n_runs, vbw, ibw = rm & 0xFFFFFFFF, (rm >> 32) & 0xFFFF, (rm >> 48) & 0xFF, (rm >> 56) & 0xFF
        assert (vbase, n_runs, vbw, ibw) == (hdr.vbase, hdr.n_runs, hdr.vbw, hdr.ibw)
        off = int(vmeta[v]) >> 8
        assert int(vmeta[v]) & 0xFF == ibw
        np.testing.assert_array_equal(bases[32 * v:32 * v + 32], np.array(hdr.bases, np.uint32))
        np.testing.assert_array_equal(words[off:off + 32 * ibw], oidx[: 32 * ibw])
        nvw = (n_runs * vbw + 31) // 32
        np.testing.assert_array_equal(words[off + 32 * ibw:off + 32 * ibw + nvw], ovals[:nvw])


@pytest.mark.parametrize("case", [(6 * 1024 + 5, 8, 11), (2048, 1, 32), (1024, 100000, 7)])
def test_rle_roundtrip_gpu(case):
    torch = T.torch_cuda()
    n, mean_run, bw = case
    x = rle_input(n, mean_run, bw)
    words, vmeta, rmeta, bases = fls.rle_encode_device(T.to_dev(x), n)
    n_vec = (n + 1023) // 1024
    out = torch.empty(n_vec * 1024, dtype=torch.int32, device="cuda")
    fls.rle_decode(words, vmeta, rmeta, bases, n_vec, out)
    padded = np.full(n_vec * 1024, x[-1], np.uint32)
    padded[:n] = x
    np.testing.assert_array_equal(u32(out), T.transpose_vectors(padded))
