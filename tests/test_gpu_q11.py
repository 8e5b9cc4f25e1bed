"""GPU parity of the fused decode + filter + aggregate (Q1.1 shape) vs the oracle."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
Q11 = (19930000, 19940000, 25, 1, 3)


def run(enc, n, params=Q11, pred=False):
    dev = [c.to_device() for c in enc]
    gt, lt, q, lo, hi = params
    return fls.query_q11(*dev, n, gt, lt, q, lo, hi, pred_load=pred)


def test_generators_match_oracle():
    torch = T.torch_cuda()
    n = 100_003
    got = [c.cpu().numpy().view(np.uint32) for c in fls.gen_c2(77, n)]
    for g, w in zip(got, T.o_gen_c2(77, n)):
        np.testing.assert_array_equal(g, w)
    np.testing.assert_array_equal(fls.gen_c1(3, 50_000, 11).cpu().numpy().view(np.uint32), T.o_gen_c1(3, 50_000, 11))
    del torch


@pytest.mark.parametrize("n", [1, 513, 1024, 4096 + 7, 300_000])
@pytest.mark.parametrize("pred", [False, True])
def test_q11_small(n, pred):
    cols = T.o_gen_c2(21, n)
    want = T.o_q11(cols, n)
    assert run(fls.encode_c2_table(*cols), n, pred=pred) == want


@pytest.mark.parametrize("params", [(0, 2**32 - 1, 40, 0, 10), (19921231, 19930102, 51, 0, 10),
                                    (19950000, 19940000, 25, 1, 3), (0, 2**32 - 1, 50, 5, 4)])
def test_q11_predicate_edges(params):
    n = 3 * 1024 + 77
    cols = T.o_gen_c2(8, n)
    assert run(fls.encode_c2_table(*cols), n, params) == T.o_q11(cols, n, params)


def test_q11_runtime_width_path():
    # widths other than (6, 4, 24) take the run-time dispatch path
    n = 50_000
    od, q, d, p = T.o_gen_c2(4, n)
    enc = [fls.delta_column(od), fls.bitpack_column(q, 7, transposed=True), fls.bitpack_column(d, 9, transposed=True),
           fls.bitpack_column(p, 31, transposed=True)]
    params = (19920000, 19990000, 30, 2, 7)
    for pred in (False, True):
        assert run(enc, n, params, pred) == T.o_q11([od, q, d, p], n, params)


def test_q11_for_date_normal_order():
    # orderdate as frame-of-reference bit-packing (fls_q11.cu:187-190) in normal row order
    n = 40 * 1024 + 5
    od, q, d, p = T.o_gen_c2(6, n)
    base = int(od.min())
    enc = [fls.bitpack_column(od, base=base), fls.bitpack_column(q), fls.bitpack_column(d), fls.bitpack_column(p)]
    assert enc[0].bw == int(od.max() - base).bit_length()
    for pred in (False, True):
        assert run(enc, n, Q11, pred) == T.o_q11([od, q, d, p], n)


def test_q11_accumulate_flag():
    torch = T.torch_cuda()
    n = 10_000
    cols = T.o_gen_c2(2, n)
    dev = [c.to_device() for c in fls.encode_c2_table(*cols)]
    out = torch.zeros(2, dtype=torch.int64, device="cuda")
    for _ in range(3):
        fls.query_q11(*dev, n, accumulate=True, d_out=out, sync=False)
    s, c = T.o_q11(cols, n)
    assert out.cpu().tolist() == [3 * s, 3 * c]


def test_q11_full_size_120m():
    n = 120_000_000
    dcols = fls.gen_c2(1, n)
    cols = [c.cpu().numpy().view(np.uint32) for c in dcols]
    del dcols
    want = T.o_q11(cols, n)
    assert want[1] > 2_000_000  # ~1.87 % selectivity
    enc = fls.encode_c2_table(*cols)
    dev = [c.to_device() for c in enc]
    assert fls.query_q11(*dev, n) == want
    assert fls.query_q11(*dev, n, pred_load=True) == want


@pytest.mark.parametrize("layout", ["rle_date", "all_delta", "rle_small_cols"])
def test_q11_rle_and_delta_columns(layout):
    """The generic path over GPU-encoded RLE and delta columns in every
    position (all share the transposed row order)."""
    n = 70 * 1024 + 333
    cols = T.o_gen_c2(31, n)

    def rle(x):
        w, vm, rm, bs = fls.rle_encode_device(T.to_dev(x), n)
        return fls.Column(fls.ENC_RLE, n, w, vmeta=vm, bases=bs, aux=rm)

    def delta(x):
        w, vm, bs = fls.delta_encode_device(T.to_dev(x), n)
        return fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)

    def packed(x):
        return fls.bitpack_column(x, transposed=True).to_device()

    kinds = {"rle_date": (rle, packed, packed, delta), "all_delta": (delta, delta, delta, delta),
             "rle_small_cols": (delta, rle, rle, packed)}[layout]
    enc = [f(c) for f, c in zip(kinds, cols)]
    want = T.o_q11(cols, n)
    for pred in (False, True):
        assert fls.query_q11(*enc, n, pred_load=pred) == want, (layout, pred)


def test_q11_reduction_slots_reused_and_concurrent():
    """150 calls in a row (overwriting and adding, both kernels, a one-block
    and a 123-block grid), then 8 streams with calls in flight at once, each
    into its own output."""
    torch = T.torch_cuda()
    sizes = [3 * 1024 + 77, 1_000_003]  # one block; 123 blocks (more than the 64 groups)
    data = []
    for i, n in enumerate(sizes):
        cols = T.o_gen_c2(40 + i, n)
        data.append((n, [c.to_device() for c in fls.encode_c2_table(*cols)], T.o_q11(cols, n)))
    acc = torch.zeros(2, dtype=torch.int64, device="cuda")
    outs = []
    for k in range(150):
        n, dev, _ = data[k % 2]
        outs.append(fls.query_q11(*dev, n, pred_load=bool(k % 3 == 0), sync=False))
        fls.query_q11(*dev, n, accumulate=True, d_out=acc, sync=False)
    for k, o in enumerate(outs):
        assert tuple(o.cpu().tolist()) == data[k % 2][2], k
    want = [75 * (data[0][2][j] + data[1][2][j]) for j in range(2)]
    assert acc.cpu().tolist() == want

    streams = [torch.cuda.Stream() for _ in range(8)]
    n, dev, w = data[1]
    torch.cuda.synchronize()
    res = []
    for rep in range(4):
        for s in streams:
            with torch.cuda.stream(s):
                out = torch.empty(2, dtype=torch.int64, device="cuda")
                res.append(fls.query_q11(*dev, n, d_out=out, stream=s.cuda_stream, sync=False))
    torch.cuda.synchronize()
    for o in res:
        assert tuple(o.cpu().tolist()) == w


def test_q11_zero_rows():
    torch = T.torch_cuda()
    n = 1024
    dev = [c.to_device() for c in fls.encode_c2_table(*T.o_gen_c2(3, n))]
    out = torch.full((2,), 7, dtype=torch.int64, device="cuda")
    assert fls.query_q11(*dev, 0, d_out=out) == (0, 0)
    out.fill_(7)
    fls.query_q11(*dev, 0, accumulate=True, d_out=out)
    assert out.cpu().tolist() == [7, 7]
