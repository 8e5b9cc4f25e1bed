"""GPU parity of the fused decode + hash-probe + group-by (Q2.1 shape) vs the oracle."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
SMALL = dict(n_date=2556, n_supp=400, n_part=8000, n_cust=10)


def dev_dims(dims_np, sizes):
    torch = T.torch_cuda()
    d = {k: torch.from_numpy(v.view(np.int32).copy()).cuda() for k, v in dims_np.items()}
    d.update({k: v for k, v in sizes.items()})
    return d


def run_both(fact, dims_np, sizes, n, enc=None):
    d = dev_dims(dims_np, sizes)
    enc = enc or fls.encode_c3_table(*fact, n_date=sizes["n_date"], n_supp=sizes["n_supp"], n_part=sizes["n_part"])
    cols = [c.to_device() for c in enc]
    ws = fls.q21_build(d)
    out_c = fls.query_q21(*cols, n, d, ws).cpu().numpy()
    # reference tables probed per row, and through their occupancy bitmaps
    pht, keep = fls.q21_build_pht(d, occupancy=False)
    out_p = fls.query_q21_pht(*cols, n, pht).cpu().numpy()
    pht_b, keep_b = fls.q21_build_pht(d)
    out_b = fls.query_q21_pht(*cols, n, pht_b).cpu().numpy()
    np.testing.assert_array_equal(out_b, out_p)
    del keep, keep_b
    return out_c, out_p


def check(out, want):
    sums, counts = want
    np.testing.assert_array_equal(out[7000:].view(np.uint64), counts)
    np.testing.assert_array_equal(out[:7000], sums)


@pytest.mark.parametrize("pair", [False, True])
@pytest.mark.parametrize("n", [1, 63, 64, 1000 + 7, 800_000])
def test_pht_occupancy_bits(pair, n):
    torch = T.torch_cuda()
    rng = np.random.default_rng(n + pair)
    words = 2 * n if pair else n
    ht = rng.integers(1, 2**32, words, dtype=np.uint64).astype(np.uint32)
    ht[rng.random(words) < 0.7] = 0
    if pair:  # a slot is empty only when both words are
        occ = (ht[0::2] != 0) | (ht[1::2] != 0)
    else:
        occ = ht != 0
    bits = fls.pht_occupancy(torch.from_numpy(ht.view(np.int32)).cuda(), n, pair)
    got = np.unpackbits(T.from_dev(bits, np.uint32).view(np.uint8), bitorder="little")[:n].astype(bool)
    np.testing.assert_array_equal(got, occ)


def test_generators_match_oracle():
    n = 70_001
    fact_d = fls.gen_c3_fact(5, n, with_customer=True)
    dims_d = fls.gen_c3_dims(5, **{k: v // 10 for k, v in fls.C3_DIMS.items()})
    fact_o, dims_o = T.o_gen_c3(5, n, **{k: v // 10 for k, v in fls.C3_DIMS.items()})
    for k, want in zip(["datefk", "suppfk", "partfk", "revenue"], fact_o):
        np.testing.assert_array_equal(fact_d[k].cpu().numpy().view(np.uint32), want)
    for k, want in dims_o.items():
        np.testing.assert_array_equal(dims_d[k].cpu().numpy().view(np.uint32)[: want.size], want)


@pytest.mark.parametrize("n", [1, 700, 1024, 50_000 + 17])
def test_q21_small(n):
    fact, dims = T.o_gen_c3(11, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    want = T.o_q21(fact, dims, n)
    out_c, out_p = run_both(fact, dims, SMALL, n)
    check(out_c, want)
    check(out_p, want)


@pytest.mark.parametrize("col", [0, 1, 2])
@pytest.mark.parametrize("fixed", [True, False])
def test_q21_keys_beyond_the_dimension_never_match(col, fixed):
    # 5 % of the rows get a date, supplier or part key past the dimension's
    # last row: the compact probe drops them (the reference's modulo hash
    # tables leave such keys undefined, so the hash-table path is not run)
    n = 40_000 + 9
    fact, dims = T.o_gen_c3(13, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    rng = np.random.default_rng(6)
    bad = rng.random(n) < 0.05
    size = (SMALL["n_date"], SMALL["n_supp"], SMALL["n_part"])[col]
    g = list(fact)
    g[col] = np.where(bad, size + rng.integers(0, 50, n), fact[col]).astype(np.uint32)
    widths = (12, 16, 20, 24) if fixed else (13, 17, 21, 25)
    enc = [fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(g[k], widths[k]), bw=widths[k], base=1)
           for k in range(3)] + [fls.bitpack_column(g[3], widths[3])]
    d = dev_dims(dims, SMALL)
    cols = [c.to_device() for c in enc]
    out = fls.query_q21(*cols, n, d, fls.q21_build(d)).cpu().numpy()
    kept = [np.ascontiguousarray(x[~bad]) for x in fact]
    check(out, T.o_q21(kept, dims, int((~bad).sum())))
    if fixed:  # both reference-table probes wrap such keys by the same modulo
        pht, keep = fls.q21_build_pht(d, occupancy=False)
        pht_b, keep_b = fls.q21_build_pht(d)
        np.testing.assert_array_equal(fls.query_q21_pht(*cols, n, pht_b).cpu().numpy(),
                                      fls.query_q21_pht(*cols, n, pht).cpu().numpy())
        del keep, keep_b


def test_q21_runtime_widths():
    # keys packed wider than needed take the run-time-width kernels
    n = 30_000
    fact, dims = T.o_gen_c3(12, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    enc = [fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(fact[0], 13), bw=13, base=1),
           fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(fact[1], 17), bw=17, base=1),
           fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(fact[2], 21), bw=21, base=1),
           fls.bitpack_column(fact[3], 25)]
    want = T.o_q21(fact, dims, n)
    out_c, out_p = run_both(fact, dims, SMALL, n, enc)
    check(out_c, want)
    check(out_p, want)


@pytest.mark.parametrize("layout", [("delta", "for", "rle", "exhaustive"), ("for", "rle", "for", "delta")])
def test_q21_any_encoding_columns(layout):
    # the compact probe decodes delta, RLE and mixed (per-vector FOR) columns
    # at run time; rows sorted by date give the delta and RLE columns runs
    n = 50_000 + 3
    fact, dims = T.o_gen_c3(13, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    order = np.argsort(fact[0], kind="stable")
    fact = [x[order] for x in fact]
    vals = [fact[0] + 1, fact[1] + 1, fact[2] + 1, fact[3]]
    cols = [T.encode_any(x, kind) for x, kind in zip(vals, layout)]
    d = dev_dims(dims, SMALL)
    out = fls.query_q21(*cols, n, d, fls.q21_build(d)).cpu().numpy()
    check(out, T.o_q21(fact, dims, n))
    with pytest.raises(fls.FlsError):  # the reference-table probe reads BITPACK only
        pht, keep = fls.q21_build_pht(d)
        fls.query_q21_pht(*cols, n, pht)


def test_q21_transposed_columns_partial_vector():
    # transposed BITPACK columns with a partial last vector: the live rows of
    # that vector are scattered over the lanes' registers, and the padding rows
    # decode to key 1 (dimension row 0), which is made to qualify so that a
    # padding row counted by mistake changes the result
    n = 50_017
    fact, dims = T.o_gen_c3(16, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    dims["s_region"][0] = 2
    dims["p_category"][0] = 12
    widths = (12, 16, 20, 24)
    enc = [fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(fact[k], widths[k], transpose=True), bw=widths[k],
                          base=1, transposed=True) for k in range(3)]
    enc.append(fls.bitpack_column(fact[3], widths[3], transposed=True))
    want = T.o_q21(fact, dims, n)
    out_c, out_p = run_both(fact, dims, SMALL, n, enc)
    check(out_c, want)
    check(out_p, want)


@pytest.mark.parametrize("n", [40_000 + 3, 33 * 1024])
def test_q21_every_row_qualifies(n):
    # all suppliers in the region and all parts in the category: the per-wave
    # queue of qualifying rows fills and flushes many times per vector pair,
    # and a half-wave past the last vector must not drop queued rows
    fact, dims = T.o_gen_c3(14, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    dims["s_region"][:] = 2
    dims["p_category"][:] = 12
    want = T.o_q21(fact, dims, n)
    assert int(want[1].sum()) == n
    out_c, out_p = run_both(fact, dims, SMALL, n)
    check(out_c, want)
    check(out_p, want)


def test_q21_more_groups_than_the_block_hash():
    # every row qualifies at the SSB widths (the compact kernel): the brands of
    # 8,000 parts span up to 1,000 x 7 year groups, beyond the probe's
    # 1,024-entry LDS hash table, so most groups take the global-atomic path
    n = 200_000 + 7
    fact, dims = T.o_gen_c3(15, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    dims["s_region"][:] = 2
    dims["p_category"][:] = 12
    want = T.o_q21(fact, dims, n)
    assert (want[1] > 0).sum() > 2048
    widths = (12, 16, 20, 24)
    enc = [fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(fact[k], widths[k]), bw=widths[k], base=1)
           for k in range(3)] + [fls.bitpack_column(fact[3], widths[3])]
    out_c, out_p = run_both(fact, dims, SMALL, n, enc)
    check(out_c, want)
    check(out_p, want)


def test_q21_full_size_and_result_rows():
    n = 120_000_000
    fd = fls.gen_c3_fact(1, n)
    fact = [fd[k].cpu().numpy().view(np.uint32) for k in ("datefk", "suppfk", "partfk", "revenue")]
    del fd
    dims_d = fls.gen_c3_dims(1)
    dims_np = {k: v.cpu().numpy().view(np.uint32) for k, v in dims_d.items() if not k.startswith("n_")}
    want = T.o_q21(fact, dims_np, n)
    assert want[1].sum() > 800_000  # ~0.8 % of the rows survive
    sizes = {k: fls.C3_DIMS[k] for k in ("n_date", "n_supp", "n_part")}
    enc = fls.encode_c3_table(*fact)
    assert [c.bw for c in enc] == [12, 16, 20, 24]
    cols = [c.to_device() for c in enc]
    ws = fls.q21_build(dims_d)
    out = fls.query_q21(*cols, n, dims_d, ws)
    check(out.cpu().numpy(), want)
    pht, keep = fls.q21_build_pht(dims_d)
    out_p = fls.query_q21_pht(*cols, n, pht)
    check(out_p.cpu().numpy(), want)
    # result rows in slot order, as runQuery compacts them
    rows = fls.q21_groups(out)
    assert len(rows) == int((want[1] > 0).sum())
    slots = [(b * 7 + (y - 1992)) % 7000 for y, b, _ in rows]
    assert slots == sorted(slots)
    del sizes, keep


def test_q21_partial_tables_reused_and_concurrent():
    """The compact probe sums its blocks' groups through sets of partial
    tables that a second launch reads and clears: 20 calls in a row (more than
    the 8 sets, overwriting and adding, the compact and the bit-filtered
    reference-table probes), then 8 streams with calls in flight at once, each
    into its own output."""
    torch = T.torch_cuda()
    n = 300_000 + 11
    fact, dims = T.o_gen_c3(23, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    want = T.o_q21(fact, dims, n)
    d = dev_dims(dims, SMALL)
    cols = [c.to_device() for c in fls.encode_c3_table(*fact, n_date=SMALL["n_date"], n_supp=SMALL["n_supp"],
                                                         n_part=SMALL["n_part"])]
    ws = fls.q21_build(d)
    pht_b, keep_b = fls.q21_build_pht(d)
    acc = torch.zeros(14000, dtype=torch.int64, device="cuda")
    outs = []
    for k in range(20):
        if k % 2:
            outs.append(fls.query_q21_pht(*cols, n, pht_b))
        else:
            outs.append(fls.query_q21(*cols, n, d, ws))
        fls.query_q21(*cols, n, d, ws, accumulate=True, d_out=acc)
    for o in outs:
        check(o.cpu().numpy(), want)
    total = acc.cpu().numpy()
    np.testing.assert_array_equal(total[:7000], 20 * want[0])
    np.testing.assert_array_equal(total[7000:].view(np.uint64), 20 * want[1])
    streams = [torch.cuda.Stream() for _ in range(8)]
    torch.cuda.synchronize()
    res = []
    for s in streams:
        with torch.cuda.stream(s):
            out = torch.empty(14000, dtype=torch.int64, device="cuda")
            res.append(fls.query_q21(*cols, n, d, ws, d_out=out, stream=s.cuda_stream))
    torch.cuda.synchronize()
    for o in res:
        check(o.cpu().numpy(), want)
    del keep_b


@pytest.mark.parametrize("accumulate", [False, True])
def test_q21_more_slots_than_the_partial_tables(accumulate):
    """A slot table larger than the partial tables (8192 slots) adds to the
    output directly: the compact probe equals the reference-table probe per
    row on the same slots, and the totals equal the oracle's."""
    torch = T.torch_cuda()
    n = 200_000 + 5
    fact, dims = T.o_gen_c3(29, n, **SMALL)
    fact[1] %= SMALL["n_supp"]
    fact[2] %= SMALL["n_part"]
    want = T.o_q21(fact, dims, n)
    d = dev_dims(dims, SMALL)
    cols = [c.to_device() for c in fls.encode_c3_table(*fact, n_date=SMALL["n_date"], n_supp=SMALL["n_supp"],
                                                         n_part=SMALL["n_part"])]
    prm = fls.q21_params(n_slots=9001)
    ws = fls.q21_build(d, prm)
    pht, keep = fls.q21_build_pht(d, occupancy=False)
    start = torch.full((2 * 9001,), 5 if accumulate else -1, dtype=torch.int64, device="cuda")
    out_c = fls.query_q21(*cols, n, d, ws, params=prm, accumulate=accumulate, d_out=start.clone()).cpu().numpy()
    out_p = fls.query_q21_pht(*cols, n, pht, params=prm, accumulate=accumulate, d_out=start.clone()).cpu().numpy()
    np.testing.assert_array_equal(out_c, out_p)
    base = 5 * 9001 if accumulate else 0
    assert int(out_c[:9001].sum()) - base == int(want[0].sum())
    assert int(out_c[9001:].sum()) - base == int(want[1].sum())
    del keep
