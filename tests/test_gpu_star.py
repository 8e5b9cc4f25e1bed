"""GPU parity of the fused decode + star joins + group-by (Q3.1 and Q4.1
shapes) against the oracle."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
SMALL = dict(n_date=2556, n_supp=400, n_part=8000, n_cust=600)
DIM_KEYS = ("d_year", "s_region", "s_nation", "c_region", "c_nation", "p_mfgr")


def small_table(seed, n, sizes=SMALL):
    f, d = T.o_gen_star(seed, n, **sizes)
    f["suppfk"] %= sizes["n_supp"]
    f["partfk"] %= sizes["n_part"]
    f["custfk"] %= sizes["n_cust"]
    return f, d


def dev_dims(d_np, sizes):
    torch = T.torch_cuda()
    d = {k: torch.from_numpy(d_np[k].view(np.int32).copy()).cuda() for k in DIM_KEYS}
    d.update(sizes)
    return d


FAST = (12, 20, 16, 20, 24, 16)  # the widths of the 120M-row table: fixed-width kernels


def encode(f, sizes, widths=None, transposed=False):
    """Minimal widths (widths None) or the given widths, in either row order."""
    if widths is None and not transposed:
        return fls.encode_star_table(*(f[k] for k in T.STAR_NAMES), n_date=sizes["n_date"],
                                     n_cust=sizes["n_cust"], n_supp=sizes["n_supp"], n_part=sizes["n_part"])
    widths = widths or [max(1, int(sizes[k] - 1).bit_length()) for k in ("n_date", "n_cust", "n_supp", "n_part")] \
        + [25, 17]
    n = f["datefk"].size
    cols = [fls.HostColumn(fls.ENC_BITPACK, n, fls.pack_column(f[k], bw, transpose=transposed), bw=bw, base=1,
                           transposed=transposed)
            for k, bw in zip(("datefk", "custfk", "suppfk", "partfk"), widths)]
    cols.append(fls.bitpack_column(f["revenue"], widths[4], transposed=transposed))
    cols.append(fls.bitpack_column(f["supplycost"], widths[5], transposed=transposed, base=fls.SUPPLYCOST_BASE))
    return cols


def run_q31(f, d_np, sizes, n, enc=None):
    d = dev_dims(d_np, sizes)
    cols = [c.to_device() for c in (enc or encode(f, sizes))]
    ws = fls.q31_build(d)
    return fls.query_q31(cols[0], cols[1], cols[2], cols[4], n, d, ws).cpu().numpy()


def run_q41(f, d_np, sizes, n, enc=None):
    d = dev_dims(d_np, sizes)
    cols = [c.to_device() for c in (enc or encode(f, sizes))]
    ws = fls.q41_build(d)
    return fls.query_q41(*cols, n, d, ws).cpu().numpy()


def check(out, want):
    sums, counts = want
    k = sums.size
    np.testing.assert_array_equal(out[k:].view(np.uint64), counts)
    np.testing.assert_array_equal(out[:k], sums)


def test_generators_match_oracle():
    n, n_part = 70_001, 80_000
    fd = fls.gen_c3_fact(6, n, with_customer=True)
    ex = fls.gen_star_extra(6, n, n_part)
    f, d = T.o_gen_star(6, n, n_part=n_part)
    np.testing.assert_array_equal(fd["custfk"].cpu().numpy().view(np.uint32), f["custfk"])
    np.testing.assert_array_equal(ex["supplycost"].cpu().numpy().view(np.uint32), f["supplycost"])
    np.testing.assert_array_equal(ex["p_mfgr"].cpu().numpy().view(np.uint32), d["p_mfgr"])


@pytest.mark.parametrize("n", [1, 700, 1024, 50_000 + 17])
def test_q31_small(n):
    f, d = small_table(31, n)
    check(run_q31(f, d, SMALL, n, encode(f, SMALL, FAST)), T.o_q31(f, d, n))


@pytest.mark.parametrize("n", [1, 700, 1024, 50_000 + 17])
def test_q41_small(n):
    f, d = small_table(41, n)
    check(run_q41(f, d, SMALL, n, encode(f, SMALL, FAST)), T.o_q41(f, d, n))


@pytest.mark.parametrize("transposed", [False, True])
def test_runtime_widths_and_row_order(transposed):
    # minimal key widths (9-13 bits here) take the run-time-width kernels
    n = 30_000 + 5
    f, d = small_table(32, n)
    enc = encode(f, SMALL, None, transposed)
    assert enc[2].bw == 9
    check(run_q31(f, d, SMALL, n, enc), T.o_q31(f, d, n))
    check(run_q41(f, d, SMALL, n, enc), T.o_q41(f, d, n))


def test_transposed_fixed_widths():
    n = 20_000 + 3
    f, d = small_table(33, n)
    enc = encode(f, SMALL, FAST, transposed=True)
    check(run_q31(f, d, SMALL, n, enc), T.o_q31(f, d, n))
    check(run_q41(f, d, SMALL, n, enc), T.o_q41(f, d, n))


@pytest.mark.parametrize("sizes,widths", [
    # supplier (Q3.1) and part (Q4.1) codes too large to stage: run-time widths
    (dict(n_date=2556, n_supp=120_000, n_part=1_100_000, n_cust=600), None),
    # part bits too large to stage next to the others (Q4.1), fixed widths
    (dict(n_date=2556, n_supp=40_000, n_part=1 << 20, n_cust=600), FAST),
])
def test_dimensions_beyond_lds(sizes, widths):
    n = 40_000 + 9
    f, d = small_table(34, n, sizes)
    enc = encode(f, sizes, widths)
    check(run_q31(f, d, sizes, n, enc), T.o_q31(f, d, n))
    check(run_q41(f, d, sizes, n, enc), T.o_q41(f, d, n))


def fold_q31(out, n_slots):
    """A Q3.1 result over n_slots >= 4375 folded onto the oracle's 4375 slots:
    every group value V = s*175 + c*7 + y is below 4,556, so it sits in slot V
    of the larger table and in slot V % 4375 of the oracle's."""
    sums, counts = np.zeros(4375, np.int64), np.zeros(4375, np.uint64)
    np.add.at(sums, np.arange(n_slots) % 4375, out[:n_slots])
    np.add.at(counts, np.arange(n_slots) % 4375, out[n_slots:].view(np.uint64))
    return np.concatenate([sums, counts.view(np.int64)])


@pytest.mark.parametrize("sizes,n_slots", [
    # SSB widths, customer bitmap too large to stage: group table, supplier
    # and date codes in LDS, customer nations from L2
    (dict(n_date=2556, n_supp=40_000, n_part=8000, n_cust=1_000_000), 4375),
    # and with 8,192 group slots only the group table fits in LDS
    (dict(n_date=2556, n_supp=65_000, n_part=8000, n_cust=1_000_000), 8192),
])
def test_q31_fixed_widths_without_customer_bitmap(sizes, n_slots):
    n = 60_000 + 11
    f, d = small_table(35, n, sizes)
    dd = dev_dims(d, sizes)
    cols = [c.to_device() for c in encode(f, sizes, FAST)]
    ws = fls.q31_build(dd)
    out = fls.query_q31(cols[0], cols[1], cols[2], cols[4], n, dd, ws, params=fls.q31_params(n_slots=n_slots))
    check(fold_q31(out.cpu().numpy(), n_slots), T.o_q31(f, d, n))


def test_q31_revenue_with_a_base():
    # a frame-of-reference base on the 24-bit revenue column: values above 24
    # bits once the base is added back, carried whole through the row queue
    n = 30_000 + 7
    f, d = small_table(37, n)
    f["revenue"] = (f["revenue"].astype(np.uint64) + (1 << 24)).astype(np.uint32)
    enc = encode(f, SMALL, FAST)
    enc[4] = fls.bitpack_column(f["revenue"], 24, base=1 << 24)
    check(run_q31(f, d, SMALL, n, enc), T.o_q31(f, d, n))


@pytest.mark.parametrize("col", ["custfk", "suppfk", "datefk", "partfk"])
@pytest.mark.parametrize("widths", [FAST, None])
def test_keys_beyond_the_dimension_never_match(col, widths):
    # 5 % of the rows get a key past the dimension's last row: those rows
    # match nothing, so the result is the oracle's over the other rows
    n = 40_000 + 13
    f, d = small_table(38, n)
    rng = np.random.default_rng(5)
    bad = rng.random(n) < 0.05
    size = {"custfk": "n_cust", "suppfk": "n_supp", "datefk": "n_date", "partfk": "n_part"}[col]
    g = dict(f)
    g[col] = np.where(bad, SMALL[size] + rng.integers(0, 50, n), f[col]).astype(np.uint32)
    enc = encode(g, SMALL, widths)
    kept = {k: np.ascontiguousarray(v[~bad]) for k, v in f.items()}
    want31 = T.o_q31(f, d, n) if col == "partfk" else T.o_q31(kept, d, int((~bad).sum()))
    check(run_q31(g, d, SMALL, n, enc), want31)
    check(run_q41(g, d, SMALL, n, enc), T.o_q41(kept, d, int((~bad).sum())))


def test_q31_more_groups_than_the_block_hash():
    # every supplier and customer in the filtered region: 25 x 25 nations x 6
    # years = 3,750 live groups, beyond the probe's dense group table (760
    # groups) and its 512-entry LDS hash table, so the probe runs the hash
    # and most groups take the global-atomic path
    n = 200_000 + 3
    f, d = small_table(36, n)
    d["s_region"][:] = 3
    d["c_region"][:] = 3
    sums, counts = T.o_q31(f, d, n)
    assert (counts > 0).sum() > 3000
    check(run_q31(f, d, SMALL, n, encode(f, SMALL, FAST)), (sums, counts))


ANY_LAYOUTS = [
    # (date, cust, supp, part, revenue, supplycost); keys are encoded 1-based
    ("delta", "for", "rle", "exhaustive", "bitpack_t", "for"),
    ("for", "exhaustive", "bitpack_t", "for", "rle", "delta"),
]


@pytest.mark.parametrize("layout", ANY_LAYOUTS)
@pytest.mark.parametrize("sort_by_date", [False, True])
def test_any_encoding_columns(layout, sort_by_date):
    # per-vector FOR keys (fls_q41_bitpacked_opt_v4.cu:29-120 reads custkey
    # and orderdate with a width, base and offset per vector), delta, RLE and
    # mixed columns take the run-time-encoding kernels; sorting the fact rows
    # by date gives the delta and RLE columns long runs
    n = 60_000 + 77
    f, d = small_table(36, n)
    if sort_by_date:
        order = np.argsort(f["datefk"], kind="stable")
        f = {k: v[order] for k, v in f.items()}
    vals = [f["datefk"] + 1, f["custfk"] + 1, f["suppfk"] + 1, f["partfk"] + 1, f["revenue"], f["supplycost"]]
    cols = [T.encode_any(x, kind) for x, kind in zip(vals, layout)]
    dd = dev_dims(d, SMALL)
    out = fls.query_q31(cols[0], cols[1], cols[2], cols[4], n, dd, fls.q31_build(dd)).cpu().numpy()
    check(out, T.o_q31(f, d, n))
    out = fls.query_q41(*cols, n, dd, fls.q41_build(dd)).cpu().numpy()
    check(out, T.o_q41(f, d, n))


def test_accumulate_and_errors():
    torch = T.torch_cuda()
    n = 10_000
    f, d_np = small_table(35, n)
    d = dev_dims(d_np, SMALL)
    cols = [c.to_device() for c in encode(f, SMALL)]
    ws = fls.q41_build(d)
    out = fls.query_q41(*cols, n, d, ws)
    fls.query_q41(*cols, n, d, ws, accumulate=True, d_out=out)
    sums, counts = T.o_q41(f, d_np, n)
    check(out.cpu().numpy(), (2 * sums, 2 * counts))
    with pytest.raises(fls.FlsError):
        fls.query_q41(*cols, n, d, ws, params=fls.q41_params(n_slots=0))
    with pytest.raises(fls.FlsError):  # one column in another row order
        bad = fls.bitpack_column(f["revenue"], 24, transposed=True).to_device()
        fls.query_q41(cols[0], cols[1], cols[2], cols[3], bad, cols[5], n, d, ws)
    with pytest.raises(fls.FlsError):  # an empty dimension
        fls.query_q41(*cols, n, dict(d, n_cust=0), ws)
    with pytest.raises(fls.FlsError):  # workspace too small
        small = torch.empty(16, dtype=torch.uint8, device="cuda")
        ds = fls.star_dims_struct(d)
        fls.call("fls_q41_build", fls.ctypes.byref(ds), fls.ctypes.byref(fls.q41_params()), small.data_ptr(), 16,
                 None)


def test_full_size():
    torch = T.torch_cuda()
    n = 120_000_000
    fd = fls.gen_c3_fact(3, n, with_customer=True)
    ex = fls.gen_star_extra(3, n)
    f = {k: fd[k].cpu().numpy().view(np.uint32) for k in ("datefk", "custfk", "suppfk", "partfk", "revenue")}
    f["supplycost"] = ex["supplycost"].cpu().numpy().view(np.uint32)
    del fd
    dims = fls.gen_c3_dims(3)
    dims["p_mfgr"] = ex.pop("p_mfgr")
    d_np = {k: dims[k].cpu().numpy().view(np.uint32) for k in DIM_KEYS}
    enc = encode(f, fls.C3_DIMS)
    assert [c.bw for c in enc] == [12, 20, 16, 20, 24, 16]
    cols = [c.to_device() for c in enc]
    del enc
    want31 = T.o_q31(f, d_np, n)
    assert want31[1].sum() > 3_000_000  # 1/5 * 1/5 * 6/7 of the rows
    out = fls.query_q31(cols[0], cols[1], cols[2], cols[4], n, dims, fls.q31_build(dims))
    check(out.cpu().numpy(), want31)
    want41 = T.o_q41(f, d_np, n)
    assert want41[1].sum() > 1_500_000  # 1/5 * 1/5 * 2/5 of the rows
    out = fls.query_q41(*cols, n, dims, fls.q41_build(dims))
    check(out.cpu().numpy(), want41)
    torch.cuda.synchronize()
