"""Pins the CPU oracle before anything is checked against it.

Each check is independent of the oracle's own formulas: the reference's golden
vectors and its compiled encoders (oracle/_ref), closed forms a mistake would
break, Python's datetime, and brute force.
"""
import ctypes
import datetime

import numpy as np
import pytest

import fls_testlib as T

RNG = np.random.default_rng(7)


@pytest.mark.parametrize("variant", ["normal", "fused"])
def test_pack_matches_reference_golden(variant):
    g = T.golden()
    rand, packed = g[f"rand_{variant}"], g[f"packed_{variant}"]
    for bw in range(33):
        np.testing.assert_array_equal(T.o_pack(rand[bw], bw), packed[bw][: 32 * bw], err_msg=f"bw={bw}")


@pytest.mark.parametrize("variant", ["normal", "fused"])
def test_unpack_golden_is_reference_kat(variant):
    # cuda_*_unpack_test.cu:37-366: unpack(pack(rand_arr_bw)) == rand_arr_bw
    g = T.golden()
    rand, packed = g[f"rand_{variant}"], g[f"packed_{variant}"]
    for bw in range(33):
        mask = (1 << bw) - 1
        assert (rand[bw].astype(np.uint64) & mask == rand[bw]).all()
        np.testing.assert_array_equal(T.o_unpack(packed[bw], bw), rand[bw], err_msg=f"bw={bw}")


def test_pack_matches_reference_build_random():
    r = T.ref()
    for bw in range(33):
        for _ in range(3):
            x = RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)  # masking exercised
            want = np.zeros(1024, np.uint32)
            r.ref_pack_u32(x, want, bw)
            np.testing.assert_array_equal(T.o_pack(x, bw), want[: 32 * bw], err_msg=f"bw={bw}")


def test_bw0_and_bw32_special_cases():
    x = RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)
    assert T.o_pack(x, 0).size == 0
    assert (T.o_unpack(np.zeros(0, np.uint32), 0) == 0).all()
    np.testing.assert_array_equal(T.o_pack(x, 32), x)  # unpack.cuh:3256-3329 is a copy
    np.testing.assert_array_equal(T.o_unpack(x, 32), x)


@pytest.mark.parametrize("bw", [8, 16])
def test_byte_and_halfword_widths(bw):
    # at bw 8 / 16 the rows of a lane are the bytes / halfwords of its words
    x = RNG.integers(0, 2**bw, 1024).astype(np.uint32)
    packed = T.o_pack(x, bw)
    per = 32 // bw
    for lane in (0, 5, 31):
        for w in range(bw):
            word = int(packed[32 * w + lane])
            for k in range(per):
                assert (word >> (bw * k)) & ((1 << bw) - 1) == x[32 * (per * w + k) + lane]


def test_transpose_i_matches_reference():
    g = T.golden()
    ar = np.arange(1024, dtype=np.uint32)
    np.testing.assert_array_equal(T.o_perm("orc_transpose_i_u32", ar), g["transpose_i_perm"])
    # textbook statement: input 64a+8b+c lands at a + 16*bitrev3(b) + 128*c
    perm = g["transpose_i_perm"]
    for k in (0, 1, 7, 8, 16, 24, 32, 63, 64, 65, 1023):
        a, b, c = k // 64, (k // 8) % 8, k % 8
        rb = int(f"{b:03b}"[::-1], 2)
        assert perm[a + 16 * rb + 128 * c] == k


def test_transpose_o_is_inverse_and_reference_defect():
    x = RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)
    t = T.o_perm("orc_transpose_i_u32", x)
    np.testing.assert_array_equal(T.o_perm("orc_transpose_o_u32", t), x)
    np.testing.assert_array_equal(T.o_perm("orc_transpose_i_u32", T.o_perm("orc_transpose_o_u32", x)), x)
    # The reference's transpose_o (transpose.cpp:5136) repeats transpose_i: it
    # is not an untranspose, so the oracle deliberately differs from it.
    g = T.golden()
    np.testing.assert_array_equal(g["transpose_o_ref"], g["transpose_i_perm"])


def test_unrsum_matches_reference_golden():
    g = T.golden()
    for k in range(g["unrsum_in"].shape[0]):
        np.testing.assert_array_equal(T.o_perm("orc_unrsum_u32", g["unrsum_in"][k]), g["unrsum_out"][k])


def test_unrsum_matches_reference_build_random():
    r = T.ref()
    x = RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32)
    want = np.zeros(1024, np.uint32)
    r.ref_unrsum_u32(x, want)
    np.testing.assert_array_equal(T.o_perm("orc_unrsum_u32", x), want)


def test_rsum_order_matches_reference():
    order = [T.oracle().orc_rsum_offset(j) for j in range(32)]
    np.testing.assert_array_equal(order, T.golden()["rsum_order"])


def test_rsum_inverts_unrsum():
    for x in (RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32),
              np.sort(RNG.integers(0, 2**31, 1024)).astype(np.uint32)):
        d = T.o_perm("orc_unrsum_u32", x)
        out = np.zeros(1024, np.uint32)
        T.oracle().orc_rsum_u32(d, out, np.ascontiguousarray(x[:32]))
        np.testing.assert_array_equal(out, x)


def test_delta_encode_matches_reference_bench():
    # fastlanes_bench_delta.cu:131-155 (value i%1024) encoded by the reference
    g = T.golden()
    x = np.arange(1024, dtype=np.uint32)
    packed, bases, bw = T.o_delta_encode(x)
    assert bw == int(g["delta_i1024_bw"][0])
    np.testing.assert_array_equal(bases, g["delta_i1024_bases"])
    np.testing.assert_array_equal(packed, g["delta_i1024_packed"][: 32 * bw])
    # decode is in transposed order (fastlanes_bench_delta.cu:183-191)
    np.testing.assert_array_equal(T.o_delta_decode(packed, bw, bases), T.o_perm("orc_transpose_i_u32", x))


def test_delta_roundtrip_sorted_and_wrapping():
    for x in (np.sort(RNG.integers(0, 2**32, 1024, dtype=np.uint64)).astype(np.uint32),
              RNG.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32),
              np.full(1024, 77, np.uint32)):
        packed, bases, bw = T.o_delta_encode(x)
        np.testing.assert_array_equal(T.o_delta_decode(packed, bw, bases), T.o_perm("orc_transpose_i_u32", x))
    packed, bases, bw = T.o_delta_encode(np.full(1024, 77, np.uint32))
    assert bw == 0


def test_bit_width_and_find_bw4():
    o = T.oracle()
    for x in (0, 1, 2, 3, 15, 16, 255, 256, 2**31 - 1, 2**31, 2**32 - 1):
        assert o.orc_bit_width(x) == int(x).bit_length()
    # ssb_utils.h:222-239: 0 -> 0, 1..4 -> 4, 5..8 -> 8, ..., 29..32 -> 32
    for w in range(33):
        arr = np.zeros(1024, np.uint32)
        if w:
            arr[3] = (1 << (w - 1))
        want = 0 if w == 0 else 4 * ((w + 3) // 4)
        assert o.orc_find_bw4(arr, 1024) == want, w


def test_checksum_brute_force():
    v = RNG.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32)
    want = sum((i + 1) * int(x) for i, x in enumerate(v)) % 2**64
    assert T.o_checksum(v) == want


def test_yyyymmdd_against_datetime():
    o = T.oracle()
    d0 = datetime.date(1992, 1, 1)
    for d in range(2557):
        want = int((d0 + datetime.timedelta(days=d)).strftime("%Y%m%d"))
        assert o.orc_yyyymmdd(d) == want
    assert o.orc_yyyymmdd(2556) == 19981231


def test_uniform_range_and_determinism():
    o = T.oracle()
    vals = [o.orc_uniform(5, 3, i, 11) for i in range(20000)]
    assert min(vals) == 0 and max(vals) == 10
    counts = np.bincount(vals, minlength=11)
    assert counts.min() > 1500  # roughly uniform
    assert vals[:50] == [o.orc_uniform(5, 3, i, 11) for i in range(50)]
    assert o.orc_uniform(5, 3, 0, 2**32) == o.orc_rand64(5, 3, 0) >> 32


def test_gen_c2_columns():
    n = 300_000
    od, q, d, p = T.o_gen_c2(1, n)
    assert (np.diff(od.astype(np.int64)) >= 0).all() and od[0] == 19920101 and od[-1] == 19981231
    assert q.min() == 1 and q.max() == 50 and d.min() == 0 and d.max() == 10
    assert p.min() >= 90000 and p.max() < 10**7
    assert int(p.max()).bit_length() == 24 and int(q.max()).bit_length() == 6 and int(d.max()).bit_length() == 4


def test_q11_brute_force():
    n = 50_000
    od, q, d, p = T.o_gen_c2(3, n)
    c = ctypes.c_uint64(0)
    s = T.oracle().orc_q11(od, q, d, p, n, 19930000, 19940000, 25, 1, 3, ctypes.byref(c))
    m = (od > 19930000) & (od < 19940000) & (q < 25) & (d >= 1) & (d <= 3)
    assert s == int((p[m].astype(np.int64) * d[m].astype(np.int64)).sum())
    assert c.value == int(m.sum()) and c.value > 0


def test_q21_brute_force():
    o = T.oracle()
    n, nd, ns, np_, nc = 40_000, 2556, 400, 8000, 600
    fk = [np.zeros(n, np.uint32) for _ in range(4)]
    o.orc_gen_c3_fact(9, n, fk[0], fk[1], fk[2], None, fk[3])
    datefk, suppfk, partfk, rev = fk
    suppfk %= ns
    partfk %= np_
    dims = [np.zeros(k, np.uint32) for k in (nd, ns, ns, np_, np_, nc, nc)]
    o.orc_gen_c3_dims(9, nd, ns, np_, nc, *dims)
    d_year, s_region, _, p_cat, p_brand, _, _ = dims
    sums = np.zeros(7000, np.int64)
    counts = np.zeros(7000, np.uint64)
    o.orc_q21(datefk, suppfk, partfk, rev, n, d_year, s_region, p_cat, p_brand, 12, 2, sums, counts)
    want = {}
    for i in range(n):
        if s_region[suppfk[i]] == 2 and p_cat[partfk[i]] == 12:
            key = (int(d_year[datefk[i]]), int(p_brand[partfk[i]]))
            want[key] = want.get(key, 0) + int(rev[i])
    got = {}
    for slot in np.nonzero(counts)[0]:
        year, brand = 1992 + slot % 7, slot // 7 if slot >= 7 else 1000
        got[(year, int(brand))] = int(sums[slot])
    assert got == want and len(want) > 10
    assert set(d_year.tolist()) == set(range(1992, 1999))


def _star_small(seed, n):
    sizes = dict(n_date=2556, n_supp=400, n_part=8000, n_cust=600)
    f, d = T.o_gen_star(seed, n, **sizes)
    f["suppfk"] %= sizes["n_supp"]
    f["partfk"] %= sizes["n_part"]
    f["custfk"] %= sizes["n_cust"]
    return f, d


def test_q31_brute_force():
    # group by (supplier nation, customer nation, year) in plain Python, then
    # decode each non-empty oracle slot back to its group: with nations 1..25,
    # V = 25*s_nation + c_nation runs over 26..650 and the slot keeps V mod 625,
    # so V = q below 26 means q + 625 (e.g. (24, 25) wraps to slots 0..6)
    n = 40_000
    f, d = _star_small(21, n)
    sums, counts = T.o_q31(f, d, n)
    want, want_n = {}, {}
    for i in range(n):
        s, c, y = int(f["suppfk"][i]), int(f["custfk"][i]), int(d["d_year"][f["datefk"][i]])
        if d["s_region"][s] == 3 and d["c_region"][c] == 3 and 1992 <= y <= 1997:
            key = (int(d["s_nation"][s]), int(d["c_nation"][c]), y)
            want[key] = want.get(key, 0) + int(f["revenue"][i])
            want_n[key] = want_n.get(key, 0) + 1
    got, got_n = {}, {}
    for slot in np.nonzero(counts)[0]:
        y = 1992 + int(slot) % 7
        q = int(slot) // 7
        v = q if q >= 26 else q + 625
        cn = (v - 1) % 25 + 1
        key = ((v - cn) // 25, cn, y)
        got[key] = int(sums[slot])
        got_n[key] = int(counts[slot])
    assert got == want and got_n == want_n and len(want) > 100
    assert max(k[2] for k in want) == 1997  # 1998 is filtered out
    assert {k[0] for k in want} | {k[1] for k in want} <= set(range(11, 16))  # region 3's nations


def test_q41_brute_force():
    n = 40_000
    f, d = _star_small(22, n)
    sums, counts = T.o_q41(f, d, n)
    want = {}
    for i in range(n):
        s, c, p = int(f["suppfk"][i]), int(f["custfk"][i]), int(f["partfk"][i])
        if d["s_region"][s] == 2 and d["c_region"][c] == 2 and 1 <= d["p_mfgr"][p] <= 2:
            key = (int(d["c_nation"][c]), int(d["d_year"][f["datefk"][i]]))
            want[key] = want.get(key, 0) + int(f["revenue"][i]) - int(f["supplycost"][i])
    got = {}
    for slot in np.nonzero(counts)[0]:
        cn = int(slot) // 7
        got[(cn if cn > 0 else 25, 1992 + int(slot) % 7)] = int(sums[slot])
    assert got == want and len(want) > 25
    assert {k[0] for k in want} <= set(range(6, 11))  # region 2's nations
    assert (f["revenue"] < f["supplycost"]).any()  # some rows add a negative profit


def test_star_extra_ranges():
    f, d = T.o_gen_star(5, 100_000, n_part=50_000)
    assert f["supplycost"].min() >= 60000 and f["supplycost"].max() < 120000
    assert f["supplycost"].max() - f["supplycost"].min() > 59000
    assert set(np.unique(d["p_mfgr"]).tolist()) == {1, 2, 3, 4, 5}
    # custfk is drawn by the same generator as the Q2.1 table's keys
    assert f["custfk"].max() < 600000 and len(np.unique(f["custfk"])) > 90_000


def test_c3_dims_follow_ssb_hierarchies():
    # a nation lies in one region, a brand in one category (SSB dbgen's
    # nation -> region and p_category -> p_brand1)
    _, d = T.o_gen_c3(4, 1, n_supp=20_000, n_part=50_000, n_cust=20_000)
    for pre in ("s", "c"):
        nat, reg = d[f"{pre}_nation"], d[f"{pre}_region"]
        assert set(np.unique(nat).tolist()) == set(range(1, 26))
        np.testing.assert_array_equal(reg, (nat - 1) // 5 + 1)
    cat, brand = d["p_category"], d["p_brand"]
    assert set(np.unique(cat).tolist()) == set(range(1, 26))
    assert ((brand - 1) // 40 + 1 == cat).all() and set(np.unique(brand).tolist()) == set(range(1, 1001))


def test_rle_roundtrip_and_run_count():
    o = T.oracle()
    for n in (1024, 1000, 1, 513):
        starts = RNG.random(1024) < 1 / 8
        starts[0] = True
        vals = RNG.integers(0, 2**13, 1024).astype(np.uint32)
        x = vals[np.cumsum(starts) - 1].astype(np.uint32)
        hdr = T.OrcRle()
        values = np.zeros(1024, np.uint32)
        idx = np.zeros(1024, np.uint32)
        o.orc_rle_encode_u32(x, n, hdr, values, idx)
        full = np.concatenate([x[:n], np.full(1024 - n, x[n - 1], np.uint32)])
        runs = 1 + int((full[1:] != full[:-1]).sum())
        assert hdr.n_runs == runs
        assert hdr.ibw == (1 if runs > 1 else 0)
        out = np.zeros(1024, np.uint32)
        o.orc_rle_decode_u32(hdr, values, idx, out)
        np.testing.assert_array_equal(out, T.o_perm("orc_transpose_i_u32", full))


def test_c4_workload_distribution():
    """The batched-segments workload: 40/30/30 bitpack/delta/RLE, widths 1..32,
    lengths log-uniform by octave in [2^10, 2^24), sorted delta segments, RLE
    runs of geometric mean 8, range = middle half of [min, max]."""
    import ctypes

    T.o_c4_counts(1, 1, 10, 24, 1)  # sets argtypes
    o = T.oracle()
    enc, bw, ln = [], [], []
    for s in range(4096):
        e, b, n = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        o.orc_c4_params(1, s, 10, 24, ctypes.byref(e), ctypes.byref(b), ctypes.byref(n))
        enc.append(e.value)
        bw.append(b.value)
        ln.append(n.value)
    enc, bw, ln = np.array(enc), np.array(bw), np.array(ln)
    frac = np.bincount(enc, minlength=3) / 4096
    assert abs(frac[0] - 0.4) < 0.03 and abs(frac[1] - 0.3) < 0.03 and abs(frac[2] - 0.3) < 0.03
    assert bw.min() == 1 and bw.max() == 32
    assert ln.min() >= 1 << 10 and ln.max() < 1 << 24
    octave = np.floor(np.log2(ln)).astype(int)
    counts = np.bincount(octave, minlength=24)[10:24]
    assert counts.min() > 4096 / 14 * 0.7  # every octave about equally likely
    small = [s for s in range(4096) if ln[s] <= 1 << 13][:40]
    runs = []
    c_all, lo_all, hi_all = T.o_c4_counts(1, max(small) + 1, 10, 24, 8)
    for s in small:
        e, b, v = T.o_c4_values(1, s, 10, 24)
        assert v.size == ln[s]
        if e != 1:
            assert int(v.max()) < (1 << b)
        else:  # sorted; consecutive steps below 2^bw
            d = np.diff(v.astype(np.int64))
            assert np.all(d >= 0) and int(d.max()) < (1 << b)
        if e == 2:
            runs.append(v.size / (1 + np.count_nonzero(np.diff(v.astype(np.int64)))))
        lo = int(v.min()) + (int(v.max()) - int(v.min())) // 4
        hi = int(v.min()) + (int(v.max()) - int(v.min())) * 3 // 4
        assert (lo_all[s], hi_all[s]) == (lo, hi)
        assert c_all[s] == np.count_nonzero((v >= lo) & (v <= hi))
    assert runs and 5 < np.mean(runs) < 11  # geometric mean 8 (equal neighbours merge)
