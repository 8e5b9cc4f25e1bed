"""Reference-shaped kernels compiled against include/fls/device alone.

tests/dropin/refshape.hip restates the reference's own query and decode
kernels call for call (QueryKernel_v1/v2, fls_q11.cu:29-135; probe_v1/v2/v3 and
build_hashtable_s/p/d, fls_q21.cu:21-263; Q3.1 probe_v1/v2 with
BlockProbeAndPHT_2_R_S, fls_q31.cu:27-160; the delta kernels with d_rsum_32,
delta_shared_memory.cu:24-60 and delta_global_memory.cu:25-40), launched the
reference's way (one tile per 32-wide workgroup).  Each is checked against the
oracle.  A primitive catalogue then runs every predicate, load, store, sum and
IsTerm form at five tile shapes against the textbook rule, and the packed
16 x 2 predicates against their per-halfword definition."""
import ctypes

import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu

LIB = T.ROOT / "tests" / "dropin" / "librefshape.so"
SMALL = dict(n_date=2556, n_supp=400, n_part=8000, n_cust=6000)


@pytest.fixture(scope="module")
def rs():
    T.torch_cuda()
    if not LIB.exists():
        pytest.fail(f"{LIB} missing: build it with `make -C tests/dropin` (done by __graft_entry__.build())")
    lib = ctypes.CDLL(str(LIB))
    return lib


def P(t):
    return ctypes.c_void_p(t.data_ptr())


def sync():
    T.torch_cuda().cuda.synchronize()


def i32(a):
    return T.to_dev(np.ascontiguousarray(a, np.uint32))


def zeros(n, dtype="i32"):
    torch = T.torch_cuda()
    return torch.zeros(max(1, n), dtype={"i32": torch.int32, "i64": torch.int64}[dtype], device="cuda")


def bw_of(v):
    return int(v.max()).bit_length() if v.size else 0


# ------------------------------------------------------------------ Q1.1
@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("n", [1024, 300_000 + 77])
def test_q11_reference_kernel(rs, version, n):
    od, qty, disc, price = T.o_gen_c2(21, n)
    want, _ = T.o_q11([od, qty, disc, price], n)
    od_min = int(od.min())
    od_for = (od - np.uint32(od_min)).astype(np.uint32)
    bws = [bw_of(od_for), 6, 4, 24]
    d_od, d_q, d_d = (i32(fls.pack_column(c, b)) for c, b in zip((od_for, qty, disc), bws))
    d_p = i32(fls.pack_column(price, 24) if version == 1 else price)
    rev = zeros(1, "i64")
    rc = rs.rs_q11(version, P(d_od), P(d_d), P(d_q), P(d_p), n, bws[0], 6, 4, 24, od_min, P(rev))
    assert rc == 0
    sync()
    assert int(rev.cpu().numpy()[0]) == want


# ------------------------------------------------------------------ Q2.1
def q21_inputs(n, sizes, seed=31):
    fact, dims = T.o_gen_c3(seed, n, **sizes)
    fact[1] %= sizes["n_supp"]
    fact[2] %= sizes["n_part"]
    return fact, dims


@pytest.mark.parametrize("version", [1, 2, 3])
@pytest.mark.parametrize("n", [256, 70_000 + 13])
def test_q21_reference_kernels(rs, version, n):
    fact, dims = q21_inputs(n, SMALL)
    want_sums, want_counts = T.o_q21(fact, dims, n)
    datefk, suppfk, partfk, rev = fact
    # 1-based keys packed as they are (the reference's SSB keys), at multiples
    # of 4 so the quarter-vector decode applies
    bws = (12, 16, 20)
    keys = [(k + 1).astype(np.uint32) for k in (datefk, suppfk, partfk)]
    d_date, d_supp, d_part = (i32(fls.pack_column(k, b)) for k, b in zip(keys, bws))
    d_rev = i32(fls.pack_column(rev, 24) if version == 1 else rev)
    ns, npart, nd = SMALL["n_supp"], SMALL["n_part"], SMALL["n_date"]
    s_region, p_cat, p_brand, d_year = (i32(dims[k]) for k in ("s_region", "p_category", "p_brand", "d_year"))
    s_key, p_key, d_key = (i32(np.arange(1, m + 1, dtype=np.uint32)) for m in (ns, npart, nd))
    ht_s, ht_p, ht_d = zeros(2 * ns), zeros(2 * npart), zeros(2 * nd)
    res = zeros(7000 * 4)
    rc = rs.rs_q21(version, P(d_date), P(d_part), P(d_supp), P(d_rev), n, 16, 20, 12, 24, P(s_region), P(s_key), ns,
                   2, P(p_cat), P(p_key), P(p_brand), npart, 12, P(d_key), P(d_year), nd, 1, P(ht_s), P(ht_p),
                   P(ht_d), P(res))
    assert rc == 0
    sync()
    r = res.cpu().numpy().reshape(7000, 4)
    sums = np.ascontiguousarray(r[:, 2:4]).view(np.int64)[:, 0]
    np.testing.assert_array_equal(sums, want_sums)
    hit = want_counts > 0
    assert hit.sum() > 0
    slots = np.nonzero(hit)[0]
    np.testing.assert_array_equal(r[slots, 0], 1992 + slots % 7)
    np.testing.assert_array_equal(r[slots, 1], np.where(slots >= 7, slots // 7, 1000))


# ------------------------------------------------------------------ Q3.1
@pytest.mark.parametrize("version", [1, 2])
def test_q31_reference_kernels(rs, version):
    n = 60_000 + 5
    f, d = T.o_gen_star(41, n, **SMALL)
    f["suppfk"] %= SMALL["n_supp"]
    f["custfk"] %= SMALL["n_cust"]
    want_sums, want_counts = T.o_q31(f, d, n, reg_c=3, reg_s=3, year_lo=1992, year_hi=1997)
    assert want_counts.sum() > 0
    bws = (12, 20, 16)
    keys = [(f[k] + 1).astype(np.uint32) for k in ("datefk", "custfk", "suppfk")]
    d_date, d_cust, d_supp = (i32(fls.pack_column(k, b)) for k, b in zip(keys, bws))
    d_rev = i32(fls.pack_column(f["revenue"], 24))
    ns, nc, nd = SMALL["n_supp"], SMALL["n_cust"], SMALL["n_date"]
    s_key, c_key, d_key = (i32(np.arange(1, m + 1, dtype=np.uint32)) for m in (ns, nc, nd))
    dev = {k: i32(d[k]) for k in ("s_region", "s_nation", "c_region", "c_nation", "d_year")}
    ht_s, ht_c, ht_d = zeros(2 * ns), zeros(2 * nc), zeros(2 * nd)
    res = zeros(4375 * 6)
    rc = rs.rs_q31(version, P(d_date), P(d_cust), P(d_supp), P(d_rev), n, 16, 20, 12, 24, P(dev["s_region"]),
                   P(s_key), P(dev["s_nation"]), ns, 3, P(dev["c_region"]), P(c_key), P(dev["c_nation"]), nc, 3,
                   P(d_key), P(dev["d_year"]), nd, 1, 1992, 1997, P(ht_s), P(ht_c), P(ht_d), P(res))
    assert rc == 0
    sync()
    r = res.cpu().numpy().reshape(4375, 6)
    sums = np.ascontiguousarray(r[:, 4:6]).view(np.int64)[:, 0]
    np.testing.assert_array_equal(sums, want_sums)
    slots = np.nonzero(want_counts > 0)[0]
    np.testing.assert_array_equal(r[slots, 0], 1992 + slots % 7)


# ------------------------------------------------------------------ delta
def test_delta_reference_kernels(rs):
    n_vec, n_tup = 40, 40 * 1024 - 300
    rng = np.random.default_rng(5)
    x = np.sort(rng.integers(0, 1 << 30, n_vec * 1024, dtype=np.uint32))
    enc = [T.o_delta_encode(x[v * 1024:(v + 1) * 1024]) for v in range(n_vec)]
    bw = max(e[2] for e in enc)
    packed = np.concatenate([T.o_pack(T.o_unpack(p, b), bw) for p, _, b in enc])
    bases = np.concatenate([e[1] for e in enc])
    out, total = zeros(n_vec * 1024), zeros(1, "i64")
    assert rs.rs_delta(P(i32(bases)), P(i32(packed)), bw, n_vec, n_tup, P(out), P(total)) == 0
    sync()
    want = T.transpose_vectors(x)
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), want)
    assert int(total.cpu().numpy()[0]) == int(want[:n_tup].astype(np.int64).sum())


# ------------------------------------------------------------------ decode names
@pytest.mark.parametrize("bw", range(33))
def test_per_width_names(rs, bw):
    n_vec = 3
    rng = np.random.default_rng(bw)
    x = rng.integers(0, 1 << bw, n_vec * 1024, dtype=np.uint64).astype(np.uint32) if bw else np.zeros(
        n_vec * 1024, np.uint32)
    d_in = i32(fls.pack_column(x, bw))
    for mode in (0, 1):
        out = zeros(n_vec * 1024)
        assert rs.rs_unpack_named(P(d_in), P(out), bw, n_vec, mode) == 0
        sync()
        np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), x, err_msg=f"mode {mode}")
    out = zeros(n_vec * 1024)
    assert rs.rs_unpack_named(P(d_in), P(out), bw, n_vec, 2) == 0
    sync()
    assert not out.cpu().numpy().any()
    qout = zeros(n_vec * 1024)
    assert rs.rs_unpack_quarters(P(d_in), P(qout), bw, n_vec, 1) == 0
    sync()
    got = qout.cpu().numpy().view(np.uint32)
    if bw % 4 == 0:
        np.testing.assert_array_equal(got, x)
        assert rs.rs_unpack_quarters(P(d_in), P(qout), bw, n_vec, 0) == 0
        sync()
        np.testing.assert_array_equal(qout.cpu().numpy().view(np.uint32), x)
    else:  # no quarter layout at this width: the decode yields zeros
        assert not got.any()


@pytest.mark.parametrize("block_threads", [32, 64, 256])
def test_unpack_global(rs, block_threads):
    n_vec, bw = 16, 11
    x = T.o_gen_c1(3, n_vec * 1024, bw)
    out = zeros(n_vec * 1024)
    assert rs.rs_unpack_global(P(i32(fls.pack_column(x, bw))), P(out), bw, n_vec, block_threads) == 0
    sync()
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), x)


# ------------------------------------------------------------------ primitives
@pytest.mark.parametrize("bt,ipt", [(32, 32), (32, 8), (64, 16), (128, 4), (256, 2)])
def test_primitive_catalogue(rs, bt, ipt):
    tile = bt * ipt
    n_tiles = 9
    n = n_tiles * tile - 37  # a partial last tile
    c = 10
    rng = np.random.default_rng(bt * 100 + ipt)
    x = rng.integers(0, 20, n).astype(np.int32)
    x[tile:2 * tile] = c + 1  # tile 1 holds no item equal to c
    f0 = rng.integers(0, 2, n).astype(np.int32)
    n_out = rs.rs_n_pred_out()
    out_flags = zeros(n_out * n)
    out_vals = zeros(n)
    sums, term = zeros(n_tiles, "i64"), zeros(n_tiles)
    assert rs.rs_prims(bt, ipt, P(i32(x)), P(i32(f0)), n, c, P(out_flags), P(out_vals), P(sums), P(term)) == 0
    sync()
    got = out_flags.cpu().numpy().reshape(n_out, n)
    ops = [x > c, x < c, x >= c, x <= c, x == c]
    want = [o.astype(np.int32) for o in ops] * 2
    want += [(f0.astype(bool) & o).astype(np.int32) for o in ops] * 2
    want.append((f0.astype(bool) | (x == c)).astype(np.int32))
    want.append((((x < c) & (x >= c // 2)) | (x == c)).astype(np.int32))
    assert len(want) == n_out
    for p in range(n_out):
        np.testing.assert_array_equal(got[p], want[p], err_msg=f"predicate output {p}")
    np.testing.assert_array_equal(out_vals.cpu().numpy(), np.where(f0 == 1, x, 0))
    xs = np.concatenate([x, np.zeros(n_tiles * tile - n, np.int32)]).reshape(n_tiles, tile)
    np.testing.assert_array_equal(sums.cpu().numpy(), xs.astype(np.int64).sum(axis=1))
    live = np.arange(n_tiles * tile).reshape(n_tiles, tile) < n
    np.testing.assert_array_equal(term.cpu().numpy(), (~((xs == c) & live).any(axis=1)).astype(np.int32))


def halves(w):
    lo, hi = (w & 0xFFFF).astype(np.uint32), (w >> 16).astype(np.uint32)
    return lo, hi, lo.astype(np.uint16).view(np.int16).astype(np.int32), hi.astype(np.uint16).view(np.int16).astype(
        np.int32)


@pytest.mark.parametrize("ipt", [32, 8])
@pytest.mark.parametrize("partial", [False, True])
def test_simd_16x2_predicates(rs, ipt, partial):
    n_tiles = 5
    half = ipt // 2
    wpt = 32 * half  # words per tile
    n_items = 301 if partial else 32 * ipt
    rng = np.random.default_rng(ipt + partial)
    w = rng.integers(0, 1 << 32, n_tiles * wpt, dtype=np.uint64).astype(np.uint32)
    w[::3] &= 0x003F003F  # small non-negative halves near the constant
    init = rng.integers(0, 1 << 32, w.size, dtype=np.uint64).astype(np.uint32)
    k = 20
    cst = (k << 16) | k
    out = zeros(6 * w.size)
    assert rs.rs_simd16(ipt, P(i32(w)), P(i32(init)), n_tiles, n_items, cst, P(out)) == 0
    sync()
    got = out.cpu().numpy().view(np.uint32).reshape(6, w.size)
    lo, hi, slo, shi = halves(w)

    def mask(a, b):
        return np.where(a, 0xFFFF, 0).astype(np.uint32) | np.where(b, 0xFFFF0000, 0).astype(np.uint32)

    gt = (np.where(slo > k, 1, 0) | np.where(shi > k, 0x10000, 0)).astype(np.uint32)
    lt = mask(slo < k, shi < k)
    le_u, ge_u = mask(lo <= k, hi <= k), mask(lo >= k, hi >= k)
    word_in_tile = np.tile(np.arange(wpt), n_tiles)
    full = n_items == 32 * ipt
    plain = np.ones(w.size, bool) if full else word_in_tile < n_items // 2
    conj = np.ones(w.size, bool) if full else word_in_tile < n_items // 2 + 1
    want = [np.where(plain, gt, init), np.where(plain, lt, init), np.where(conj, init & lt, init),
            np.where(conj, init & le_u, init), np.where(conj, init & ge_u, init),
            np.where(plain, ge_u & le_u, init)]
    for j in range(6):
        np.testing.assert_array_equal(got[j], want[j], err_msg=f"simd output {j}")
