"""The oracle's CPU baseline (decode + query over encoded columns) is pinned
against the bit-by-bit oracle functions before it is used as a baseline."""
import numpy as np

import fastlanesgpu_amd as fls
import fls_testlib as T

RNG = np.random.default_rng(5)


def test_lanewise_unpack_matches_bitwise():
    o = T.oracle()
    o.orc_unpack_lanewise_u32.argtypes = [T.U32P, T.U32P, T.ctypes.c_uint]
    for bw in range(33):
        x = RNG.integers(0, 2**bw, 1024, dtype=np.uint64).astype(np.uint32)
        packed = np.zeros(1024, np.uint32)
        packed[: 32 * bw] = T.o_pack(x, bw)
        out = np.zeros(1024, np.uint32)
        o.orc_unpack_lanewise_u32(packed, out, bw)
        np.testing.assert_array_equal(out, x)


def test_q11_encoded_matches_raw_evaluation():
    for n in (1, 700, 1024, 20 * 1024 + 333):
        cols = T.o_gen_c2(13, n)
        enc = fls.encode_c2_table(*cols)
        assert T.o_q11_encoded(enc, n, threads=4) == T.o_q11(cols, n)
    # a selective window, so the sum is non-trivial at small n
    n = 5000
    cols = T.o_gen_c2(13, n)
    enc = fls.encode_c2_table(*cols)
    params = (0, 2**31, 40, 0, 10)
    got = T.o_q11_encoded(enc, n, threads=2, params=params)
    assert got == T.o_q11(cols, n, params) and got[1] > 0


# ------------------------------------------------ the other workloads' baselines
class Col:
    """Column-like holder of numpy arrays for the oracle's orc_col view."""

    def __init__(self, encoding, bw, words, vmeta=None, bases=None, aux=None, transposed=False, base=0):
        self.encoding, self.bw, self.words, self.vmeta, self.bases, self.aux = encoding, bw, words, vmeta, bases, aux
        self.transposed, self.base = transposed, base


def rle_column(x):
    """FastLanes RLE column (the library's vector format) assembled from the
    oracle's per-vector RLE encoder."""
    o = T.oracle()
    words, vmeta, bases, aux, off = [], [], [], [], 0
    for v in range((x.size + 1023) // 1024):
        vec = np.ascontiguousarray(x[v * 1024:(v + 1) * 1024])
        hdr = T.OrcRle()
        vals, idx = np.zeros(1024, np.uint32), np.zeros(1024, np.uint32)
        o.orc_rle_encode_u32(vec, vec.size, T.ctypes.byref(hdr), vals, idx)
        payload = np.concatenate([idx[:32 * hdr.ibw], vals[:(hdr.n_runs * hdr.vbw + 31) // 32]])
        vmeta.append((off << 8) | hdr.ibw)
        aux.append(hdr.vbase | (hdr.n_runs << 32) | (hdr.vbw << 48) | (hdr.ibw << 56))
        bases.append(np.array(hdr.bases, np.uint32))
        words.append(payload)
        off += payload.size
    return Col(2, 0, np.concatenate(words + [np.zeros(1, np.uint32)]), np.array(vmeta, np.uint64),
               np.concatenate(bases), np.array(aux, np.uint64), transposed=True)


def test_c1_encoded_matches_values_and_checksum():
    for n, bw in ((1, 3), (5000, 11), (3 * 1024, 32), (2048, 0)):
        x = T.o_gen_c1(3, n, bw)
        out, s = T.o_c1_encoded(fls.pack_column(x, bw), n, bw, threads=4)
        np.testing.assert_array_equal(out, x)
        assert s == T.o_checksum(x)


def test_rle_vector_decode_matches_transposed_values():
    lens = RNG.geometric(1 / 8, 5000)
    x = np.repeat(RNG.integers(0, 1 << 20, 5000), lens)[:3 * 1024].astype(np.uint32)
    c = rle_column(x)
    o = T.oracle()
    o.orc_decode_vector.argtypes = [T.ctypes.POINTER(T.OrcCol), T.ctypes.c_uint64, T.U32P]
    oc, keep = T.orc_col(c)
    for v in range(3):
        out = np.zeros(1024, np.uint32)
        o.orc_decode_vector(T.ctypes.byref(oc), v, out)
        np.testing.assert_array_equal(out, T.o_perm("orc_transpose_i_u32", x[v * 1024:(v + 1) * 1024]))
    del keep


def test_c4_encoded_counts_match_numpy():
    segs, want = [], []
    for k, n in enumerate((1, 1000, 4096, 7000 + 3)):
        x = RNG.integers(0, 1 << 12, n).astype(np.uint32)
        lo, hi = 1000, 3000
        segs.append((Col(0, 12, fls.pack_column(x, 12)), n, lo, hi))
        want.append(int(((x >= lo) & (x <= hi)).sum()))
        s = np.sort(RNG.integers(0, 1 << 30, n)).astype(np.uint32)
        d = fls.delta_column(s)
        lo, hi = int(s[n // 4]), int(s[3 * n // 4])
        segs.append((Col(1, 0, d.words, d.vmeta, d.bases, transposed=True), n, lo, hi))
        want.append(int(((s >= lo) & (s <= hi)).sum()))
        r = np.repeat(RNG.integers(0, 1 << 16, n), RNG.geometric(1 / 8, n))[:n].astype(np.uint32)
        segs.append((rle_column(r), n, 1 << 14, 1 << 15))
        want.append(int(((r >= 1 << 14) & (r <= 1 << 15)).sum()))
    np.testing.assert_array_equal(T.o_c4_encoded(segs, threads=3), np.array(want, np.uint64))


def test_q21_encoded_matches_plain():
    sizes = dict(n_date=2556, n_supp=400, n_part=8000, n_cust=10)
    for n in (700, 20 * 1024 + 5):
        fact, dims = T.o_gen_c3(9, n, **sizes)
        fact[1] %= sizes["n_supp"]
        fact[2] %= sizes["n_part"]
        enc = fls.encode_c3_table(*fact, n_date=2556, n_supp=400, n_part=8000)
        got = T.o_q21_encoded(enc, n, dims, sizes, threads=3)
        want = T.o_q21(fact, dims, n)
        np.testing.assert_array_equal(got[0], want[0])
        np.testing.assert_array_equal(got[1], want[1])
        assert want[1].sum() > 0


def test_star_encoded_match_plain():
    sizes = dict(n_date=2556, n_supp=400, n_part=8000, n_cust=6000)
    n = 30 * 1024 + 17
    f, d = T.o_gen_star(8, n, **sizes)
    f["suppfk"] %= sizes["n_supp"]
    f["custfk"] %= sizes["n_cust"]
    f["partfk"] %= sizes["n_part"]
    enc = fls.encode_star_table(*(f[k] for k in T.STAR_NAMES), n_date=2556, n_cust=6000, n_supp=400, n_part=8000)
    got31 = T.o_q31_encoded([enc[0], enc[1], enc[2], enc[4]], n, d, sizes, threads=3)
    want31 = T.o_q31(f, d, n)
    np.testing.assert_array_equal(got31[0], want31[0])
    np.testing.assert_array_equal(got31[1], want31[1])
    assert want31[1].sum() > 0
    got41 = T.o_q41_encoded(enc, n, d, sizes, threads=3)
    want41 = T.o_q41(f, d, n)
    np.testing.assert_array_equal(got41[0], want41[0])
    np.testing.assert_array_equal(got41[1], want41[1])
    assert want41[1].sum() > 0


def test_exhaustive_classes_match_per_vector():
    cdf = np.linspace(1, 2**32 - 1, 4096).astype(np.uint32)
    vecs = np.concatenate([T.o_train_vector(4, v, cdf) for v in range(40)])
    cls = T.o_exhaustive_classes(vecs, threads=3)
    assert [int(c) for c in cls] == [T.o_exhaustive_class(vecs[v * 1024:(v + 1) * 1024]) for v in range(40)]
