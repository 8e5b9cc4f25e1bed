"""Shared helpers for the tests: the CPU oracle (ctypes), the reference's own
encoders built into oracle/_ref (when present), and device plumbing."""
import ctypes
import pathlib

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libflsref.so"
GOLDEN = ROOT / "tests" / "golden"

U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C")
U64P = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C")
I64P = np.ctypeslib.ndpointer(dtype=np.int64, flags="C")


class OrcRle(ctypes.Structure):
    _fields_ = [("n_runs", ctypes.c_uint32), ("vbase", ctypes.c_uint32), ("vbw", ctypes.c_uint32),
                ("ibw", ctypes.c_uint32), ("bases", ctypes.c_uint32 * 32)]


_oracle = None


def oracle():
    global _oracle
    if _oracle is not None:
        return _oracle
    if not ORACLE_SO.exists():
        raise FileNotFoundError(f"{ORACLE_SO} missing: run `make -C oracle`")
    o = ctypes.CDLL(str(ORACLE_SO))
    u, u64, c_u = ctypes.c_uint, ctypes.c_uint64, ctypes.c_uint32
    o.orc_pack_u32.argtypes = [U32P, U32P, u]
    o.orc_unpack_u32.argtypes = [U32P, U32P, u]
    for n in ("orc_transpose_i_u32", "orc_transpose_o_u32", "orc_unrsum_u32"):
        getattr(o, n).argtypes = [U32P, U32P]
    o.orc_rsum_u32.argtypes = [U32P, U32P, U32P]
    o.orc_rsum_offset.argtypes = [u]
    o.orc_rsum_offset.restype = u
    o.orc_bit_width.argtypes = [c_u]
    o.orc_bit_width.restype = u
    o.orc_find_bw4.argtypes = [U32P, u]
    o.orc_find_bw4.restype = u
    o.orc_delta_encode_u32.argtypes = [U32P, U32P, U32P]
    o.orc_delta_encode_u32.restype = u
    o.orc_delta_decode_u32.argtypes = [U32P, u, U32P, U32P]
    o.orc_rle_encode_u32.argtypes = [U32P, u, ctypes.POINTER(OrcRle), U32P, U32P]
    o.orc_rle_decode_u32.argtypes = [ctypes.POINTER(OrcRle), U32P, U32P, U32P]
    o.orc_linear_extract.argtypes = [U32P, u64, u]
    o.orc_linear_extract.restype = c_u
    o.orc_checksum_u32.argtypes = [U32P, u64]
    o.orc_checksum_u32.restype = u64
    o.orc_mix64.argtypes = [u64]
    o.orc_mix64.restype = u64
    o.orc_rand64.argtypes = [u64, u64, u64]
    o.orc_rand64.restype = u64
    o.orc_uniform.argtypes = [u64, u64, u64, u64]
    o.orc_uniform.restype = c_u
    o.orc_yyyymmdd.argtypes = [c_u]
    o.orc_yyyymmdd.restype = c_u
    o.orc_gen_c1.argtypes = [u64, u64, u, U32P]
    o.orc_gen_c2.argtypes = [u64, u64, U32P, U32P, U32P, U32P]
    o.orc_gen_c3_fact.argtypes = [u64, u64, U32P, U32P, U32P, ctypes.c_void_p, U32P]
    o.orc_gen_c3_dims.argtypes = [u64, c_u, c_u, c_u, c_u] + [U32P] * 7
    o.orc_q11.argtypes = [U32P, U32P, U32P, U32P, u64, c_u, c_u, c_u, c_u, c_u, ctypes.POINTER(u64)]
    o.orc_q11.restype = ctypes.c_int64
    o.orc_q21.argtypes = [U32P, U32P, U32P, U32P, u64, U32P, U32P, U32P, U32P, c_u, c_u, I64P, U64P]
    o.orc_gen_star_extra.argtypes = [u64, u64, U32P, c_u, U32P]
    o.orc_tile_max_words.argtypes = [u64]
    o.orc_tile_max_words.restype = u64
    for name in ("orc_tile_for_encode", "orc_tile_dfor_encode"):
        getattr(o, name).argtypes = [U32P, u64, U32P, U32P]
        getattr(o, name).restype = u64
    for name in ("orc_tile_for_decode", "orc_tile_dfor_decode"):
        getattr(o, name).argtypes = [U32P, U32P, u64, U32P]
    o.orc_tile_rfor_encode.argtypes = [U32P, u64, U32P, U32P, U32P, U32P, ctypes.POINTER(u64), ctypes.POINTER(u64)]
    o.orc_tile_rfor_decode.argtypes = [U32P, U32P, U32P, U32P, u64, U32P]
    o.orc_q31.argtypes = [U32P, U32P, U32P, U32P, u64, U32P, U32P, U32P, U32P, U32P, c_u, c_u, c_u, c_u, I64P, U64P]
    o.orc_q41.argtypes = [U32P, U32P, U32P, U32P, U32P, U32P, u64, U32P, U32P, U32P, U32P, U32P, c_u, c_u, c_u, c_u,
                          I64P, U64P]
    _oracle = o
    return o


def ref():
    """The reference's own encoders (oracle/_ref), or skip when not built."""
    if not REF_SO.exists():
        pytest.skip("oracle/_ref not built (reference checkout absent)")
    r = ctypes.CDLL(str(REF_SO))
    r.ref_pack_u32.argtypes = [U32P, U32P, ctypes.c_uint]
    for n in ("ref_transpose_i_u32", "ref_transpose_o_u32", "ref_unrsum_u32"):
        getattr(r, n).argtypes = [U32P, U32P]
    return r


def golden():
    return np.load(GOLDEN / "fls_layout_golden.npz")


# ----------------------------------------------------------------- oracle wrappers
def o_pack(x, bw):
    out = np.zeros(1024, np.uint32)
    oracle().orc_pack_u32(np.ascontiguousarray(x, np.uint32), out, bw)
    return out[: 32 * bw]


def o_unpack(words, bw):
    buf = np.zeros(1024, np.uint32)
    buf[: 32 * bw] = words[: 32 * bw]
    out = np.zeros(1024, np.uint32)
    oracle().orc_unpack_u32(buf, out, bw)
    return out


def o_perm(name, x):
    out = np.zeros(1024, np.uint32)
    getattr(oracle(), name)(np.ascontiguousarray(x, np.uint32), out)
    return out


def o_delta_encode(x):
    packed = np.zeros(1024, np.uint32)
    bases = np.zeros(32, np.uint32)
    bw = oracle().orc_delta_encode_u32(np.ascontiguousarray(x, np.uint32), packed, bases)
    return packed[: 32 * bw], bases, bw


def o_delta_decode(packed, bw, bases):
    buf = np.zeros(1024, np.uint32)
    buf[: 32 * bw] = packed[: 32 * bw]
    out = np.zeros(1024, np.uint32)
    oracle().orc_delta_decode_u32(buf, bw, np.ascontiguousarray(bases, np.uint32), out)
    return out


def o_gen_c1(seed, n, bw):
    v = np.zeros(n, np.uint32)
    oracle().orc_gen_c1(seed, n, bw, v)
    return v


def o_gen_c2(seed, n):
    cols = [np.zeros(n, np.uint32) for _ in range(4)]
    oracle().orc_gen_c2(seed, n, *cols)
    return cols


def o_checksum(v):
    return int(oracle().orc_checksum_u32(np.ascontiguousarray(v, np.uint32), v.size))


def transpose_vectors(col):
    """Apply the oracle transpose_i to every 1024-value vector of a padded column."""
    n_vec = col.size // 1024
    out = np.empty_like(col)
    for v in range(n_vec):
        out[v * 1024:(v + 1) * 1024] = o_perm("orc_transpose_i_u32", col[v * 1024:(v + 1) * 1024])
    return out


# ----------------------------------------------------------------- device helpers
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def to_dev(a):
    torch = torch_cuda()
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a.copy()).cuda()


def dev_empty(n, dtype="u32"):
    torch = torch_cuda()
    dt = {"u32": torch.int32, "u64": torch.int64, "i64": torch.int64}[dtype]
    return torch.empty(max(1, n), dtype=dt, device="cuda")


def from_dev(t, dtype=np.uint32):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


# ----------------------------------------------------------------- CPU baseline (oracle)
class OrcCol(ctypes.Structure):
    _fields_ = [("encoding", ctypes.c_uint32), ("bw", ctypes.c_uint32), ("base", ctypes.c_uint32),
                ("transposed", ctypes.c_uint32), ("words", ctypes.c_void_p), ("vmeta", ctypes.c_void_p),
                ("bases", ctypes.c_void_p), ("aux", ctypes.c_void_p)]


def orc_col(hc):
    """oracle view (orc_col) of a fastlanesgpu_amd.HostColumn (or any object
    with encoding, bw, base, transposed, words, vmeta, bases, aux numpy
    attributes); keeps arrays alive via the returned tuple."""
    aux = getattr(hc, "aux", None)
    keep = [a for a in (hc.words, hc.vmeta, hc.bases, aux) if a is not None]
    ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
    c = OrcCol(hc.encoding, hc.bw, hc.base, int(hc.transposed), ptr(hc.words), ptr(hc.vmeta), ptr(hc.bases),
               ptr(aux))
    return c, keep


def _baseline_sigs():
    o = oracle()
    if hasattr(o, "_baseline_set"):
        return o
    P = ctypes.POINTER(OrcCol)
    u32, u64, i = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    o.orc_c1_encoded.argtypes = [U32P, u64, ctypes.c_uint, U32P, i]
    o.orc_c1_encoded.restype = u64
    o.orc_q21_encoded.argtypes = [P] * 4 + [u64, u64, u64] + [U32P] * 4 + [u32] * 6 + [i, I64P, U64P]
    o.orc_q31_encoded.argtypes = [P] * 4 + [u64, u64, u64] + [U32P] * 5 + [u32] * 8 + [i, I64P, U64P]
    o.orc_q41_encoded.argtypes = [P] * 6 + [u64, u64, u64] + [U32P] * 5 + [u32] * 9 + [i, I64P, U64P]
    o.orc_c4_encoded.argtypes = [P, U64P, U32P, U32P, u32, i, U64P]
    o.orc_exhaustive_classes.argtypes = [U32P, u64, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C"), i]
    o._baseline_set = True
    return o


def o_c1_encoded(words, n, bw, threads=8):
    out = np.zeros(max(1, n), np.uint32)
    s = _baseline_sigs().orc_c1_encoded(np.ascontiguousarray(words, np.uint32), n, bw, out, threads)
    return out[:n], int(s)


def _vrange(n_rows, v_begin, v_end):
    return v_begin, (n_rows + 1023) // 1024 if v_end is None else v_end


def o_q21_encoded(cols, n_rows, dims, sizes, keys_min=1, cat=12, reg=2, threads=8, v_begin=0, v_end=None):
    cs = [orc_col(c) for c in cols]
    sums, counts = np.zeros(7000, np.int64), np.zeros(7000, np.uint64)
    d = [np.ascontiguousarray(dims[k], np.uint32) for k in ("d_year", "s_region", "p_category", "p_brand")]
    _baseline_sigs().orc_q21_encoded(*(ctypes.byref(c[0]) for c in cs), n_rows, *_vrange(n_rows, v_begin, v_end),
                                     *d, sizes["n_date"], sizes["n_supp"], sizes["n_part"], keys_min, cat, reg,
                                     threads, sums, counts)
    return sums, counts


def o_q31_encoded(cols, n_rows, dims, sizes, keys_min=1, reg_c=3, reg_s=3, year_lo=1992, year_hi=1997, threads=8,
                  v_begin=0, v_end=None):
    """cols: date, customer, supplier, revenue."""
    cs = [orc_col(c) for c in cols]
    sums, counts = np.zeros(4375, np.int64), np.zeros(4375, np.uint64)
    d = [np.ascontiguousarray(dims[k], np.uint32) for k in ("d_year", "s_region", "s_nation", "c_region",
                                                            "c_nation")]
    _baseline_sigs().orc_q31_encoded(*(ctypes.byref(c[0]) for c in cs), n_rows, *_vrange(n_rows, v_begin, v_end),
                                     *d, sizes["n_date"], sizes["n_supp"], sizes["n_cust"], keys_min, reg_c, reg_s,
                                     year_lo, year_hi, threads, sums, counts)
    return sums, counts


def o_q41_encoded(cols, n_rows, dims, sizes, keys_min=1, reg_c=2, reg_s=2, mfgr_lo=1, mfgr_hi=2, threads=8,
                  v_begin=0, v_end=None):
    """cols: date, customer, supplier, part, revenue, supplycost."""
    cs = [orc_col(c) for c in cols]
    sums, counts = np.zeros(175, np.int64), np.zeros(175, np.uint64)
    d = [np.ascontiguousarray(dims[k], np.uint32) for k in ("d_year", "s_region", "c_region", "c_nation", "p_mfgr")]
    _baseline_sigs().orc_q41_encoded(*(ctypes.byref(c[0]) for c in cs), n_rows, *_vrange(n_rows, v_begin, v_end),
                                     *d, sizes["n_date"], sizes["n_supp"], sizes["n_cust"], sizes["n_part"],
                                     keys_min, reg_c, reg_s, mfgr_lo, mfgr_hi, threads, sums, counts)
    return sums, counts


def o_c4_encoded(segs, threads=8):
    """segs: list of (column-like with numpy arrays, n_values, lo, hi)."""
    n = len(segs)
    arr = (OrcCol * max(1, n))()
    keep = []
    for k, (c, _, _, _) in enumerate(segs):
        oc, kp = orc_col(c)
        arr[k] = oc
        keep.append(kp)
    nv = np.array([s[1] for s in segs], np.uint64)
    lo = np.array([s[2] for s in segs], np.uint32)
    hi = np.array([s[3] for s in segs], np.uint32)
    counts = np.zeros(max(1, n), np.uint64)
    _baseline_sigs().orc_c4_encoded(arr, nv, lo, hi, n, threads, counts)
    return counts[:n]


def o_exhaustive_classes(vectors, threads=8):
    v = np.ascontiguousarray(vectors, np.uint32)
    n_vec = v.size // 1024
    cls = np.zeros(max(1, n_vec), np.uint8)
    _baseline_sigs().orc_exhaustive_classes(v, n_vec, cls, threads)
    return cls[:n_vec]


def o_q11_encoded(cols, n_rows, v_begin=0, v_end=None, threads=8, params=(19930000, 19940000, 25, 1, 3)):
    o = oracle()
    if not hasattr(o.orc_q11_encoded, "_set"):
        P = ctypes.POINTER(OrcCol)
        o.orc_q11_encoded.argtypes = [P, P, P, P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64] + \
            [ctypes.c_uint32] * 5 + [ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)]
        o.orc_q11_encoded.restype = ctypes.c_int64
        o.orc_q11_encoded._set = True
    if v_end is None:
        v_end = (n_rows + 1023) // 1024
    cs = [orc_col(c) for c in cols]
    cnt = ctypes.c_uint64(0)
    s = o.orc_q11_encoded(*(ctypes.byref(c[0]) for c in cs), n_rows, v_begin, v_end, *params, threads,
                          ctypes.byref(cnt))
    return int(s), int(cnt.value)


def o_q11(cols, n, params=(19930000, 19940000, 25, 1, 3)):
    c = ctypes.c_uint64(0)
    s = oracle().orc_q11(*[np.ascontiguousarray(x, np.uint32) for x in cols], n, *params, ctypes.byref(c))
    return int(s), int(c.value)


def o_gen_c3(seed, n, n_date=2556, n_supp=40000, n_part=800000, n_cust=600000):
    o = oracle()
    fact = [np.zeros(n, np.uint32) for _ in range(4)]
    o.orc_gen_c3_fact(seed, n, fact[0], fact[1], fact[2], None, fact[3])
    dims = [np.zeros(max(1, k), np.uint32) for k in (n_date, n_supp, n_supp, n_part, n_part, n_cust, n_cust)]
    o.orc_gen_c3_dims(seed, n_date, n_supp, n_part, n_cust, *dims)
    names = ["d_year", "s_region", "s_nation", "p_category", "p_brand", "c_region", "c_nation"]
    return fact, dict(zip(names, dims))


def o_q21(fact, dims, n, cat=12, reg=2):
    sums = np.zeros(7000, np.int64)
    counts = np.zeros(7000, np.uint64)
    datefk, suppfk, partfk, rev = [np.ascontiguousarray(x, np.uint32) for x in fact]
    oracle().orc_q21(datefk, suppfk, partfk, rev, n, dims["d_year"], dims["s_region"], dims["p_category"],
                     dims["p_brand"], cat, reg, sums, counts)
    return sums, counts


STAR_NAMES = ("datefk", "custfk", "suppfk", "partfk", "revenue", "supplycost")


def o_gen_star(seed, n, n_date=2556, n_supp=40000, n_part=800000, n_cust=600000):
    """Star-join fact columns (0-based keys) and dimensions, as the GPU
    generators write them: fact dict over STAR_NAMES, dims dict with p_mfgr."""
    o = oracle()
    f = {k: np.zeros(max(1, n), np.uint32) for k in STAR_NAMES}
    o.orc_gen_c3_fact(seed, n, f["datefk"], f["suppfk"], f["partfk"], f["custfk"].ctypes.data, f["revenue"])
    dims = [np.zeros(max(1, k), np.uint32) for k in (n_date, n_supp, n_supp, n_part, n_part, n_cust, n_cust)]
    o.orc_gen_c3_dims(seed, n_date, n_supp, n_part, n_cust, *dims)
    names = ["d_year", "s_region", "s_nation", "p_category", "p_brand", "c_region", "c_nation"]
    d = dict(zip(names, dims))
    d["p_mfgr"] = np.zeros(max(1, n_part), np.uint32)
    o.orc_gen_star_extra(seed, n, f["supplycost"], n_part, d["p_mfgr"])
    return {k: v[:n] for k, v in f.items()}, d


def o_q31(f, d, n, reg_c=3, reg_s=3, year_lo=1992, year_hi=1997):
    sums = np.zeros(4375, np.int64)
    counts = np.zeros(4375, np.uint64)
    c = [np.ascontiguousarray(f[k], np.uint32) for k in ("datefk", "custfk", "suppfk", "revenue")]
    oracle().orc_q31(*c, n, d["d_year"], d["s_region"], d["s_nation"], d["c_region"], d["c_nation"], reg_c, reg_s,
                     year_lo, year_hi, sums, counts)
    return sums, counts


def o_q41(f, d, n, reg_c=2, reg_s=2, mfgr_lo=1, mfgr_hi=2):
    sums = np.zeros(175, np.int64)
    counts = np.zeros(175, np.uint64)
    c = [np.ascontiguousarray(f[k], np.uint32) for k in ("datefk", "custfk", "suppfk", "partfk", "revenue",
                                                          "supplycost")]
    oracle().orc_q41(*c, n, d["d_year"], d["s_region"], d["c_region"], d["c_nation"], d["p_mfgr"], reg_c, reg_s,
                     mfgr_lo, mfgr_hi, sums, counts)
    return sums, counts


def o_c4_counts(seed, n_seg, lmin, lmax, threads=16):
    o = oracle()
    if not hasattr(o.orc_c4_counts, "_set"):
        o.orc_c4_counts.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, U64P, U32P,
                                    U32P, ctypes.c_int]
        o.orc_c4_values.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                    ctypes.c_uint64, U32P]
        o.orc_c4_params.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32] + \
            [ctypes.POINTER(ctypes.c_uint32)] * 2 + [ctypes.POINTER(ctypes.c_uint64)]
        o.orc_c4_counts._set = True
    counts = np.zeros(n_seg, np.uint64)
    lo = np.zeros(n_seg, np.uint32)
    hi = np.zeros(n_seg, np.uint32)
    o.orc_c4_counts(seed, n_seg, lmin, lmax, counts, lo, hi, threads)
    return counts, lo, hi


def o_c4_values(seed, s, lmin, lmax):
    o_c4_counts(seed, 1, lmin, lmax, 1)  # sets argtypes
    o = oracle()
    enc, bw, n = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    o.orc_c4_params(seed, s, lmin, lmax, ctypes.byref(enc), ctypes.byref(bw), ctypes.byref(n))
    out = np.zeros(n.value, np.uint32)
    o.orc_c4_values(seed, s, enc.value, bw.value, n.value, out)
    return enc.value, bw.value, out


# ----------------------------------------------------------------- selector oracle
F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C")
F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C")
U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C")


def _sel():
    o = oracle()
    if not hasattr(o, "_sel_set"):
        u, u64 = ctypes.c_uint, ctypes.c_uint64
        o.orc_width_bucket.argtypes = [u]
        o.orc_width_bucket.restype = u
        o.orc_mix_encode_u32.argtypes = [U32P, u, U32P, ctypes.POINTER(u), ctypes.POINTER(u64)]
        o.orc_mix_encode_u32.restype = u
        o.orc_mix_sizes.argtypes = [U32P, U32P, U32P]
        o.orc_exhaustive_class.argtypes = [U32P]
        o.orc_exhaustive_class.restype = u
        o.orc_vector_features.argtypes = [U32P, F32P]
        o.orc_gen_train_vector.argtypes = [u64, u64, U32P, U32P]
        o.orc_mlp_param_count.restype = u
        o.orc_bf16.argtypes = [ctypes.c_double]
        o.orc_bf16.restype = ctypes.c_double
        o.orc_mlp_forward.argtypes = [F64P, F32P, u64, ctypes.c_int, F64P]
        o.orc_mlp_grad.argtypes = [F64P, F32P, U8P, u64, ctypes.c_double, ctypes.c_int, F64P]
        o.orc_mlp_grad.restype = ctypes.c_double
        o._sel_set = True
    return o


def o_mix_encode(vec, enc):
    """(payload words, width, aux) of one 1024-value vector."""
    words = np.zeros(1088, np.uint32)
    w, aux = ctypes.c_uint(), ctypes.c_uint64()
    n = _sel().orc_mix_encode_u32(np.ascontiguousarray(vec, np.uint32), enc, words, ctypes.byref(w),
                                  ctypes.byref(aux))
    return words[:n].copy(), w.value, aux.value


def o_mix_sizes(vec):
    s, w = np.zeros(4, np.uint32), np.zeros(4, np.uint32)
    _sel().orc_mix_sizes(np.ascontiguousarray(vec, np.uint32), s, w)
    return s, w


def o_exhaustive_class(vec):
    return int(_sel().orc_exhaustive_class(np.ascontiguousarray(vec, np.uint32)))


def o_features(vec):
    f = np.zeros(16, np.float32)
    _sel().orc_vector_features(np.ascontiguousarray(vec, np.uint32), f)
    return f


def o_train_vector(seed, v, cdf):
    out = np.zeros(1024, np.uint32)
    _sel().orc_gen_train_vector(seed, v, np.ascontiguousarray(cdf, np.uint32), out)
    return out


def o_mlp_forward(w, feat, round_bf16):
    feat = np.ascontiguousarray(feat, np.float32)
    n = feat.shape[0]
    out = np.zeros((n, 40), np.float64)
    _sel().orc_mlp_forward(np.ascontiguousarray(w, np.float64), feat, n, int(round_bf16), out)
    return out


def o_mlp_grad(w, feat, label, round_bf16, inv_batch=None):
    feat = np.ascontiguousarray(feat, np.float32)
    n = feat.shape[0]
    g = np.zeros(int(_sel().orc_mlp_param_count()), np.float64)
    loss = _sel().orc_mlp_grad(np.ascontiguousarray(w, np.float64), feat, np.ascontiguousarray(label, np.uint8), n,
                               1.0 / n if inv_batch is None else inv_batch, int(round_bf16), g)
    return loss, g


def pad_vectors(a):
    """Pad a column to whole vectors by repeating its last value (the GPU
    encoders' convention for a partial last vector)."""
    a = np.ascontiguousarray(a, np.uint32)
    nv = (a.size + 1023) // 1024
    out = np.empty(nv * 1024, np.uint32)
    out[:a.size] = a
    out[a.size:] = a[-1] if a.size else 0
    return out


def o_uniform_py(seed, stream, i, rng_range):
    return int(oracle().orc_uniform(seed, stream, i, rng_range))


def encode_any(x, kind):
    """Device column of the uint32 values x in one of the formats the probe
    kernels decode at run time (all in transposed row order):
      "delta"      FastLanes delta (GPU encoder)
      "rle"        FastLanes RLE (GPU encoder)
      "for"        MIXED with every vector FOR: a width, base and offset per
                   vector, the layout of the per-vector FOR key columns
      "exhaustive" MIXED with each vector's smallest encoding
      "bitpack_t"  BITPACK at the smallest width, transposed"""
    import fastlanesgpu_amd as fls
    from fastlanesgpu_amd import selector as sel

    x = np.ascontiguousarray(x, np.uint32)
    n = x.size
    d = to_dev(x)
    if kind == "delta":
        w, vm, bs = fls.delta_encode_device(d, n)
        return fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)
    if kind == "rle":
        w, vm, rm, bs = fls.rle_encode_device(d, n)
        return fls.Column(fls.ENC_RLE, n, w, vmeta=vm, bases=bs, aux=rm)
    if kind in ("for", "exhaustive"):
        ws = sel.MixedWorkspace(n)
        if kind == "for":
            ws.cls.fill_(sel.MIX_FOR * 10)
            sel.mixed_encode_classes(d, n, ws)
        else:
            sel.mixed_encode_exhaustive(d, n, ws)
        return ws.column(n)
    if kind == "bitpack_t":
        return fls.bitpack_column(x, transposed=True).to_device()
    raise ValueError(kind)


# ----------------------------------------------------- tile-based formats (oracle)
def o_tile_encode(kind, x):
    """GPU-FOR ("for") / GPU-DFOR ("dfor") words and block starts of x."""
    o = oracle()
    x = np.ascontiguousarray(x, np.uint32)
    n = x.size
    out = np.zeros(int(o.orc_tile_max_words(n)), np.uint32)
    starts = np.zeros((n + 127) // 128 + 8, np.uint32)
    nw = getattr(o, f"orc_tile_{kind}_encode")(x, n, out, starts)
    nb = (n + 127) // 128 if kind == "for" else 4 * ((n + 511) // 512)
    return out[:nw].copy(), starts[:nb + 1].copy()


def o_tile_decode(kind, words, starts, n):
    out = np.zeros(max(1, n), np.uint32)
    getattr(oracle(), f"orc_tile_{kind}_decode")(np.ascontiguousarray(words, np.uint32),
                                                  np.ascontiguousarray(starts, np.uint32), n, out)
    return out[:n]


def o_rfor_encode(x):
    """GPU-RFOR: (values, val_starts, run_lengths, rl_starts)."""
    o = oracle()
    x = np.ascontiguousarray(x, np.uint32)
    n = x.size
    cap = int(o.orc_tile_max_words(n))
    vals, rls = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    nt = (n + 511) // 512
    vs, rs = np.zeros(nt + 1, np.uint32), np.zeros(nt + 1, np.uint32)
    nv, nr = ctypes.c_uint64(0), ctypes.c_uint64(0)
    o.orc_tile_rfor_encode(x, n, vals, vs, rls, rs, ctypes.byref(nv), ctypes.byref(nr))
    return vals[:nv.value].copy(), vs, rls[:nr.value].copy(), rs


def o_rfor_decode(vals, vs, rls, rs, n):
    out = np.zeros(max(1, n), np.uint32)
    oracle().orc_tile_rfor_decode(vals, vs, rls, rs, n, out)
    return out[:n]
