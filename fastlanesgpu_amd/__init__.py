"""MI355X (gfx950) FastLanes decode + Crystal query library: Python host API.

A thin layer over the C ABI of libfls.so (include/fls/fls.h).  Names follow the
reference's host interface: `pack`, `transpose_i`, `transpose_o`, `unrsum`
(fastlanes/src/include/fls_gen/{pack,transpose,unrsum}/*.hpp) and the grid-level
`unpack_global` (fastlanes/src/include/fls_gen/unpack/unpack.cuh:3436).

Host arrays are numpy uint32; device buffers are torch CUDA tensors (any
4-byte dtype) or raw device addresses (int).  Device calls run on torch's
current stream unless `stream` is given.
"""
import ctypes

import numpy as np

from ._lib import FlsError, LIB_PATH, call, load  # noqa: F401

VECTOR_SIZE = 1024
LANES = 32
ENC_FLAG_TRANSPOSE = 1


def _hptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _u32(a):
    a = np.ascontiguousarray(a)
    if a.dtype != np.uint32:
        a = a.astype(np.uint32)
    return a


def dptr(x):
    """Device address of a torch tensor (or pass an int through)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def current_stream():
    try:
        import torch

        if torch.cuda.is_available():
            return torch.cuda.current_stream().cuda_stream
    except ImportError:  # pragma: no cover
        pass
    return None


def _stream(s):
    return current_stream() if s is None else s


# ----------------------------------------------------------- host encoders
def pack(values, bw):
    """generated::pack::fallback::scalar::pack for one 1024-value vector."""
    v = _u32(values)
    assert v.size == VECTOR_SIZE
    out = np.zeros(VECTOR_SIZE, np.uint32)
    call("fls_pack_u32", _hptr(v), _hptr(out), bw)
    return out[: 32 * bw].copy()


def transpose_i(values):
    v = _u32(values)
    out = np.empty(VECTOR_SIZE, np.uint32)
    call("fls_transpose_i_u32", _hptr(v), _hptr(out))
    return out


def transpose_o(values):
    v = _u32(values)
    out = np.empty(VECTOR_SIZE, np.uint32)
    call("fls_transpose_o_u32", _hptr(v), _hptr(out))
    return out


def unrsum(values):
    v = _u32(values)
    out = np.empty(VECTOR_SIZE, np.uint32)
    call("fls_unrsum_u32", _hptr(v), _hptr(out))
    return out


def packed_words(n, bw):
    return int(load().fls_packed_words(n, bw))


def pack_column(values, bw, transpose=False):
    v = _u32(values)
    out = np.zeros(max(1, packed_words(v.size, bw)), np.uint32)
    call("fls_pack_column_u32", _hptr(v), v.size, bw, ENC_FLAG_TRANSPOSE if transpose else 0, _hptr(out))
    return out[: packed_words(v.size, bw)]


def delta_encode_column(values):
    """FastLanes delta of a column -> (words, vmeta, bases)."""
    v = _u32(values)
    n_vec = (v.size + VECTOR_SIZE - 1) // VECTOR_SIZE
    vmeta = np.zeros(max(1, n_vec), np.uint64)
    bases = np.zeros(max(1, 32 * n_vec), np.uint32)
    used = ctypes.c_uint64(0)
    lib = load()
    # first call sizes the output (FLS_ERR_CAPACITY with out_cap 0), second fills it
    st = lib.fls_delta_encode_column_u32(_hptr(v), v.size, None, 0, _hptr(vmeta), _hptr(bases), ctypes.byref(used))
    if st not in (0, -7):
        raise FlsError("fls_delta_encode_column_u32", st, lib)
    words = np.zeros(max(1, used.value), np.uint32)
    call("fls_delta_encode_column_u32", _hptr(v), v.size, _hptr(words), words.size, _hptr(vmeta), _hptr(bases),
         ctypes.byref(used))
    return words[: used.value], vmeta[:n_vec], bases[: 32 * n_vec]


# ----------------------------------------------------------- device decode
def unpack_global(d_in, d_out, n_vec, bw, stream=None):
    """Grid-level decode of n_vec fixed-width vectors (normal order)."""
    call("fls_unpack_u32", dptr(d_in), dptr(d_out), n_vec, bw, _stream(stream))


CHECKSUM_WORDS = 1056  # FLS_CHECKSUM_WORDS: the result and the in-launch reduction's state


def checksum_buffer():
    """A zeroed device buffer for unpack_checksum (the result lands in word 0;
    the rest is the reduction's state, which every call leaves zero)."""
    import torch

    return torch.zeros(CHECKSUM_WORDS, dtype=torch.int64, device="cuda")


def unpack_checksum(d_in, d_out, n_values, bw, d_checksum, stream=None):
    """Decode + checksum into d_checksum[0]; d_checksum is a checksum_buffer()
    (CHECKSUM_WORDS 64-bit words, state zero before the first call)."""
    if not isinstance(d_checksum, int) and d_checksum.numel() * d_checksum.element_size() < 8 * CHECKSUM_WORDS:
        raise ValueError(f"d_checksum must hold {CHECKSUM_WORDS} 64-bit words (use checksum_buffer())")
    call("fls_unpack_u32_checksum", dptr(d_in), dptr(d_out), n_values, bw, dptr(d_checksum), _stream(stream))


def delta_decode(d_words, d_vmeta, d_bases, n_vec, d_out, stream=None):
    call("fls_delta_decode_u32", dptr(d_words), dptr(d_vmeta), dptr(d_bases), n_vec, dptr(d_out), _stream(stream))


# ----------------------------------------------------------- columns
ENC_BITPACK, ENC_DELTA, ENC_RLE, ENC_MIXED = 0, 1, 2, 3
Q_PRED_LOAD = 1
Q_ACCUMULATE = 2

from ._lib import FlsColumn, Q11Params  # noqa: E402


class Column:
    """A compressed column resident on the device (torch tensors own the memory)."""

    def __init__(self, encoding, n_rows, words, bw=0, base=0, transposed=False, vmeta=None, bases=None,
                 n_words=None, aux=None):
        self.encoding, self.n_rows, self.bw, self.base = encoding, n_rows, bw, base
        self.transposed = bool(transposed) or encoding in (ENC_DELTA, ENC_RLE, ENC_MIXED)
        self.words, self.vmeta, self.bases, self.aux = words, vmeta, bases, aux
        self.n_words = int(n_words if n_words is not None else (words.numel() if words is not None else 0))

    def struct(self):
        return FlsColumn(self.encoding, self.bw, self.base, int(self.transposed), self.n_words,
                         dptr(self.words), dptr(self.vmeta), dptr(self.bases), dptr(self.aux))

    def nbytes(self):
        s = self.struct()
        return int(load().fls_column_bytes(ctypes.byref(s), self.n_rows))

    def to_host(self):
        """Copy the column's device arrays to a HostColumn (for serialization)."""
        def host(t, dt, count=None):
            if t is None:
                return None
            a = t.detach().cpu().numpy().view(dt)
            return a[:count].copy() if count is not None else a.copy()

        nv = (self.n_rows + VECTOR_SIZE - 1) // VECTOR_SIZE
        meta = self.encoding != ENC_BITPACK
        return HostColumn(self.encoding, self.n_rows, host(self.words, np.uint32, self.n_words), bw=self.bw,
                          base=self.base, transposed=self.transposed,
                          vmeta=host(self.vmeta, np.uint64, nv) if meta else None,
                          bases=host(self.bases, np.uint32, 32 * nv) if self.encoding in (ENC_DELTA, ENC_RLE) else None,
                          aux=host(self.aux, np.uint64, nv) if self.encoding in (ENC_RLE, ENC_MIXED) else None)


def _torch_dev(a):
    import torch

    a = np.ascontiguousarray(a)
    view = {np.dtype(np.uint32): np.int32, np.dtype(np.uint64): np.int64}.get(a.dtype)
    if view is not None:
        a = a.view(view)
    if a.size == 0:
        a = np.zeros(1, a.dtype)
    return torch.from_numpy(a.copy()).cuda()


class HostColumn:
    """A compressed column in host memory (numpy arrays), as the encoders emit it."""

    def __init__(self, encoding, n_rows, words, bw=0, base=0, transposed=False, vmeta=None, bases=None, aux=None):
        self.encoding, self.n_rows, self.bw, self.base = encoding, n_rows, bw, base
        self.transposed = bool(transposed) or encoding in (ENC_DELTA, ENC_RLE, ENC_MIXED)
        self.words, self.vmeta, self.bases, self.aux = words, vmeta, bases, aux

    def to_device(self):
        opt = lambda a: None if a is None else _torch_dev(a)  # noqa: E731
        return Column(self.encoding, self.n_rows, _torch_dev(self.words), bw=self.bw, base=self.base,
                      transposed=self.transposed, vmeta=opt(self.vmeta), bases=opt(self.bases),
                      n_words=self.words.size, aux=opt(self.aux))

    def struct(self):
        """fls_column over the host arrays (valid while this object lives)."""
        ptr = lambda a: None if a is None else a.ctypes.data  # noqa: E731
        return FlsColumn(self.encoding, self.bw, self.base, int(self.transposed), self.words.size,
                         ptr(self.words), ptr(self.vmeta), ptr(self.bases), ptr(self.aux))

    def to_bytes(self):
        """The serialized column image (fls_column_serialize): header + sections."""
        s = self.struct()
        n = int(load().fls_column_image_bytes(ctypes.byref(s), self.n_rows))
        if n == 0:
            raise FlsError("fls_column_image_bytes", -5, load())
        out = np.zeros(n, np.uint8)
        call("fls_column_serialize", ctypes.byref(s), self.n_rows, out.ctypes.data, n)
        return out

    @classmethod
    def from_bytes(cls, image):
        """Parse and validate an image (fls_column_deserialize); the arrays are
        copied out of it."""
        img = np.ascontiguousarray(np.frombuffer(image, np.uint8) if not isinstance(image, np.ndarray) else image,
                                   np.uint8)
        c = FlsColumn()
        rows = ctypes.c_uint64()
        call("fls_column_deserialize", img.ctypes.data, img.size, ctypes.byref(c), ctypes.byref(rows))
        n = rows.value
        nv = (n + VECTOR_SIZE - 1) // VECTOR_SIZE

        def take(ptr, count, dt):
            if not ptr or count == 0:
                return None
            off = ptr - img.ctypes.data
            return img[off:off + count * np.dtype(dt).itemsize].view(dt).copy()

        words = take(c.words, c.n_words, np.uint32)
        words = words if words is not None else np.zeros(0, np.uint32)
        has_meta = c.encoding != ENC_BITPACK
        return cls(c.encoding, n, words, bw=c.bw, base=c.base, transposed=bool(c.transposed),
                   vmeta=take(c.vmeta, nv if has_meta else 0, np.uint64),
                   bases=take(c.bases, 32 * nv if c.encoding in (ENC_DELTA, ENC_RLE) else 0, np.uint32),
                   aux=take(c.aux, nv if c.encoding in (ENC_RLE, ENC_MIXED) else 0, np.uint64))

    def save(self, path):
        self.to_bytes().tofile(path)

    @classmethod
    def load(cls, path):
        return cls.from_bytes(np.fromfile(path, np.uint8))

    def nbytes(self):
        s = FlsColumn(self.encoding, self.bw, self.base, int(self.transposed), self.words.size)
        return int(load().fls_column_bytes(ctypes.byref(s), self.n_rows))


def bitpack_column(values, bw=None, transposed=False, base=0):
    """Host-encode a fixed-width (FOR) column; bw defaults to the smallest width."""
    v = _u32(values)
    if base:
        v = (v - np.uint32(base)).astype(np.uint32)
    if bw is None:
        bw = int(v.max()).bit_length() if v.size else 0
    return HostColumn(ENC_BITPACK, v.size, pack_column(v, bw, transpose=transposed), bw=bw, base=base,
                      transposed=transposed)


def delta_column(values):
    """Host-encode a FastLanes delta column."""
    v = _u32(values)
    words, vmeta, bases = delta_encode_column(v)
    return HostColumn(ENC_DELTA, v.size, words, vmeta=vmeta, bases=bases)


def bitpack_column_to_device(values, bw=None, transposed=False, base=0):
    return bitpack_column(values, bw, transposed, base).to_device()


def delta_column_to_device(values):
    return delta_column(values).to_device()


def query_q11(orderdate, quantity, discount, price, n_rows, date_gt=19930000, date_lt=19940000, qty_lt=25,
              disc_lo=1, disc_hi=3, pred_load=False, accumulate=False, d_out=None, stream=None, sync=True):
    """Fused decode + filter + SUM(price*discount) (fls_q11.cu QueryKernel_v1 shape).

    Returns (revenue, count) as Python ints when sync=True; otherwise the
    2-element int64 device tensor [revenue, count] that the kernel fills
    (adds into, with accumulate=True)."""
    import torch

    if d_out is None:
        d_out = torch.zeros(2, dtype=torch.int64, device="cuda")
    cols = [c.struct() for c in (orderdate, quantity, discount, price)]
    p = Q11Params(date_gt, date_lt, qty_lt, disc_lo, disc_hi)
    base = d_out.data_ptr()
    flags = (Q_PRED_LOAD if pred_load else 0) | (Q_ACCUMULATE if accumulate else 0)
    call("fls_query_q11", *(ctypes.byref(c) for c in cols), n_rows, ctypes.byref(p), flags, base, base + 8,
         _stream(stream))
    if not sync:
        return d_out
    h = d_out.cpu().numpy()
    return int(h[0]), int(h[1])


def gen_c2(seed, n, stream=None):
    """Generate the C2 fact table columns on the device -> 4 int32 tensors."""
    import torch

    cols = [torch.empty(n, dtype=torch.int32, device="cuda") for _ in range(4)]
    call("fls_gen_c2", seed, n, *(c.data_ptr() for c in cols), _stream(stream))
    return cols


def gen_c1(seed, n, bw, stream=None):
    import torch

    out = torch.empty(n, dtype=torch.int32, device="cuda")
    call("fls_gen_c1", seed, n, bw, out.data_ptr(), _stream(stream))
    return out


def encode_c2_table(orderdate, quantity, discount, price):
    """Encode the C2 fact table in FastLanes' unified transposed layout ->
    4 HostColumns: orderdate as FastLanes delta, quantity/discount/price
    bit-packed at their smallest widths after transpose_i (so their rows align
    with the delta-decoded orderdate)."""
    return [delta_column(orderdate)] + [bitpack_column(c, transposed=True) for c in (quantity, discount, price)]


# ----------------------------------------------------------- Q2.1 shape
from ._lib import Q21Dims, Q21Params, Q21Pht  # noqa: E402

C3_DIMS = {"n_date": 2556, "n_supp": 40000, "n_part": 800000, "n_cust": 600000}


def q21_params(category=12, region=2, year_base=1992, n_years=7, n_slots=7000):
    return Q21Params(category, region, year_base, n_years, n_slots)


def gen_c3_fact(seed, n, with_customer=False, stream=None):
    """C3 fact columns on the device (0-based dense keys) -> dict of int32 tensors."""
    import torch

    names = ["datefk", "suppfk", "partfk", "custfk", "revenue"]
    cols = {k: torch.empty(n, dtype=torch.int32, device="cuda") for k in names if with_customer or k != "custfk"}
    call("fls_gen_c3_fact", seed, n, cols["datefk"].data_ptr(), cols["suppfk"].data_ptr(),
         cols["partfk"].data_ptr(), cols["custfk"].data_ptr() if with_customer else None,
         cols["revenue"].data_ptr(), _stream(stream))
    return cols


def gen_c3_dims(seed, n_date=2556, n_supp=40000, n_part=800000, n_cust=600000, stream=None):
    """C3 dimension columns on the device -> dict of int32 tensors."""
    import torch

    e = lambda k: torch.empty(max(1, k), dtype=torch.int32, device="cuda")  # noqa: E731
    d = {"d_year": e(n_date), "s_region": e(n_supp), "s_nation": e(n_supp), "p_category": e(n_part),
         "p_brand": e(n_part), "c_region": e(n_cust), "c_nation": e(n_cust)}
    call("fls_gen_c3_dims", seed, n_date, n_supp, n_part, n_cust,
         *(d[k].data_ptr() for k in ("d_year", "s_region", "s_nation", "p_category", "p_brand", "c_region",
                                     "c_nation")), _stream(stream))
    d.update(n_date=n_date, n_supp=n_supp, n_part=n_part, n_cust=n_cust)
    return d


def key_column(fk0, n_keys):
    """Frame-of-reference column of 1-based keys (key = fk0 + 1, base 1) at
    bw = ceil(log2 n_keys): the packed values are the 0-based keys."""
    fk0 = _u32(fk0)
    bw = max(1, int(n_keys - 1).bit_length())
    return HostColumn(ENC_BITPACK, fk0.size, pack_column(fk0, bw), bw=bw, base=1)


def encode_c3_table(datefk, suppfk, partfk, revenue, n_date=2556, n_supp=40000, n_part=800000):
    """C3 fact table: FOR-packed keys at ceil(log2 n) bits, revenue at its width."""
    return [key_column(datefk, n_date), key_column(suppfk, n_supp), key_column(partfk, n_part),
            bitpack_column(revenue)]


def q21_dims_struct(d, keys_min=1):
    return Q21Dims(d["s_region"].data_ptr(), d["n_supp"], d["p_category"].data_ptr(), d["p_brand"].data_ptr(),
                   d["n_part"], d["d_year"].data_ptr(), d["n_date"], keys_min)


def q21_build(dims, params=None, stream=None):
    """Build the compact dimension workspace (device tensor) for query_q21."""
    import torch

    params = params or q21_params()
    nbytes = int(load().fls_q21_workspace_bytes(dims["n_supp"], dims["n_part"], dims["n_date"]))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    ds = q21_dims_struct(dims)
    call("fls_q21_build", ctypes.byref(ds), ctypes.byref(params), ws.data_ptr(), nbytes, _stream(stream))
    return ws


def _q21_out(params, d_out):
    import torch

    if d_out is None:
        d_out = torch.zeros(2 * params.n_slots, dtype=torch.int64, device="cuda")
    return d_out, d_out.data_ptr(), d_out.data_ptr() + 8 * params.n_slots


def query_q21(date, supp, part, rev, n_rows, dims, ws, params=None, accumulate=False, d_out=None, stream=None):
    """Fused decode + semi-joins + group-by over the compact workspace.
    Returns the [sums(n_slots) | counts(n_slots)] int64 device tensor."""
    params = params or q21_params()
    d_out, ps, pc = _q21_out(params, d_out)
    cols = [c.struct() for c in (date, supp, part, rev)]
    ds = q21_dims_struct(dims)
    call("fls_query_q21", *(ctypes.byref(c) for c in cols), n_rows, ctypes.byref(ds), ws.data_ptr(),
         ctypes.byref(params), ps, pc, Q_ACCUMULATE if accumulate else 0, _stream(stream))
    return d_out


def query_q21_noalloc(date, supp, part, rev, n_rows, dims, params=None, accumulate=False, d_out=None, stream=None):
    """query_q21 without a workspace: the library builds the compact tables
    into its static per-GPU scratch array (fls_query_q21_noalloc; calls that
    overlap on one GPU race).  Allocates nothing itself when d_out is given."""
    params = params or q21_params()
    d_out, ps, pc = _q21_out(params, d_out)
    cols = [c.struct() for c in (date, supp, part, rev)]
    ds = q21_dims_struct(dims)
    call("fls_query_q21_noalloc", *(ctypes.byref(c) for c in cols), n_rows, ctypes.byref(ds), ctypes.byref(params),
         ps, pc, Q_ACCUMULATE if accumulate else 0, _stream(stream))
    return d_out


def build_pht(d_filter, filter_eq, d_keys, d_vals, n, d_ht, ht_len, keys_min, stream=None):
    call("fls_build_pht_u32", dptr(d_filter), filter_eq, dptr(d_keys), dptr(d_vals), n, dptr(d_ht), ht_len,
         keys_min, _stream(stream))


def pht_occupancy(d_ht, ht_len, pair, stream=None):
    """Occupancy bitmap of a reference hash table (fls_pht_occupancy_u32)."""
    import torch

    bits = torch.zeros(max(1, (ht_len + 31) // 32), dtype=torch.int32, device="cuda")
    call("fls_pht_occupancy_u32", dptr(d_ht), ht_len, int(pair), dptr(bits), _stream(stream))
    return bits


def q21_build_pht(dims, params=None, stream=None, occupancy=True):
    """Reference-layout hash tables (build_hashtable_s/p/d) -> (Q21Pht, tensors);
    with occupancy, also the supplier and part tables' occupancy bitmaps."""
    import torch

    params = params or q21_params()
    z = lambda k: torch.zeros(max(1, k), dtype=torch.int32, device="cuda")  # noqa: E731
    keys = {k: torch.arange(1, dims[k] + 1, dtype=torch.int32, device="cuda") for k in ("n_supp", "n_part", "n_date")}
    ht_s, ht_p, ht_d = z(dims["n_supp"]), z(2 * dims["n_part"]), z(2 * dims["n_date"])
    build_pht(dims["s_region"], params.region, keys["n_supp"], None, dims["n_supp"], ht_s, dims["n_supp"], 1, stream)
    build_pht(dims["p_category"], params.category, keys["n_part"], dims["p_brand"], dims["n_part"], ht_p,
              dims["n_part"], 1, stream)
    build_pht(None, 0, keys["n_date"], dims["d_year"], dims["n_date"], ht_d, dims["n_date"], 1, stream)
    keep = [ht_s, ht_p, ht_d]
    occ_s = occ_p = None
    if occupancy:
        occ_s = pht_occupancy(ht_s, dims["n_supp"], False, stream)
        occ_p = pht_occupancy(ht_p, dims["n_part"], True, stream)
        keep += [occ_s, occ_p]
    pht = Q21Pht(ht_s.data_ptr(), dims["n_supp"], 1, ht_p.data_ptr(), dims["n_part"], 1, ht_d.data_ptr(),
                 dims["n_date"], 1, dptr(occ_s) if occ_s is not None else None,
                 dptr(occ_p) if occ_p is not None else None)
    return pht, tuple(keep)


def query_q21_pht(date, supp, part, rev, n_rows, pht, params=None, accumulate=False, d_out=None, stream=None):
    params = params or q21_params()
    d_out, ps, pc = _q21_out(params, d_out)
    cols = [c.struct() for c in (date, supp, part, rev)]
    call("fls_query_q21_pht", *(ctypes.byref(c) for c in cols), n_rows, ctypes.byref(pht), ctypes.byref(params),
         ps, pc, Q_ACCUMULATE if accumulate else 0, _stream(stream))
    return d_out


def q21_groups(d_out, params=None):
    """Result rows (year, brand, sum) of the non-empty slots in slot order --
    the compaction of runQuery (fls_q21.cu:341-347)."""
    params = params or q21_params()
    h = d_out.cpu().numpy()
    sums, counts = h[: params.n_slots], h[params.n_slots:]
    rows = []
    for slot in np.nonzero(counts)[0]:
        brand = int(slot) // params.n_years
        if brand == 0:  # brand n_slots/n_years wraps onto slot 0..n_years-1
            brand = params.n_slots // params.n_years
        rows.append((params.year_base + int(slot) % params.n_years, brand, int(sums[slot])))
    return rows


# ----------------------------------------------------- Q3.1 / Q4.1 shapes
from ._lib import Q31Params, Q41Params, StarDims  # noqa: E402

SUPPLYCOST_BASE = 60000


def q31_params(c_region=3, s_region=3, year_lo=1992, year_hi=1997, year_base=1992, n_years=7, n_nations=25,
               n_slots=4375):
    """Q3.1 constants (fls_q31.cu:84-85) with the region code of ASIA in the
    1-based synthetic dimension domain."""
    return Q31Params(c_region, s_region, year_lo, year_hi, year_base, n_years, n_nations, n_slots)


def q41_params(c_region=2, s_region=2, mfgr_lo=1, mfgr_hi=2, year_base=1992, n_years=7, n_slots=175):
    """Q4.1 constants (fls_q41.cu:97) with AMERICA and MFGR#1/#2 in the 1-based
    synthetic dimension domain."""
    return Q41Params(c_region, s_region, mfgr_lo, mfgr_hi, year_base, n_years, n_slots)


def gen_star_extra(seed, n, n_part=800000, stream=None):
    """Supplycost per fact row and manufacturer per part on the device."""
    import torch

    d = {"supplycost": torch.empty(max(1, n), dtype=torch.int32, device="cuda"),
         "p_mfgr": torch.empty(max(1, n_part), dtype=torch.int32, device="cuda")}
    call("fls_gen_star_extra", seed, n, d["supplycost"].data_ptr(), n_part, d["p_mfgr"].data_ptr(),
         _stream(stream))
    return d


def encode_star_table(datefk, custfk, suppfk, partfk, revenue, supplycost, n_date=2556, n_cust=600000,
                      n_supp=40000, n_part=800000):
    """Star-join fact table (date, customer, supplier, part, revenue,
    supplycost): FOR-packed 1-based keys, revenue at its width, supplycost
    frame-of-reference from 60000."""
    return [key_column(datefk, n_date), key_column(custfk, n_cust), key_column(suppfk, n_supp),
            key_column(partfk, n_part), bitpack_column(revenue), bitpack_column(supplycost, base=SUPPLYCOST_BASE)]


def star_dims_struct(d, keys_min=1):
    p = lambda k: d[k].data_ptr() if k in d else None  # noqa: E731
    return StarDims(p("s_region"), p("s_nation"), d["n_supp"], p("c_region"), p("c_nation"), d["n_cust"],
                    p("p_mfgr"), d.get("n_part", 0) if "p_mfgr" in d else 0, p("d_year"), d["n_date"], keys_min)


def _star_build(fn, dims, params, stream):
    import torch

    ds = star_dims_struct(dims)
    nbytes = int(load().fls_star_workspace_bytes(ctypes.byref(ds)))
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    call(fn, ctypes.byref(ds), ctypes.byref(params), ws.data_ptr(), nbytes, _stream(stream))
    return ws


def q31_build(dims, params=None, stream=None):
    """Compact Q3.1 workspace (supplier / customer nation codes, year codes)."""
    return _star_build("fls_q31_build", dims, params or q31_params(), stream)


def q41_build(dims, params=None, stream=None):
    """Compact Q4.1 workspace (supplier codes, customer nations, part bits, years);
    dims needs "p_mfgr" (from gen_star_extra)."""
    return _star_build("fls_q41_build", dims, params or q41_params(), stream)


def query_q31(date, cust, supp, rev, n_rows, dims, ws, params=None, accumulate=False, d_out=None, stream=None):
    """Fused decode + three joins + group-by (Q3.1 shape).  Returns the
    [sums(n_slots) | counts(n_slots)] int64 device tensor."""
    params = params or q31_params()
    d_out, ps, pc = _q21_out(params, d_out)
    cols = [c.struct() for c in (date, cust, supp, rev)]
    ds = star_dims_struct(dims)
    call("fls_query_q31", *(ctypes.byref(c) for c in cols), n_rows, ctypes.byref(ds), ws.data_ptr(),
         ctypes.byref(params), ps, pc, Q_ACCUMULATE if accumulate else 0, _stream(stream))
    return d_out


def query_q41(date, cust, supp, part, rev, cost, n_rows, dims, ws, params=None, accumulate=False, d_out=None,
              stream=None):
    """Fused decode + four joins + group-by of revenue - supplycost (Q4.1
    shape).  Returns the [sums(n_slots) | counts(n_slots)] int64 device tensor."""
    params = params or q41_params()
    d_out, ps, pc = _q21_out(params, d_out)
    cols = [c.struct() for c in (date, cust, supp, part, rev, cost)]
    ds = star_dims_struct(dims)
    call("fls_query_q41", *(ctypes.byref(c) for c in cols), n_rows, ctypes.byref(ds), ws.data_ptr(),
         ctypes.byref(params), ps, pc, Q_ACCUMULATE if accumulate else 0, _stream(stream))
    return d_out


# ------------------------------------------------ untranspose, materialize
ORDER_STORED, ORDER_ROWS = 0, 1


def transpose_vectors(d_in, n_vec, inverse=False, out=None, stream=None):
    """transpose_i (or, inverse, transpose_o) of n_vec whole 1024-value vectors
    on the device; returns the output tensor."""
    if out is None:
        out = _dev_u32(max(1, n_vec * VECTOR_SIZE))
    call("fls_transpose_u32", dptr(d_in), n_vec, 1 if inverse else 0, dptr(out), _stream(stream))
    return out


def decode_column(col, n_rows=None, order=ORDER_ROWS, out=None, stream=None):
    """Decode a device column of any encoding to plain uint32 values (int32
    tensor): original row order (n_rows values) or the stored order (whole
    vectors)."""
    n_rows = col.n_rows if n_rows is None else n_rows
    n = n_rows if order == ORDER_ROWS else (n_rows + VECTOR_SIZE - 1) // VECTOR_SIZE * VECTOR_SIZE
    if out is None:
        out = _dev_u32(max(1, n))
    s = col.struct()
    call("fls_decode_column_u32", ctypes.byref(s), n_rows, order, dptr(out), _stream(stream))
    return out[:n]


# ----------------------------------------------------------- GPU encoders
def _dev_u32(n):
    import torch

    return torch.empty(max(1, n), dtype=torch.int32, device="cuda")


def _dev_u64(n):
    import torch

    return torch.empty(max(1, n), dtype=torch.int64, device="cuda")


def pack_column_device(d_in, n, bw, base=0, transposed=False, stream=None):
    """GPU bit-packing of n values (minus base) -> int32 tensor of packed words."""
    out = _dev_u32(packed_words(n, bw))
    call("fls_pack_column_device", dptr(d_in), n, bw, base, ENC_FLAG_TRANSPOSE if transposed else 0, dptr(out),
         _stream(stream))
    return out


def delta_encode_device(d_in, n, stream=None):
    """GPU FastLanes delta: plan (widths, bases, offsets), then pack.
    Returns (words, vmeta, bases) device tensors."""
    import torch

    n_vec = (n + VECTOR_SIZE - 1) // VECTOR_SIZE
    vmeta, bases = _dev_u64(n_vec), _dev_u32(32 * n_vec)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    scratch = torch.empty(max(16, int(load().fls_encode_scratch_bytes(n))), dtype=torch.uint8, device="cuda")
    call("fls_delta_plan_device", dptr(d_in), n, dptr(vmeta), dptr(bases), dptr(total), dptr(scratch),
         _stream(stream))
    words = _dev_u32(int(total.item()))
    call("fls_delta_pack_device", dptr(d_in), n, dptr(vmeta), dptr(words), _stream(stream))
    return words[: max(1, int(total.item()))], vmeta[:max(1, n_vec)], bases


def rle_encode_device(d_in, n, stream=None):
    """GPU FastLanes RLE: plan then pack.  Returns (words, vmeta, rmeta, bases)
    device tensors; rmeta holds one 8-byte fls_rle_meta per vector."""
    import torch

    n_vec = (n + VECTOR_SIZE - 1) // VECTOR_SIZE
    vmeta, rmeta, bases = _dev_u64(n_vec), _dev_u64(n_vec), _dev_u32(32 * n_vec)
    total = torch.zeros(1, dtype=torch.int64, device="cuda")
    scratch = torch.empty(max(16, int(load().fls_encode_scratch_bytes(n))), dtype=torch.uint8, device="cuda")
    call("fls_rle_plan_device", dptr(d_in), n, dptr(vmeta), dptr(rmeta), dptr(bases), dptr(total), dptr(scratch),
         _stream(stream))
    words = _dev_u32(int(total.item()))
    call("fls_rle_pack_device", dptr(d_in), n, dptr(vmeta), dptr(rmeta), dptr(words), _stream(stream))
    return words, vmeta, rmeta, bases


def rle_decode(d_words, d_vmeta, d_rmeta, d_bases, n_vec, d_out, stream=None):
    call("fls_rle_decode_u32", dptr(d_words), dptr(d_vmeta), dptr(d_rmeta), dptr(d_bases), n_vec, dptr(d_out),
         _stream(stream))


# ----------------------------------------------------------- batched segments
from ._lib import BatchInfo, Segment, SegmentInfo, Tune  # noqa: E402

SCHED_STATIC, SCHED_DYNAMIC = 0, 1


class Batch:
    """Owner of an fls_batch handle: many compressed segments decoded and
    range-filtered in one persistent launch (plus graph, autotuner, baseline)."""

    def __init__(self, handle, keep=None):
        self._h = handle
        self._keep = keep  # device buffers a caller-described batch borrows

    @property
    def handle(self):
        return self._h

    def info(self):
        i = BatchInfo()
        call("fls_batch_get_info", self._h, ctypes.byref(i))
        return {k: getattr(i, k) for k, _ in BatchInfo._fields_}

    def segment(self, s):
        i = SegmentInfo()
        call("fls_batch_segment", self._h, s, ctypes.byref(i))
        return {k: getattr(i, k) for k, _ in SegmentInfo._fields_}

    def segment_host(self, s):
        """Segment s copied to host memory: a HostColumn-like object (numpy
        words, vmeta, bases and RLE metadata as uint64 aux) plus n_values, lo,
        hi (fls_batch_segment_column, then device-to-host copies)."""
        import numpy as np

        seg = Segment()
        call("fls_batch_segment_column", self._h, s, ctypes.byref(seg))
        c = seg.col
        nv = (seg.n_values + 1023) // 1024
        fetch = lambda ptr, n, dt: _device_to_host(ptr, n, dt) if ptr and n else None  # noqa: E731
        hc = HostColumn(c.encoding, seg.n_values, fetch(c.words, c.n_words, np.uint32), bw=c.bw, base=c.base,
                        transposed=bool(c.transposed),
                        vmeta=fetch(c.vmeta, nv, np.uint64) if c.encoding != ENC_BITPACK else None,
                        bases=fetch(c.bases, 32 * nv, np.uint32) if c.encoding in (ENC_DELTA, ENC_RLE) else None,
                        aux=fetch(c.aux, nv, np.uint64) if c.encoding in (ENC_RLE, ENC_MIXED) else None)
        return hc, seg.n_values, seg.lo, seg.hi

    def run(self, tune, d_counts, stream=None):
        call("fls_batch_run", self._h, ctypes.byref(tune), dptr(d_counts), _stream(stream))

    def run_per_segment(self, tune, d_counts, stream=None):
        call("fls_batch_run_per_segment", self._h, ctypes.byref(tune), dptr(d_counts), _stream(stream))

    def graph_launch(self, tune, d_counts, stream=None):
        call("fls_batch_graph_launch", self._h, ctypes.byref(tune), dptr(d_counts), _stream(stream))

    def autotune(self, d_counts, cache_path=None, force=False, stream=None):
        t = Tune()
        ms = ctypes.c_double(0)
        call("fls_batch_autotune", self._h, cache_path.encode() if cache_path else None, int(force), dptr(d_counts),
             ctypes.byref(t), ctypes.byref(ms), _stream(stream))
        return t, ms.value

    def close(self):
        if self._h:
            load().fls_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover
            pass


def tune(tile_vectors=64, block_threads=256, schedule=SCHED_STATIC, class_tiles=None):
    """Launch shape of a batch; class_tiles: optional {class: tile size} with
    class = 2*encoding + (more than 8 encoded bits per value)."""
    packed = 0
    for c, tv in (class_tiles or {}).items():
        assert tv & (tv - 1) == 0 and 2 <= tv <= 32768, tv
        packed |= (tv.bit_length() - 1) << (4 * c)
    return Tune(tile_vectors, block_threads, schedule, packed)


def _device_to_host(ptr, n, dtype):
    """n elements of dtype at device address ptr, copied into a new numpy
    array (hipMemcpy on the HIP runtime this process already uses)."""
    import numpy as np

    out = np.empty(n, dtype)
    # through libfls.so's own handle: dlsym finds the HIP runtime libfls.so is
    # bound to (the one torch loaded), never a second copy
    memcpy = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int)(
        ("hipMemcpy", load()))
    st = memcpy(out.ctypes.data, ptr, out.nbytes, 2)  # hipMemcpyDeviceToHost
    if st != 0:
        raise RuntimeError(f"hipMemcpy device-to-host failed: {st}")
    return out


def batch_c4(seed, n_segments=4096, log2_min_len=10, log2_max_len=24, stream=None):
    """Build the batched-segments workload on the GPU (see fls_batch_create_c4)."""
    h = ctypes.c_void_p()
    call("fls_batch_create_c4", seed, n_segments, log2_min_len, log2_max_len, ctypes.byref(h), _stream(stream))
    return Batch(h.value)


def c4_segment_params(seed, s, log2_min_len=10, log2_max_len=24):
    enc, bw, n = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
    load().fls_c4_segment_params(seed, s, log2_min_len, log2_max_len, ctypes.byref(enc), ctypes.byref(bw),
                                 ctypes.byref(n))
    return enc.value, bw.value, n.value


def gen_c4_segment(seed, s, log2_min_len=10, log2_max_len=24, stream=None):
    import torch

    n = c4_segment_params(seed, s, log2_min_len, log2_max_len)[2]
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    call("fls_gen_c4_segment", seed, s, log2_min_len, log2_max_len, out.data_ptr(), _stream(stream))
    return out


def batch_from_segments(segments):
    """segments: list of dicts {col: Column, rmeta: tensor|None, n_values, lo, hi}."""
    arr = (Segment * len(segments))()
    keep = []
    for i, s in enumerate(segments):
        arr[i].col = s["col"].struct()
        arr[i].rmeta = dptr(s.get("rmeta"))
        arr[i].n_values, arr[i].lo, arr[i].hi = s["n_values"], s["lo"], s["hi"]
        keep.append(s)
    h = ctypes.c_void_p()
    call("fls_batch_create", arr, len(segments), ctypes.byref(h))
    return Batch(h.value, keep)


# ------------------------------------- tile-based competitor formats (comparison)
class TileColumn:
    """A GPU-FOR / GPU-DFOR / GPU-RFOR encoded column (host numpy arrays, or
    device tensors after to_device()).  kind: "for", "dfor" or "rfor"."""

    def __init__(self, kind, n, words, starts, rl_words=None, rl_starts=None):
        self.kind, self.n, self.words, self.starts = kind, n, words, starts
        self.rl_words, self.rl_starts = rl_words, rl_starts

    def nbytes(self):
        extra = 0 if self.rl_words is None else 4 * (len(self.rl_words) + len(self.rl_starts))
        return 4 * (len(self.words) + len(self.starts)) + extra

    def to_device(self):
        import torch

        dev = lambda a: None if a is None else torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()  # noqa
        return TileColumn(self.kind, self.n, dev(self.words), dev(self.starts), dev(self.rl_words),
                          dev(self.rl_starts))


def tile_encode(values, kind):
    """Host encode of uint32 values into the tile-based format `kind`."""
    x = _u32(values)
    n = x.size
    cap = int(load().fls_tile_max_words(n))
    if kind == "rfor":
        nt = (n + 511) // 512
        v, r = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
        vs, rs = np.zeros(nt + 1, np.uint32), np.zeros(nt + 1, np.uint32)
        nv, nr = ctypes.c_uint64(0), ctypes.c_uint64(0)
        call("fls_tile_rfor_encode_u32", x.ctypes.data, n, v.ctypes.data, vs.ctypes.data, r.ctypes.data,
             rs.ctypes.data, ctypes.byref(nv), ctypes.byref(nr))
        return TileColumn(kind, n, v[:nv.value], vs, r[:nr.value], rs)
    nb = (n + 127) // 128 if kind == "for" else 4 * ((n + 511) // 512)
    out, starts = np.zeros(cap, np.uint32), np.zeros(nb + 1, np.uint32)
    nw = ctypes.c_uint64(0)
    call(f"fls_tile_{kind}_encode_u32", x.ctypes.data, n, out.ctypes.data, starts.ctypes.data, ctypes.byref(nw))
    return TileColumn(kind, n, out[:nw.value], starts)


def tile_decode(col, out=None, d_sum=None, values=True, stream=None):
    """Decode a device TileColumn: returns (values tensor or None, sum tensor or None)."""
    import torch

    if values and out is None:
        out = torch.empty(max(1, col.n), dtype=torch.int32, device="cuda")
    if not values:
        out = None
    if d_sum is None:
        d_sum = torch.zeros(1, dtype=torch.int64, device="cuda")
    if col.kind == "rfor":
        call("fls_tile_rfor_decode_u32", dptr(col.words), dptr(col.starts), dptr(col.rl_words), dptr(col.rl_starts),
             col.n, dptr(out), dptr(d_sum), _stream(stream))
    else:
        call(f"fls_tile_{col.kind}_decode_u32", dptr(col.words), dptr(col.starts), col.n, dptr(out), dptr(d_sum),
             _stream(stream))
    return (out[:col.n] if out is not None else None), d_sum
