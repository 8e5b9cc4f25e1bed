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


def unpack_checksum(d_in, d_out, n_values, bw, d_checksum, stream=None):
    call("fls_unpack_u32_checksum", dptr(d_in), dptr(d_out), n_values, bw, dptr(d_checksum), _stream(stream))


def delta_decode(d_words, d_vmeta, d_bases, n_vec, d_out, stream=None):
    call("fls_delta_decode_u32", dptr(d_words), dptr(d_vmeta), dptr(d_bases), n_vec, dptr(d_out), _stream(stream))


# ----------------------------------------------------------- columns
ENC_BITPACK, ENC_DELTA, ENC_RLE = 0, 1, 2
Q_PRED_LOAD = 1
Q_ACCUMULATE = 2

from ._lib import FlsColumn, Q11Params  # noqa: E402


class Column:
    """A compressed column resident on the device (torch tensors own the memory)."""

    def __init__(self, encoding, n_rows, words, bw=0, base=0, transposed=False, vmeta=None, bases=None,
                 n_words=None):
        self.encoding, self.n_rows, self.bw, self.base = encoding, n_rows, bw, base
        self.transposed = bool(transposed) or encoding in (ENC_DELTA, ENC_RLE)
        self.words, self.vmeta, self.bases = words, vmeta, bases
        self.n_words = int(n_words if n_words is not None else (words.numel() if words is not None else 0))

    def struct(self):
        return FlsColumn(self.encoding, self.bw, self.base, int(self.transposed), self.n_words,
                         dptr(self.words), dptr(self.vmeta), dptr(self.bases))

    def nbytes(self):
        s = self.struct()
        return int(load().fls_column_bytes(ctypes.byref(s), self.n_rows))


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

    def __init__(self, encoding, n_rows, words, bw=0, base=0, transposed=False, vmeta=None, bases=None):
        self.encoding, self.n_rows, self.bw, self.base = encoding, n_rows, bw, base
        self.transposed = bool(transposed) or encoding in (ENC_DELTA, ENC_RLE)
        self.words, self.vmeta, self.bases = words, vmeta, bases

    def to_device(self):
        return Column(self.encoding, self.n_rows, _torch_dev(self.words), bw=self.bw, base=self.base,
                      transposed=self.transposed,
                      vmeta=None if self.vmeta is None else _torch_dev(self.vmeta),
                      bases=None if self.bases is None else _torch_dev(self.bases), n_words=self.words.size)

    def nbytes(self):
        s = FlsColumn(self.encoding, self.bw, self.base, int(self.transposed), self.words.size, None, None, None)
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
