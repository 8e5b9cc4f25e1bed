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
