"""ctypes binding of libfls.so (the C ABI declared in include/fls/fls.h).

The library is built in-tree by `make -C fastlanesgpu_amd`; nothing here falls
back to another implementation: if libfls.so is missing or a symbol is absent
the load raises.
"""
import ctypes
import os
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = _HERE / "libfls.so"

c_u32p = ctypes.POINTER(ctypes.c_uint32)
c_u64p = ctypes.POINTER(ctypes.c_uint64)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_vp = ctypes.c_void_p
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int

# name -> (restype, argtypes).  Pointers are passed as c_void_p (ints).
SIGNATURES = {
    "fls_status_string": (ctypes.c_char_p, [c_int]),
    "fls_last_hip_error": (c_int, []),
    "fls_version": (ctypes.c_char_p, []),
    "fls_pack_u32": (c_int, [c_vp, c_vp, c_u32]),
    "fls_transpose_i_u32": (c_int, [c_vp, c_vp]),
    "fls_transpose_o_u32": (c_int, [c_vp, c_vp]),
    "fls_unrsum_u32": (c_int, [c_vp, c_vp]),
    "fls_packed_words": (c_u64, [c_u64, c_u32]),
    "fls_pack_column_u32": (c_int, [c_vp, c_u64, c_u32, c_u32, c_vp]),
    "fls_delta_encode_column_u32": (c_int, [c_vp, c_u64, c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_unpack_u32": (c_int, [c_vp, c_vp, c_u64, c_u32, c_vp]),
    "fls_unpack_u32_checksum": (c_int, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp]),
    "fls_delta_decode_u32": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp, c_vp]),
}



class FlsColumn(ctypes.Structure):
    """struct fls_column (include/fls/fls.h)."""
    _fields_ = [("encoding", c_u32), ("bw", c_u32), ("base", c_u32), ("transposed", c_u32),
                ("n_words", c_u64), ("words", c_vp), ("vmeta", c_vp), ("bases", c_vp), ("aux", c_vp)]


class Q11Params(ctypes.Structure):
    """struct fls_q11_params."""
    _fields_ = [("date_gt", c_u32), ("date_lt", c_u32), ("qty_lt", c_u32), ("disc_lo", c_u32), ("disc_hi", c_u32)]


COLP = ctypes.POINTER(FlsColumn)

SIGNATURES.update({
    "fls_column_image_bytes": (c_u64, [COLP, c_u64]),
    "fls_column_serialize": (c_int, [COLP, c_u64, c_vp, c_u64]),
    "fls_column_deserialize": (c_int, [c_vp, c_u64, COLP, c_u64p]),
    "fls_query_q11": (c_int, [COLP, COLP, COLP, COLP, c_u64, ctypes.POINTER(Q11Params), c_u32, c_vp, c_vp, c_vp]),
    "fls_column_bytes": (c_u64, [COLP, c_u64]),
    "fls_gen_c1": (c_int, [c_u64, c_u64, c_u32, c_vp, c_vp]),
    "fls_copy_bytes": (c_int, [c_vp, c_vp, c_u64, c_vp]),
    "fls_read_bytes": (c_int, [c_vp, c_u64, c_vp, c_vp]),
    "fls_gen_c2": (c_int, [c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_gen_c3_fact": (c_int, [c_u64, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_gen_c3_dims": (c_int, [c_u64, c_u32, c_u32, c_u32, c_u32] + [c_vp] * 8),
})

_lib = None


class FlsError(RuntimeError):
    def __init__(self, func, status, lib):
        msg = lib.fls_status_string(status).decode()
        hip = lib.fls_last_hip_error() if status == -4 else 0
        super().__init__(f"{func} failed: status {status} ({msg})" + (f", hip error {hip}" if hip else ""))
        self.status = status


def load(path=None):
    """Load libfls.so and declare every C-ABI signature.  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = pathlib.Path(path) if path else pathlib.Path(os.environ.get("FLS_LIB", LIB_PATH))
    if not p.exists():
        raise FileNotFoundError(f"{p} is not built: run `make -C {_HERE}` (or __graft_entry__.build())")
    # torch wheels ship their own libamdhip64.so.7.  Loading torch first makes
    # libfls.so bind to that already-loaded runtime, so the process has one
    # HIP runtime; loading libfls.so first would pull in the system copy and
    # leave two runtimes competing for the device.
    try:
        import torch  # noqa: F401
    except ImportError:  # pragma: no cover - C/C++ hosts do not need torch
        pass
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if the export is missing
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(lib, name, status):
    if status != 0:
        raise FlsError(name, status, lib)


def call(name, *args):
    """Call a status-returning C-ABI function and raise FlsError on failure."""
    lib = load()
    st = getattr(lib, name)(*args)
    check(lib, name, st)
    return st


class Q21Params(ctypes.Structure):
    """struct fls_q21_params."""
    _fields_ = [("category", c_u32), ("region", c_u32), ("year_base", c_u32), ("n_years", c_u32),
                ("n_slots", c_u32)]


class Q21Pht(ctypes.Structure):
    """struct fls_q21_pht."""
    _fields_ = [("ht_s", c_vp), ("s_len", c_u32), ("s_min", c_u32), ("ht_p", c_vp), ("p_len", c_u32),
                ("p_min", c_u32), ("ht_d", c_vp), ("d_len", c_u32), ("d_min", c_u32), ("occ_s", c_vp),
                ("occ_p", c_vp)]


class Q21Dims(ctypes.Structure):
    """struct fls_q21_dims."""
    _fields_ = [("s_region", c_vp), ("n_supp", c_u32), ("p_category", c_vp), ("p_brand", c_vp), ("n_part", c_u32),
                ("d_year", c_vp), ("n_date", c_u32), ("keys_min", c_u32)]


_Q21P = ctypes.POINTER(Q21Params)
SIGNATURES.update({
    "fls_build_pht_u32": (c_int, [c_vp, c_u32, c_vp, c_vp, c_u32, c_vp, c_u32, c_u32, c_vp]),
    "fls_query_q21_pht": (c_int, [COLP, COLP, COLP, COLP, c_u64, ctypes.POINTER(Q21Pht), _Q21P, c_vp, c_vp, c_u32,
                                  c_vp]),
    "fls_pht_occupancy_u32": (c_int, [c_vp, c_u32, c_u32, c_vp, c_vp]),
    "fls_q21_workspace_bytes": (c_u64, [c_u32, c_u32, c_u32]),
    "fls_q21_build": (c_int, [ctypes.POINTER(Q21Dims), _Q21P, c_vp, c_u64, c_vp]),
    "fls_query_q21": (c_int, [COLP, COLP, COLP, COLP, c_u64, ctypes.POINTER(Q21Dims), c_vp, _Q21P, c_vp, c_vp,
                              c_u32, c_vp]),
    "fls_query_q21_noalloc": (c_int, [COLP, COLP, COLP, COLP, c_u64, ctypes.POINTER(Q21Dims), _Q21P, c_vp, c_vp,
                                      c_u32, c_vp]),
})

class StarDims(ctypes.Structure):
    """struct fls_star_dims."""
    _fields_ = [("s_region", c_vp), ("s_nation", c_vp), ("n_supp", c_u32), ("c_region", c_vp), ("c_nation", c_vp),
                ("n_cust", c_u32), ("p_mfgr", c_vp), ("n_part", c_u32), ("d_year", c_vp), ("n_date", c_u32),
                ("keys_min", c_u32)]


class Q31Params(ctypes.Structure):
    """struct fls_q31_params."""
    _fields_ = [("c_region", c_u32), ("s_region", c_u32), ("year_lo", c_u32), ("year_hi", c_u32),
                ("year_base", c_u32), ("n_years", c_u32), ("n_nations", c_u32), ("n_slots", c_u32)]


class Q41Params(ctypes.Structure):
    """struct fls_q41_params."""
    _fields_ = [("c_region", c_u32), ("s_region", c_u32), ("mfgr_lo", c_u32), ("mfgr_hi", c_u32),
                ("year_base", c_u32), ("n_years", c_u32), ("n_slots", c_u32)]


SIGNATURES.update({
    "fls_transpose_u32": (c_int, [c_vp, c_u64, c_u32, c_vp, c_vp]),
    "fls_decode_column_u32": (c_int, [COLP, c_u64, c_u32, c_vp, c_vp]),
})

SIGNATURES.update({
    "fls_tile_max_words": (c_u64, [c_u64]),
    "fls_tile_for_encode_u32": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_tile_dfor_encode_u32": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_tile_rfor_encode_u32": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_tile_for_decode_u32": (c_int, [c_vp, c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_tile_dfor_decode_u32": (c_int, [c_vp, c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_tile_rfor_decode_u32": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp]),
})

_SDP = ctypes.POINTER(StarDims)
SIGNATURES.update({
    "fls_gen_star_extra": (c_int, [c_u64, c_u64, c_vp, c_u32, c_vp, c_vp]),
    "fls_star_workspace_bytes": (c_u64, [_SDP]),
    "fls_q31_build": (c_int, [_SDP, ctypes.POINTER(Q31Params), c_vp, c_u64, c_vp]),
    "fls_query_q31": (c_int, [COLP, COLP, COLP, COLP, c_u64, _SDP, c_vp, ctypes.POINTER(Q31Params), c_vp, c_vp,
                              c_u32, c_vp]),
    "fls_q41_build": (c_int, [_SDP, ctypes.POINTER(Q41Params), c_vp, c_u64, c_vp]),
    "fls_query_q41": (c_int, [COLP, COLP, COLP, COLP, COLP, COLP, c_u64, _SDP, c_vp, ctypes.POINTER(Q41Params),
                              c_vp, c_vp, c_u32, c_vp]),
})

SIGNATURES.update({
    "fls_encode_scratch_bytes": (c_u64, [c_u64]),
    "fls_pack_column_device": (c_int, [c_vp, c_u64, c_u32, c_u32, c_u32, c_vp, c_vp]),
    "fls_delta_plan_device": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_delta_pack_device": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_rle_plan_device": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_rle_pack_device": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp, c_vp]),
    "fls_rle_decode_u32": (c_int, [c_vp, c_vp, c_vp, c_vp, c_u64, c_vp, c_vp]),
})

SIGNATURES.update({
    "fls_exhaustive_select": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp]),
    "fls_vector_features": (c_int, [c_vp, c_u64, c_vp, c_vp]),
    "fls_mixed_max_words": (c_u64, [c_u64]),
    "fls_mixed_plan_device": (c_int, [c_vp, c_u64, c_vp, c_vp, ctypes.c_float, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_trial_workspace_bytes": (c_u64, [c_u64]),
    "fls_mixed_encode_trial": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_mixed_pack_device": (c_int, [c_vp, c_u64, c_vp, c_vp, c_vp, c_vp]),
    "fls_mixed_encode_device": (c_int, [c_vp, c_u64, c_vp, c_vp, ctypes.c_float, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "fls_mixed_decode_u32": (c_int, [c_vp, c_vp, c_vp, c_u64, c_vp, c_vp]),
    "fls_gen_train_vectors": (c_int, [c_u64, c_u64, c_u64, c_vp, c_vp, c_vp]),
    "fls_mlp_param_count": (c_u32, []),
    "fls_mlp_to_bf16": (c_int, [c_vp, c_vp, c_vp]),
    "fls_mlp_forward": (c_int, [c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp]),
    "fls_mlp_grad_workspace_bytes": (c_u64, [c_u64]),
    "fls_mlp_grad": (c_int, [c_vp, c_vp, c_vp, c_u64, ctypes.c_float, c_vp, c_vp, c_vp, c_vp]),
    "fls_mlp_adam": (c_int, [c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                             ctypes.c_float, c_vp]),
    "fls_mlp_mfma_selftest": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "fls_mlp_mfma32_selftest": (c_int, [c_vp, c_vp, c_vp, c_vp]),
})


class Tune(ctypes.Structure):
    """struct fls_tune."""
    _fields_ = [("tile_vectors", c_u32), ("block_threads", c_u32), ("schedule", c_u32), ("class_tiles", c_u32)]


class BatchInfo(ctypes.Structure):
    """struct fls_batch_info."""
    _fields_ = [("n_segments", c_u32), ("n_values", c_u64), ("n_vectors", c_u64), ("encoded_bytes", c_u64),
                ("n_bitpack", c_u32), ("n_delta", c_u32), ("n_rle", c_u32)]


class SegmentInfo(ctypes.Structure):
    """struct fls_segment_info."""
    _fields_ = [("encoding", c_u32), ("bw", c_u32), ("n_values", c_u64), ("lo", c_u32), ("hi", c_u32),
                ("n_words", c_u64)]


class Segment(ctypes.Structure):
    """struct fls_segment."""
    _fields_ = [("col", FlsColumn), ("rmeta", c_vp), ("n_values", c_u64), ("lo", c_u32), ("hi", c_u32)]


_TP = ctypes.POINTER(Tune)
SIGNATURES.update({
    "fls_batch_create": (c_int, [ctypes.POINTER(Segment), c_u32, ctypes.POINTER(c_vp)]),
    "fls_batch_create_c4": (c_int, [c_u64, c_u32, c_u32, c_u32, ctypes.POINTER(c_vp), c_vp]),
    "fls_c4_segment_params": (None, [c_u64, c_u32, c_u32, c_u32, c_u32p, c_u32p, c_u64p]),
    "fls_gen_c4_segment": (c_int, [c_u64, c_u32, c_u32, c_u32, c_vp, c_vp]),
    "fls_batch_get_info": (c_int, [c_vp, ctypes.POINTER(BatchInfo)]),
    "fls_batch_segment": (c_int, [c_vp, c_u32, ctypes.POINTER(SegmentInfo)]),
    "fls_batch_segment_column": (c_int, [c_vp, c_u32, ctypes.POINTER(Segment)]),
    "fls_batch_run": (c_int, [c_vp, _TP, c_vp, c_vp]),
    "fls_batch_run_per_segment": (c_int, [c_vp, _TP, c_vp, c_vp]),
    "fls_batch_graph_launch": (c_int, [c_vp, _TP, c_vp, c_vp]),
    "fls_batch_autotune": (c_int, [c_vp, ctypes.c_char_p, c_int, c_vp, _TP, ctypes.POINTER(ctypes.c_double), c_vp]),
    "fls_batch_destroy": (None, [c_vp]),
})
