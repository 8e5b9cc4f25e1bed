"""libfls.so loads and exports every function that include/fls/*.h declares."""
import re

import fastlanesgpu_amd as fls
import fls_testlib as T


def declared_functions():
    names = set()
    for h in (T.ROOT / "include" / "fls").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(fls_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported():
    lib = fls.load()
    names = declared_functions()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_every_declared_symbol_has_a_python_signature():
    from fastlanesgpu_amd._lib import SIGNATURES

    missing = sorted(declared_functions() - set(SIGNATURES))
    assert not missing, missing


def test_version():
    assert fls.load().fls_version().startswith(b"fls-mi355x")


def test_reference_shaped_test_library_loads():
    # the reference-shaped kernels compiled against include/fls/device (run by
    # tests/test_gpu_dropin.py) load without a GPU and export every launcher
    import ctypes

    lib = ctypes.CDLL(str(T.ROOT / "tests" / "dropin" / "librefshape.so"))
    for name in ("rs_q11", "rs_q21", "rs_q31", "rs_delta", "rs_unpack_named", "rs_unpack_quarters",
                 "rs_unpack_global", "rs_prims", "rs_simd16", "rs_n_pred_out"):
        assert hasattr(lib, name), name
    assert lib.rs_n_pred_out() == 22


def test_mlp_forward_rejects_unaligned_features():
    # the argument checks run before any device work, so they run here too
    lib = fls.load()
    err_arg = -6  # FLS_ERR_ARG
    got = lib.fls_mlp_forward(0x10000, 0x10004, 16, 0x20000, None, None, None)
    assert got == err_arg, got
