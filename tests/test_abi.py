"""CPU: the C-ABI library loads, exports every symbol include/hasty_hip.h
declares, and rejects invalid arguments with the documented status codes
before touching the GPU."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hasty_hip.h")


@pytest.fixture(scope="module")
def lib():
    from hastycucompute_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from hastycucompute_amd import build
        build.build()
    return _lib.lib()


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hasty_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("hasty_nufft_makeplan", "hasty_nufft_setpts", "hasty_nufft_execute",
                 "hasty_nufft_destroy", "hasty_toeplitz_kernel", "hasty_toeplitz_multiply",
                 "hasty_toeplitz_multiplication", "hasty_normal_create", "hasty_normal_apply",
                 "hasty_cg_step", "hasty_dot"):
        assert must in names


def test_every_declared_symbol_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    from hastycucompute_amd._lib import SIGNATURES
    assert set(SIGNATURES) == set(declared_functions())


def test_version(lib):
    assert lib.hasty_version() >= 100


def _err(lib):
    return lib.hasty_last_error().decode()


def test_makeplan_rejects_invalid_arguments(lib):
    from hastycucompute_amd._lib import NufftOpts, i64_array
    h = ctypes.c_void_p()
    modes = i64_array([8, 8, 8])
    o = NufftOpts()
    cases = [
        (3, 3, 1, 1, 1e-5, 0),    # type 3
        (1, 4, 1, 1, 1e-5, 0),    # dim 4
        (1, 3, 1, 0, 1e-5, 0),    # ntransf 0
        (1, 3, 2, 1, 1e-5, 0),    # sign 2
        (1, 3, 1, 1, 0.0, 0),     # tol 0
        (1, 3, 1, 1, 1e-5, 7),    # precision
    ]
    for typ, dim, sign, nt, tol, prec in cases:
        st = lib.hasty_nufft_makeplan(typ, dim, modes, sign, nt, tol, prec, ctypes.byref(o), None,
                                      ctypes.byref(h))
        assert st == 1, (typ, dim, sign, nt, tol, prec, st)
        assert _err(lib)


def test_null_handles_are_rejected(lib):
    assert lib.hasty_nufft_execute(None, None, None) == 1
    assert lib.hasty_nufft_setpts(None, 0, None, None, None) == 1
    assert lib.hasty_normal_apply(None, None, None) == 1
    assert "NULL" in _err(lib)


def test_toeplitz_validation(lib):
    from hastycucompute_amd._lib import i32_array, i64_array
    assert lib.hasty_toeplitz_scratch_bytes(3, i64_array([7, 8, 8]), 1) == -1   # 2n = 14 not smooth
    assert lib.hasty_toeplitz_scratch_bytes(3, i64_array([8, 8, 8]), 2) == 2 * (8 * 8 * 16 + 8 * 16 * 16) * 8
    bad_types = i32_array([0, 3, 0, 0, 0])
    dummy = ctypes.c_void_p(16)
    st = lib.hasty_toeplitz_multiplication(3, i64_array([8, 8, 8]), 1, dummy, dummy, dummy, None, 0,
                                           None, None, bad_types, 0, None)
    assert st == 1 and "ToeplitzMultType" in _err(lib)
    st = lib.hasty_toeplitz_multiplication(3, i64_array([8, 8, 8]), 1, dummy, dummy, dummy, None, 0,
                                           None, None, i32_array([0] * 5), 2, None)
    assert st == 1 and "Accumulate" in _err(lib)
    assert lib.hasty_transform_toeplitz_kernel(4, i64_array([8, 8, 8, 8]), dummy, None) == 1


def test_cg_validation(lib):
    assert lib.hasty_cg_step(None, None, None, None, 10, None, None) == 1
    assert lib.hasty_dot(None, None, -1, None, None) == 1
