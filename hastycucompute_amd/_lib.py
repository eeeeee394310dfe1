"""ctypes binding of ``libhasty_hip.so`` (the C-ABI declared in include/hasty_hip.h).

The library is loaded from this package directory only; there is no
fallback.  A missing or unloadable library raises ``RuntimeError`` so that no
caller can silently run another implementation.
"""
from __future__ import annotations

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhasty_hip.so")

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_i32 = ctypes.POINTER(ctypes.c_int32)
P_dbl = ctypes.POINTER(ctypes.c_double)
P_int = ctypes.POINTER(ctypes.c_int)

HASTY_OK = 0
STATUS_NAMES = {1: "invalid argument", 2: "allocation failed", 3: "HIP error", 4: "invalid state",
                5: "internal error"}
CG_STATE_DOUBLES = 1040


class HastyError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hasty: {STATUS_NAMES.get(code, 'error')} ({code}): {msg}")
        self.code = code


class NufftOpts(ctypes.Structure):
    _fields_ = [("upsamp", ctypes.c_double), ("method", ctypes.c_int32),
                ("tile", ctypes.c_int32 * 3), ("coil_group", ctypes.c_int32),
                ("max_batch", ctypes.c_int32), ("max_subproblem", ctypes.c_int32)]


# name -> (restype, argtypes)
SIGNATURES = {
    "hasty_last_error": (ctypes.c_char_p, []),
    "hasty_version": (c_int, []),
    "hasty_device_synchronize": (c_int, []),
    "hasty_profile_enable": (c_int, [c_int]),
    "hasty_profile_read": (c_int, [ctypes.c_char_p, c_i64, P_dbl, P_i32, c_i32, P_i32]),
    "hasty_nufft_makeplan": (c_int, [c_int, c_int, P_i64, c_int, c_int, c_dbl, c_int,
                                     ctypes.POINTER(NufftOpts), c_vp, ctypes.POINTER(c_vp)]),
    "hasty_nufft_setpts": (c_int, [c_vp, c_i64, c_vp, c_vp, c_vp]),
    "hasty_nufft_execute": (c_int, [c_vp, c_vp, c_vp]),
    "hasty_nufft_set_stream": (c_int, [c_vp, c_vp]),
    "hasty_nufft_coord_grad": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "hasty_nufft_destroy": (c_int, [c_vp]),
    "hasty_nufft_info": (c_int, [c_vp, P_int, P_dbl, P_i64, P_i32, P_i64]),
    "hasty_nufft_default_opts": (c_int, [ctypes.POINTER(NufftOpts)]),
    "hasty_nufft_kernel_width": (c_int, [c_dbl, c_dbl, c_int, c_int, P_int]),
    "hasty_toeplitz_kernel": (c_int, [c_int, P_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp]),
    "hasty_toeplitz_multiply": (c_int, [c_int, P_i64, c_i64, c_vp, c_vp, c_vp]),
    "hasty_transform_toeplitz_kernel": (c_int, [c_int, P_i64, c_vp, c_vp]),
    "hasty_toeplitz_multiplication": (c_int, [c_int, P_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64,
                                              c_vp, c_vp, P_i32, c_int, c_vp]),
    "hasty_toeplitz_scratch_bytes": (c_i64, [c_int, P_i64, c_i64]),
    "hasty_normal_create": (c_int, [c_int, P_i64, c_int, c_vp, c_vp, c_int, c_vp, ctypes.POINTER(c_vp)]),
    "hasty_normal_idt_create": (c_int, [c_int, P_i64, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_vp,
                                        ctypes.POINTER(c_vp)]),
    "hasty_normal_t2_create": (c_int, [c_int, P_i64, c_int, c_vp, c_int, c_vp, c_vp, c_vp, c_int, c_vp,
                                       ctypes.POINTER(c_vp)]),
    "hasty_normal_frames_create": (c_int, [c_int, P_i64, c_int, c_int, c_vp, c_vp, c_int, c_vp,
                                           ctypes.POINTER(c_vp)]),
    "hasty_sense_forward": (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "hasty_sense_adjoint": (c_int, [c_vp, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_int]),
    "hasty_normal_apply": (c_int, [c_vp, c_vp, c_vp]),
    "hasty_normal_set_stream": (c_int, [c_vp, c_vp]),
    "hasty_normal_destroy": (c_int, [c_vp]),
    "hasty_dot": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "hasty_dot_noalloc": (c_int, [c_vp, c_vp, c_i64, c_vp, c_vp]),
    "hasty_cg_init": (c_int, [c_vp, c_vp, c_i64, c_dbl, c_vp, c_vp]),
    "hasty_cg_step": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "hasty_cg_init_batched": (c_int, [c_vp, c_vp, c_i64, c_int, c_dbl, c_vp, c_vp]),
    "hasty_cg_step_batched": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_vp]),
    "hasty_cg_status": (c_int, [c_vp, P_dbl, P_int, c_vp]),
    "hasty_pcg_init": (c_int, [c_vp, c_vp, c_vp, c_i64, c_dbl, c_vp, c_vp]),
    "hasty_pcg_update": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp]),
    "hasty_pcg_direction": (c_int, [c_vp, c_vp, c_vp, c_i64, c_int, c_vp, c_vp]),
    "hasty_scale_by_device": (c_int, [c_vp, c_i64, c_vp, c_int, c_vp]),
    "hasty_norm": (c_int, [c_vp, c_i64, c_vp, c_vp]),
    "hasty_axpy": (c_int, [ctypes.c_double, ctypes.c_double, c_vp, c_vp, c_i64, c_vp]),
}

_lib = None
_lock = threading.Lock()


def lib():
    """Load (once) and return the C library with signatures applied."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m hastycucompute_amd.build` "
                    "(there is no CPU fallback)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(status: int) -> None:
    if status != HASTY_OK:
        msg = lib().hasty_last_error()
        raise HastyError(status, msg.decode() if msg else "")


def i64_array(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * len(vals))(*vals)


def i32_array(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int32 * len(vals))(*vals)


def profile_enable(on: bool = True) -> None:
    check(lib().hasty_profile_enable(1 if on else 0))


def profile_read(max_entries: int = 64) -> dict:
    """{kernel name: (total ms, launches)} of the launches recorded since the
    last read on this thread (synchronises on their events)."""
    buf = ctypes.create_string_buffer(8192)
    ms = (ctypes.c_double * max_entries)()
    cnt = (ctypes.c_int32 * max_entries)()
    n = ctypes.c_int32()
    check(lib().hasty_profile_read(buf, len(buf), ms, cnt, max_entries, ctypes.byref(n)))
    names = buf.value.decode().split("\n")
    return {names[i]: (ms[i], cnt[i]) for i in range(n.value)}
