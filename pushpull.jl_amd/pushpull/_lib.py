"""ctypes binding of libpushpull_hip.so (the C ABI declared in include/pushpull_hip.h).

The library is the only compute path: if it is missing, loading raises. There is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(os.path.dirname(_PKG), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libpushpull_hip.so")
HEADER = os.path.join(os.path.dirname(os.path.dirname(_PKG)), "include", "pushpull_hip.h")

PP_MAX_CHAN_BND = 64
PP_OK, PP_EINVAL, PP_EDEVICE, PP_ENOMEM, PP_ERANGE = 0, -1, -2, -3, -4
PP_OPT_TILING = 1
PP_OPT_TILE_SHAPE = 2
PP_OPT_USE_TUNED = 3
PP_OPT_BOX_CAP = 4


class PPDesc(ct.Structure):
    """Mirror of ``pp_desc`` (include/pushpull_hip.h)."""

    _fields_ = [
        ("d0", ct.c_int64 * 3),
        ("d1", ct.c_int64 * 3),
        ("nchan", ct.c_int32),
        ("nbatch", ct.c_int32),
        ("deg", ct.c_int32 * 3),
        ("ext", ct.c_int32),
        ("nbnd", ct.c_int32),
        ("bnd", ct.c_int32 * (3 * PP_MAX_CHAN_BND)),
    ]


class PPError(RuntimeError):
    pass


_P = ct.c_void_p
_U = ct.c_ulonglong
_TIER_A = ["pull", "pullg", "pullh", "push", "pushg"]

# name -> (argtypes, restype)
_SIGS = {
    **{n: ([_P, _P, _P, _P, _U, _P, _P, ct.c_int], None) for n in _TIER_A},
    "pull_affine": ([_P, _P, _P, _U, _P, _P, _P, _P, ct.c_int], None),
    "push_affine": ([_P, _P, _P, _U, _P, _P, _P, _P, ct.c_int], None),
    "pp_affine_pull": ([_P, _P, _P, _P, _P], ct.c_int),
    "pp_affine_push": ([_P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_pull": ([_P, _P, _P, _P, _P], ct.c_int),
    "pp_pull_grad": ([_P, _P, _P, _P, _P], ct.c_int),
    "pp_pull_hess": ([_P, _P, _P, _P, _P], ct.c_int),
    "pp_push": ([_P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_push_grad": ([_P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_pull_backward": ([_P, _P, _P, _P, _P, _P], ct.c_int),
    "pp_pull_grad_backward": ([_P, _P, _P, _P, _P, _P], ct.c_int),
    "pp_objective_grad": ([_P, _P, _P, _P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_objective_grad_step": ([_P, _P, _P, _P, _P, _P, _P, _P, ct.c_float, ct.c_int, _P], ct.c_int),
    "pp_pull_vjp": ([_P, _P, _P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_push_count": ([_P, _P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_affine_push_count": ([_P, _P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_compose": ([_P, _P, _P, ct.POINTER(ct.c_longlong), ct.POINTER(ct.c_longlong), ct.c_int, _P], ct.c_int),
    "pp_exp": ([_P, _P, _P, ct.POINTER(ct.c_longlong), ct.c_int, ct.c_int, _P], ct.c_int),
    "vel2mom": ([_P] * 10, None),
    "relax": ([_P] * 11, None),
    "pp_sparse_create": ([ct.POINTER(ct.c_void_p), ct.POINTER(ct.c_int64), ct.c_int, ct.POINTER(ct.c_int32), _P, _P, _P,
                          _P, ct.c_int64], ct.c_int),
    "pp_sparse_destroy": ([_P], ct.c_int),
    "pp_vel2mom": ([_P, _P, _P, _P, ct.c_int, _P], ct.c_int),
    "pp_relax": ([_P, _P, _P, _P, ct.c_int, _P, ct.c_int, _P], ct.c_int),
    "pp_malloc": ([ct.POINTER(ct.c_void_p), ct.c_size_t], ct.c_int),
    "pp_free": ([_P], ct.c_int),
    "pp_memcpy_h2d": ([_P, _P, ct.c_size_t, _P], ct.c_int),
    "pp_memcpy_d2h": ([_P, _P, ct.c_size_t, _P], ct.c_int),
    "pp_memset_zero": ([_P, ct.c_size_t, _P], ct.c_int),
    "pp_axpy": ([_P, _P, ct.c_float, ct.c_size_t, _P], ct.c_int),
    "pp_lds_atomic_calibrate": ([_P, ct.c_int, ct.c_int, _P], ct.c_int),
    "pp_stream_sync": ([_P], ct.c_int),
    "pp_device_sync": ([], ct.c_int),
    "pp_set_option": ([ct.c_int, ct.c_int], ct.c_int),
    "pp_get_option": ([ct.c_int], ct.c_int),
    "pp_tune": ([_P, _P, _P, _P, _P, _P, _P, _P, _P, ct.c_int, ct.POINTER(ct.c_int)], ct.c_int),
    "pp_tune_results": ([ct.c_int, _P, _P, _P], ct.c_int),
    "pp_tune_save": ([ct.c_char_p], ct.c_int),
    "pp_tune_load": ([ct.c_char_p], ct.c_int),
    "pp_tune_clear": ([], ct.c_int),
    "pp_last_error": ([], ct.c_char_p),
    "pp_last_status": ([], ct.c_int),
    "pp_version": ([], ct.c_char_p),
}

_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PPError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the pull/push path has no CPU fallback)"
            )
        L = ct.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().pp_last_error().decode(errors="replace")
        raise PPError(f"{what} failed ({status}): {msg}")


def last_error() -> str:
    return lib().pp_last_error().decode(errors="replace")


def declared_functions(header: str = HEADER) -> list[str]:
    """Function names declared in include/pushpull_hip.h (for the export check)."""
    import re

    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|char)\s*\*?\s*(\w+)\s*\(", txt, flags=re.M)
    return names


def set_option(key: int, value: int) -> None:
    check(lib().pp_set_option(key, value), "pp_set_option")


def get_option(key: int) -> int:
    return int(lib().pp_get_option(key))
