"""CPU oracle of the sparse regulariser (TEST INFRASTRUCTURE ONLY).

Only tests/ and the ``cpu_baseline`` leg of bench.py import this module, as the checker.

* ``vel2mom`` / ``relax``: the restatement in sp_oracle.c (built into
  ``_build/libpp_oracle.so``) of artifacts/C/sparse_operator.c:45-135 and
  artifacts/C/sparse_operator_dev.cu:12-204, bit-exact FP32.
* ``ref_vel2mom`` / ``ref_relax``: the reference's own artifacts/C/sparse_operator.c,
  compiled from source by ``make ref`` into ``_ref/libsparse_ref.so`` (present only where
  that tree was available).

Arrays: velocity fields are C-ordered ``(C, Z, Y, X)`` float32 (Julia ``(X, Y, Z, C)``);
the kernel is the package's ``SparseKernel`` (any object with the same attributes);
``bnd`` is ``(C, 3)`` with the (x, y, z) codes of each channel.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

import oracle as _o

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libsparse_ref.so")
_P = ct.c_void_p
_ref = None
_bound = False


def _lib():
    global _bound
    L = _o.lib()
    if not _bound:
        L.spo_vel2mom_export.argtypes = [_P] * 10 + [ct.c_int]
        L.spo_vel2mom_export.restype = None
        L.spo_relax_export.argtypes = [_P, _P, _P, ct.c_int, _P, ct.c_int, _P, _P, _P, _P, _P, _P, ct.c_int]
        L.spo_relax_export.restype = None
        _bound = True
    return L


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_SO}")
        L = ct.CDLL(REF_SO)
        L.vel2mom.argtypes = [_P] * 10
        L.vel2mom.restype = None
        L.relax.argtypes = [_P] * 11
        L.relax.restype = None
        _ref = L
    return _ref


def _p(a):
    return a.ctypes.data_as(_P)


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a


def _common(kernel, bnd):
    dims = np.array([int(s) for s in kernel.d[::-1]], dtype=np.uint64)
    dp = np.array([int(s) for s in kernel.stride[::-1]], dtype=np.int32)
    b = np.ascontiguousarray(np.asarray(bnd, dtype=np.int32).reshape(-1))
    return dims, dp, b


def vel2mom(v, kernel, bnd, u=None, accum_edges=False):
    """u = L v (the restatement). With ``accum_edges`` the voxels outside the interior block
    add to the given u, as the reference's vel2mom does."""
    v = _f32(v)
    out = np.zeros_like(v) if u is None else _f32(u).copy()
    dims, dp, b = _common(kernel, bnd)
    _lib().spo_vel2mom_export(_p(out), _p(v), _p(dims), int(kernel.nchan), _p(kernel.offset), _p(kernel.length),
                              _p(kernel.values), _p(kernel.patch_indices), _p(dp), _p(b), int(bool(accum_edges)))
    return out


def _hcode(h, nc):
    if h is None or np.asarray(h).size == 0:
        return 0, None
    h = _f32(h)
    ch = 1 if h.ndim == 3 else h.shape[0]
    return (1 if ch == nc else 2), h


def relax(v, g, h, kernel, bnd, nit):
    """``nit`` sweeps of the restatement; returns the updated copy of v."""
    out = _f32(v).copy()
    g = _f32(g)
    hc, h = _hcode(h, kernel.nchan)
    dims, dp, b = _common(kernel, bnd)
    _lib().spo_relax_export(_p(out), _p(g), _p(h) if hc else None, hc, _p(dims), int(kernel.nchan),
                            _p(kernel.offset), _p(kernel.length), _p(kernel.values), _p(kernel.patch_indices), _p(dp),
                            _p(b), int(nit))
    return out


def _ref_d(kernel, hcode):
    return np.array([*[int(s) for s in kernel.d[::-1]], int(kernel.nchan), hcode], dtype=np.uint64)


def ref_vel2mom(v, kernel, bnd, u=None):
    """The reference's vel2mom (sparse_operator.c:45): edges add to u, the interior is stored."""
    v = _f32(v)
    out = np.zeros_like(v) if u is None else _f32(u).copy()
    dp = np.array([int(s) for s in kernel.stride[::-1]], dtype=np.uint64)
    b = np.ascontiguousarray(np.asarray(bnd, dtype=np.int32).reshape(-1))
    ref().vel2mom(_p(out), _p(v), _p(_ref_d(kernel, 1)), _p(kernel.offset), _p(kernel.length), _p(kernel.values),
                  _p(kernel.indices), _p(kernel.patch_indices), _p(dp), _p(b))
    return out


def ref_relax(v, g, h, kernel, bnd, nit):
    """``nit`` calls of the reference's relax (one sweep each, as src/operator_sparse.jl:50-60)."""
    out = _f32(v).copy()
    g = _f32(g)
    hc, h = _hcode(h, kernel.nchan)
    dp = np.array([int(s) for s in kernel.stride[::-1]], dtype=np.uint64)
    b = np.ascontiguousarray(np.asarray(bnd, dtype=np.int32).reshape(-1))
    hp = _p(h) if hc else _p(np.zeros(1, dtype=np.float32))
    for _ in range(int(nit)):
        ref().relax(_p(out), _p(_ref_d(kernel, hc)), _p(g), hp, _p(kernel.offset), _p(kernel.length),
                    _p(kernel.values), _p(kernel.indices), _p(kernel.patch_indices), _p(dp), _p(b))
    return out
