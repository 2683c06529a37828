"""affine_pull / affine_push (src/pushpull_cpu.jl:268-362): resampling through a 4x4
(or 3x4) voxel-to-voxel affine instead of a deformation field. The sample coordinates
are computed inside the kernels from the matrix, so no deformation array is read.

``Aff`` is indexed like the Julia matrix (row, column) and maps 1-based output voxel
indices to 1-based source indices; ``adjust_affine`` turns it into the 0-based 3x4
matrix the kernels take. Output/input sizes are given in array order (Z, Y, X).
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .ops import _desc, _dims5, _f32, _hostptr, _is_dev, _ptr, _req, _stream
from .settings import Settings

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def adjust_affine(Aff) -> np.ndarray:
    """src/pushpull_cpu.jl:357-362: rows 1-3; column 4 becomes the row sum minus 1 (1-based
    to 0-based indices). Float32 throughout, row sums accumulated left to right."""
    A = np.asarray(Aff, dtype=np.float32)
    _req((A.shape[0] in (3, 4)) and A.shape[1] == 4, "Aff must be 3x4 or 4x4")
    A = A[:3, :].copy()
    for r in range(3):
        s = np.float32(0)
        for c in range(4):
            s = np.float32(s + A[r, c])
        A[r, 3] = np.float32(s - np.float32(1))
    return A


def _aff12(Aff):
    """The 12 floats of the adjusted matrix in Julia's column-major order."""
    return np.ascontiguousarray(adjust_affine(Aff).T.ravel())


def affine_pull(f0, Aff, d1, sett: Settings | None = None):
    """f0 (B, C, Z0, Y0, X0) -> (B, C) + d1."""
    sett = sett or Settings()
    f0 = _f32(f0, "f₀")
    lead, d0 = _dims5(f0.shape, 3, 5, "f₀")
    nc = f0.shape[-4] if f0.ndim >= 4 else 1
    nb = f0.shape[0] if f0.ndim == 5 else 1
    _req(len(sett.bnd) == nc or len(sett.bnd) == 1, "length(sett.bnd) must be 1 or the number of channels")
    d1 = tuple(int(v) for v in d1)
    a12 = _aff12(Aff)
    if _is_dev(f0):
        out = torch.empty(lead + d1, dtype=torch.float32, device=f0.device)
        d = _desc(sett, d0, d1, nc, nb)
        _lib.check(_lib.lib().pp_affine_pull(ct.byref(d), _ptr(out), _ptr(f0), _hostptr(a12), _stream(f0)),
                   "pp_affine_pull")
        return out
    out = np.empty(lead + d1, dtype=np.float32)
    n0, n1 = int(np.prod(d0)), int(np.prod(d1))
    dp = np.ascontiguousarray([g + 1 for g in sett.deg], dtype=np.uint64)
    dd0 = np.ascontiguousarray(d0[::-1], dtype=np.uint64)
    dd1 = np.ascontiguousarray(d1[::-1], dtype=np.uint64)
    for b in range(nb):
        for c in range(nc):
            bnd = np.ascontiguousarray(sett.bnd_for(c), dtype=np.int32)
            _lib.lib().pull_affine(_hostptr(out, n1 * (nc * b + c)), _hostptr(f0, n0 * (nc * b + c)), _hostptr(dd0),
                                   n1, _hostptr(dd1), _hostptr(a12), _hostptr(bnd), _hostptr(dp), int(sett.ext))
            st = _lib.lib().pp_last_status()
            if st != 0:
                raise _lib.PPError(f"pull_affine failed ({st}): {_lib.last_error()}")
    return out


def affine_push(f1, Aff, d0, sett: Settings | None = None):
    """f1 (B, C, Z1, Y1, X1) -> (B, C) + d0: the adjoint of affine_pull."""
    sett = sett or Settings()
    f1 = _f32(f1, "f₁")
    lead, d1 = _dims5(f1.shape, 3, 5, "f₁")
    nc = f1.shape[-4] if f1.ndim >= 4 else 1
    nb = f1.shape[0] if f1.ndim == 5 else 1
    _req(len(sett.bnd) == nc or len(sett.bnd) == 1, "length(sett.bnd) must be 1 or the number of channels")
    d0 = tuple(int(v) for v in d0)
    a12 = _aff12(Aff)
    if _is_dev(f1):
        out = torch.empty(lead + d0, dtype=torch.float32, device=f1.device)
        d = _desc(sett, d0, d1, nc, nb)
        _lib.check(_lib.lib().pp_affine_push(ct.byref(d), _ptr(out), _ptr(f1), _hostptr(a12), 0, _stream(f1)),
                   "pp_affine_push")
        return out
    out = np.zeros(lead + d0, dtype=np.float32)
    n0, n1 = int(np.prod(d0)), int(np.prod(d1))
    dp = np.ascontiguousarray([g + 1 for g in sett.deg], dtype=np.uint64)
    dd0 = np.ascontiguousarray(d0[::-1], dtype=np.uint64)
    dd1 = np.ascontiguousarray(d1[::-1], dtype=np.uint64)
    for b in range(nb):
        for c in range(nc):
            bnd = np.ascontiguousarray(sett.bnd_for(c), dtype=np.int32)
            _lib.lib().push_affine(_hostptr(out, n0 * (nc * b + c)), _hostptr(f1, n1 * (nc * b + c)), _hostptr(dd0),
                                   n1, _hostptr(dd1), _hostptr(a12), _hostptr(bnd), _hostptr(dp), int(sett.ext))
            st = _lib.lib().pp_last_status()
            if st != 0:
                raise _lib.PPError(f"push_affine failed ({st}): {_lib.last_error()}")
    return out
