"""Push with a count image, and the resampling steps of the multigrid callers
(src/multigrid.jl): the right-hand side of invert_def and restrict / prolong. The solver
around them (hessian_pyramid, relax!, vcycle!) is not part of this package.

``push_count`` is ``pushc`` (artifacts/C/pushpull.cu:128-142): the push of an image and the
push of ones through the same deformation, which the callers use as a normaliser. On
device tensors both come from one launch (``pp_push_count`` / ``pp_affine_push_count``);
numpy arrays go through the host entry points, one push at a time.

Shapes follow this package's order: images (B, C, Z, Y, X), deformations (B, 3, Z, Y, X),
sizes ``dim`` as (Z, Y, X).
"""
from __future__ import annotations

import ctypes as ct
import math

import numpy as np

from . import _lib
from .affine import _aff12, affine_pull, affine_push
from .ops import _desc, _f32, _hostptr, _is_dev, _phi_dims, _ptr, _req, _stream, push_raw
from .settings import Settings, id as _identity

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

SLIDING = [(2, 1, 1), (1, 2, 1), (1, 1, 2)]  # src/multigrid.jl:268 (and :49)


def _count_desc(sett, d0, d1, nc, nb, count_bnd):
    d = _desc(sett, d0, d1, nc, nb)
    if count_bnd is not None:
        full = [sett.bnd_for(c) for c in range(nc)] + [tuple(int(v) for v in count_bnd)]
        _req(len(full) <= _lib.PP_MAX_CHAN_BND, "too many boundary triples")
        d.nbnd = len(full)
        for c, t in enumerate(full):
            for q in range(3):
                d.bnd[3 * c + q] = int(t[q])
    return d


def _lead(shape_f1, nc, nb):
    return ((nb,) if len(shape_f1) == 5 else ()) + ((nc,) if len(shape_f1) >= 4 else ())


def push_count(f1, phi, dim, sett: Settings | None = None, count_bnd=None):
    """(push(f1, phi, dim, sett), push(ones, phi, dim)) -- the count uses ``count_bnd``
    (default: the first triple of ``sett``). ``f1=None`` pushes the identity of phi's grid
    (three channels, the 1-based index along each axis). Returns (f0, c0); c0 has one
    channel."""
    sett = sett or Settings()
    phi = _f32(phi, "ϕ")
    nb, d1 = _phi_dims(phi)
    dim = tuple(int(v) for v in dim)
    cb = tuple(int(v) for v in (count_bnd if count_bnd is not None else sett.bnd_for(0)))
    if f1 is None:
        nc = 3
        _req(len(sett.bnd) in (1, 3), "pushing the identity takes 1 or 3 boundary triples")
        shape_f1 = ((nb,) if phi.ndim == 5 else ()) + (3,) + d1
    else:
        f1 = _f32(f1, "f₁")
        _req(tuple(f1.shape[-3:]) == d1, "`f₁` and `ϕ` must have the same grid")
        nc = f1.shape[-4] if f1.ndim >= 4 else 1
        shape_f1 = tuple(f1.shape)
    lead = _lead(shape_f1, nc, nb)
    clead = lead[:-1] + (1,) if len(lead) and len(shape_f1) >= 4 else lead
    if _is_dev(phi):
        f0 = torch.empty(lead + dim, dtype=torch.float32, device=phi.device)
        c0 = torch.empty(clead + dim, dtype=torch.float32, device=phi.device)
        d = _count_desc(sett, dim, d1, nc, nb, cb)
        _lib.check(_lib.lib().pp_push_count(ct.byref(d), _ptr(f0), _ptr(c0), _ptr(phi),
                                            _ptr(f1) if f1 is not None else None, 0, _stream(phi)), "pp_push_count")
        return f0, c0
    src = f1
    if src is None:
        src = _identity(dim=d1[::-1])
        if phi.ndim == 5:
            src = np.broadcast_to(src, (nb, 3) + d1).copy()
    f0 = push_raw(src, phi, dim, sett)
    ones = np.ones(((nb,) if phi.ndim == 5 else ()) + d1, np.float32)
    c0 = push_raw(ones, phi, dim, Settings(sett.deg, cb, sett.ext))
    return f0, c0.reshape(clead + dim)


def affine_push_count(f1, Aff, dim, sett: Settings | None = None, count_bnd=None):
    """push_count through an affine (src/multigrid.jl:26-30): f1 (B, C, Z1, Y1, X1)."""
    sett = sett or Settings()
    f1 = _f32(f1, "f₁")
    dim = tuple(int(v) for v in dim)
    d1 = tuple(f1.shape[-3:])
    nc = f1.shape[-4] if f1.ndim >= 4 else 1
    nb = f1.shape[0] if f1.ndim == 5 else 1
    cb = tuple(int(v) for v in (count_bnd if count_bnd is not None else sett.bnd_for(0)))
    lead = _lead(tuple(f1.shape), nc, nb)
    clead = lead[:-1] + (1,) if f1.ndim >= 4 else lead
    if _is_dev(f1):
        f0 = torch.empty(lead + dim, dtype=torch.float32, device=f1.device)
        c0 = torch.empty(clead + dim, dtype=torch.float32, device=f1.device)
        d = _count_desc(sett, dim, d1, nc, nb, cb)
        _lib.check(_lib.lib().pp_affine_push_count(ct.byref(d), _ptr(f0), _ptr(c0), _ptr(f1), _hostptr(_aff12(Aff)),
                                                   0, _stream(f1)), "pp_affine_push_count")
        return f0, c0
    f0 = affine_push(f1, Aff, dim, sett)
    ones = np.ones(((nb,) if f1.ndim == 5 else ()) + d1, np.float32)
    c0 = affine_push(ones, Aff, dim, Settings(sett.deg, cb, sett.ext))
    return f0, c0.reshape(clead + dim)


def invert_def_rhs(phi):
    """The pushes of invert_def (src/multigrid.jl:266-275) for a deformation (3, Z, Y, X) or
    a batch (B, 3, Z, Y, X): g = push(Id, phi, d, Settings(1, [2 1 1; 1 2 1; 1 1 2], 0)),
    h = push(ones, phi, d, Settings(1, 1, 0)), g .-= h .* Id. Returns (g, h) -- the
    right-hand side and Hessian diagonal the multigrid solve starts from."""
    phi = _f32(phi, "ϕ")
    _, d = _phi_dims(phi)
    g, h = push_count(None, phi, d, Settings(1, SLIDING, 0), count_bnd=(1, 1, 1))
    Id = _identity(dim=d[::-1])
    if _is_dev(g):
        Id = torch.from_numpy(Id).to(g.device)
    if g.ndim == 5:
        Id = Id[None]
    g -= h * Id
    return g, h


def _zoom_affine(d_from, d_to):
    """Mat of restrict / prolong (src/multigrid.jl:23-25, 45-47): zm = d_to ./ d_from per
    Julia axis (X, Y, Z), os = (1 - zm) / 2, converted to Float32; sizes in (Z, Y, X)."""
    zm = [d_to[2 - q] / d_from[2 - q] for q in range(3)]
    os_ = [(1 - z) / 2 for z in zm]
    M = np.zeros((4, 4))
    for q in range(3):
        M[q, q] = zm[q]
        M[q, 3] = os_[q]
    M[3, 3] = 1
    return M.astype(np.float32)


def restrict(f1, d0=None, deg: int = 2):
    """restrict (src/multigrid.jl:21-35): push onto the coarser grid d0 (default ceil(d/2))
    and divide by the pushed count. f1 (…, Z1, Y1, X1)."""
    d1 = tuple(int(s) for s in f1.shape[-3:])
    d0 = tuple(int(v) for v in d0) if d0 is not None else tuple(int(math.ceil(s / 2)) for s in d1)
    f0, c0 = affine_push_count(f1, _zoom_affine(d1, d0), d0, Settings(deg, (1, 1, 1), True))
    f0 /= c0
    return f0


def prolong(f0, d1=None, deg: int = 2):
    """prolong (src/multigrid.jl:43-52): pull onto the finer grid d1 (default 2d) with the
    sliding boundary triples, channel c antisymmetric along axis c."""
    d0 = tuple(int(s) for s in f0.shape[-3:])
    d1 = tuple(int(v) for v in d1) if d1 is not None else tuple(2 * s for s in d0)
    return affine_pull(f0, _zoom_affine(d1, d0), d1, Settings(deg, SLIDING, True))
