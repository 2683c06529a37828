"""compose, compose! and Exp (src/compositions.jl:1-49).

``compose(phi, theta)``: ``psi = phi o theta`` for a displacement ``phi`` (three channels)
sampled at the 1-based deformation ``theta``: channel c is pulled with degree 1,
extrapolation and boundary 1 on every axis except 2 on axis c (the "sliding" matrix
``[2 1 1; 1 2 1; 1 1 2]`` of compositions.jl:33). ``psi`` has ``theta``'s size.

``Exp(v, nsq=8)``: scaling and squaring as the reference writes it (compositions.jl:41-49):
``u = v / 2^nsq``, then ``nsq`` times ``u = compose(u, Id + u)``, then ``u + Id``. (The
reference's update keeps only ``u o (Id + u)``, not ``u + u o (Id + u)``; this package
reproduces it as written.)

Arrays are in this package's order: a displacement or deformation is ``(3, Z, Y, X)`` or
``(B, 3, Z, Y, X)`` (Julia ``(X, Y, Z, 3[, B])``), component 0 along X. Device tensors run
through ``pp_compose`` / ``pp_exp`` (all three channels in one launch per composition);
numpy arrays through the host entry points, one pull per channel as the reference does.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .ops import _f32, _is_dev, _ptr, _req, _stream, pull_raw
from .settings import Settings, id as _identity

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

SLIDING = [(2, 1, 1), (1, 2, 1), (1, 1, 2)]  # compositions.jl:33, column c -> channel c


def _field_dims(a, name):
    _req(a.ndim in (4, 5), f"`{name}` must have 4 or 5 dimensions")
    _req(a.shape[-4] == 3, f"`{name}` must have three channels")
    nb = a.shape[0] if a.ndim == 5 else 1
    return nb, tuple(int(s) for s in a.shape[-3:])


def _xyz(d):
    return (ct.c_longlong * 3)(*(int(s) for s in d[::-1]))


def compose_(psi, phi, theta):
    """compose! (compositions.jl:27-39): writes phi o theta into psi (same size as theta)."""
    phi = _f32(phi, "ϕ")
    theta = _f32(theta, "θ")
    nb, d0 = _field_dims(phi, "ϕ")
    nb_t, d1 = _field_dims(theta, "θ")
    _req(nb == nb_t, "`ϕ` & `θ` must have the same batch size")
    _req(tuple(psi.shape) == tuple(theta.shape), "`ψ` and `θ` must have the same dimensions")
    if _is_dev(phi):
        _req(_is_dev(psi) and psi.is_contiguous() and psi.dtype == torch.float32, "`ψ` must be a contiguous device Float32")
        _lib.check(_lib.lib().pp_compose(_ptr(psi), _ptr(phi), _ptr(theta), _xyz(d0), _xyz(d1), nb, _stream(phi)),
                   "pp_compose")
        return psi
    psi[...] = pull_raw(phi, theta, Settings(1, SLIDING, True))
    return psi


def compose(phi, theta):
    """compose (compositions.jl:9-15)."""
    if _is_dev(phi):
        psi = torch.empty(theta.shape, dtype=torch.float32, device=phi.device)
    else:
        psi = np.empty(np.shape(theta), dtype=np.float32)
    return compose_(psi, phi, theta)


def Exp(v, nsq: int = 8):  # noqa: N802 - Julia name
    """Exp (compositions.jl:41-49): the 1-based deformation of velocity v (3, Z, Y, X) or
    (B, 3, Z, Y, X); nsq squarings (the reference uses 8)."""
    v = _f32(v, "v")
    nb, d = _field_dims(v, "v")
    _req(1 <= int(nsq) <= 30, "nsq must be in 1..30")
    if _is_dev(v):
        out = torch.empty_like(v)
        work = torch.empty_like(v)
        _lib.check(_lib.lib().pp_exp(_ptr(out), _ptr(v), _ptr(work), _xyz(d), nb, int(nsq), _stream(v)), "pp_exp")
        return out
    Id = _identity(dim=d[::-1])
    if v.ndim == 5:
        Id = Id[None]
    u = (v / np.float32(2.0 ** nsq)).astype(np.float32)
    for _ in range(int(nsq)):
        u = compose(u, (Id + u).astype(np.float32))
    return (u + Id).astype(np.float32)
