"""Settings and the identity grid, mirroring src/pushpull_general.jl.

``Settings(deg, bnd, ext)`` normalises exactly as the Julia constructor
(src/pushpull_general.jl:15-60):

* ``deg``: an integer or a 3-tuple, each clamped to 0..4 (``:27-30``);
* ``bnd``: an integer, a 3-tuple, a list of 3-tuples (one per channel) or a 3 x C
  matrix; every code must be 0 (circulant), 1 (reflect) or 2 (antisymmetric), else an
  ``AssertionError`` as Julia's ``@assert`` raises (``:33-54``);
* ``ext``: clamped to 0/1 (``:57``).

Defaults are ``deg=(1,1,1)``, ``bnd=[(1,1,1)]``, ``ext=1`` (``:19-24``).
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

_ALLOWED = (0, 1, 2)


def _require(cond: bool, msg: str = "AssertionError") -> None:
    if not cond:
        raise AssertionError(msg)


class Settings:
    __slots__ = ("deg", "bnd", "ext")

    def __init__(self, deg=(1, 1, 1), bnd=None, ext=True):
        if isinstance(deg, (int, np.integer)):
            deg = (deg, deg, deg)
        deg = tuple(int(d) for d in deg)
        _require(len(deg) == 3, "deg must be an integer or a 3-tuple")
        self.deg = tuple(max(min(d, 4), 0) for d in deg)

        if bnd is None:
            bnd1 = [(1, 1, 1)]
        elif isinstance(bnd, (int, np.integer)):
            _require(int(bnd) in _ALLOWED)
            b = int(bnd)
            bnd1 = [(b, b, b)]
        elif isinstance(bnd, np.ndarray) and bnd.ndim == 2:
            _require(bnd.shape[0] == 3)
            _require(all(int(v) in _ALLOWED for v in bnd.ravel()))
            bnd1 = [tuple(int(v) for v in bnd[:, i]) for i in range(bnd.shape[1])]
        elif isinstance(bnd, list):
            bnd1 = []
            for t in bnd:
                t = tuple(int(v) for v in t)
                _require(len(t) == 3 and all(v in _ALLOWED for v in t))
                bnd1.append(t)
        else:  # a 3-tuple
            t = tuple(int(v) for v in bnd)
            _require(len(t) == 3 and all(v in _ALLOWED for v in t))
            bnd1 = [t]
        self.bnd = bnd1
        self.ext = int(max(min(int(ext), 1), 0))

    def __repr__(self):
        return f"Settings(deg={self.deg}, bnd={self.bnd}, ext={self.ext})"

    def bnd_for(self, c: int):
        """Boundary triple of channel c (src/pushpull_cpu.jl:43)."""
        return self.bnd[0] if len(self.bnd) == 1 else self.bnd[c]


def id(v=None, affine=None, dim: Sequence[int] | None = None, like=None):  # noqa: A001 - Julia name
    """Identity deformation, 1-based (src/pushpull_general.jl:81-99).

    ``dim`` is the Julia volume size ``(X, Y, Z)``; it defaults to the last three axes
    of ``v`` reversed (``v`` in this package's (..., Z, Y, X) convention). Returns an
    array of shape ``(3, Z, Y, X)`` (Julia ``(X, Y, Z, 3)``): numpy by default, or a
    torch tensor on ``v``'s device when ``v`` is a tensor.
    """
    if dim is None:
        dim = tuple(int(s) for s in v.shape[-3:][::-1])
    X, Y, Z = (int(s) for s in dim)
    x = np.arange(1, X + 1, dtype=np.float64)[None, None, :]
    y = np.arange(1, Y + 1, dtype=np.float64)[None, :, None]
    z = np.arange(1, Z + 1, dtype=np.float64)[:, None, None]
    phi = np.empty((3, Z, Y, X), dtype=np.float32)
    if affine is None:
        phi[0] = x
        phi[1] = y
        phi[2] = z
    else:
        A = np.asarray(affine, dtype=np.float64)
        for r in range(3):
            phi[r] = (A[r, 0] * x + A[r, 1] * y + A[r, 2] * z + A[r, 3]).astype(np.float32)
    try:
        import torch

        if isinstance(v, torch.Tensor):
            return torch.from_numpy(phi).to(v.device)
    except ImportError:  # pragma: no cover
        pass
    return phi
