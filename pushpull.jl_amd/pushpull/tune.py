"""Autotuner front end (``pp_tune``): pick the fastest kernel choice for an op, degree
and volume-size class by timing every candidate on the caller's own device buffers.

Variants: 0 = one work-item per sample; 1, 2, 3 = LDS-tiled with 16x8x8, 32x8x4, 16x16x4
tiles of the deformation grid, each at the op's default LDS box budget and at the budgets
that let 6, 5, 4 or 3 workgroups share a CU. Choices live in the library (per process)
and can be saved to / loaded from a small text table.
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import _lib, ops
from .settings import Settings

OPS = {"pull": 0, "pull_grad": 1, "push": 3, "push_grad": 4, "pull_backward": 5, "objective": 7, "pull_vjp": 8}
VARIANTS = ("per-sample", "tiled 16x8x8", "tiled 32x8x4", "tiled 16x16x4")


def _p(t):
    return ct.c_void_p(t.data_ptr()) if t is not None else ct.c_void_p(0)


def tune(op: str, f, phi, sett: Settings | None = None, aux=None, dim=None, reps: int = 3) -> int:
    """Time the variants of ``op`` on these device tensors; returns the chosen variant."""
    sett = sett or Settings()
    f = f.contiguous()
    phi = phi.contiguous()
    nb = phi.shape[0] if phi.ndim == 5 else 1
    d1 = tuple(phi.shape[-3:])
    if op in ("push", "push_grad"):
        C = f.shape[-5] if (op == "push_grad" and f.ndim >= 5) else (f.shape[-4] if f.ndim >= 4 and op == "push" else 1)
        d0 = tuple(dim)
    else:
        C = f.shape[-4] if f.ndim >= 4 else 1
        d0 = tuple(f.shape[-3:])
    desc = ops._desc(sett, d0, d1, C, nb)
    dev = f.device
    out2 = loss = None
    if op == "pull":
        out = torch.empty((nb, C) + d1, device=dev)
    elif op == "pull_grad":
        out = torch.empty((nb, C, 3) + d1, device=dev)
    elif op in ("push", "push_grad"):
        out = torch.empty((nb, C) + d0, device=dev)
    else:
        out = torch.empty((nb, 3) + d1, device=dev)
        if op in ("objective", "pull_vjp"):
            out2 = torch.empty((nb, C) + d0, device=dev)
        if op == "objective":
            loss = torch.empty((), dtype=torch.float64, device=dev)
    v = ct.c_int(-1)
    s = ct.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    L = _lib.lib()
    _lib.check(L.pp_tune(OPS[op], ct.byref(desc), _p(out), _p(out2), _p(loss), _p(phi), _p(f),
                         _p(aux.contiguous() if aux is not None else None), s, int(reps), ct.byref(v)), "pp_tune")
    return int(v.value)


def results() -> list[dict]:
    """The candidates of this thread's last ``tune``: variant, box budget (0: the op's
    default) and mean ms per call, in the order timed."""
    L = _lib.lib()
    n = L.pp_tune_results(0, None, None, None)
    v, c, m = (ct.c_int * max(n, 1))(), (ct.c_int * max(n, 1))(), (ct.c_float * max(n, 1))()
    L.pp_tune_results(n, v, c, m)
    return [dict(variant=int(v[i]), name=VARIANTS[v[i]], cap=int(c[i]), ms=float(m[i])) for i in range(n)]


def save(path: str) -> None:
    _lib.check(_lib.lib().pp_tune_save(path.encode()), "pp_tune_save")


def load(path: str) -> None:
    _lib.check(_lib.lib().pp_tune_load(path.encode()), "pp_tune_load")


def clear() -> None:
    _lib.check(_lib.lib().pp_tune_clear(), "pp_tune_clear")
