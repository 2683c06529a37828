"""pull / push / pull_grad / push_grad / pull_hess over the HIP C ABI.

Mirrors the Julia API of PushPull.jl (src/pushpull_cpu.jl for host arrays,
ext/pushpull_gpu.jl for device arrays), with the same argument order, shape rules and
``AssertionError`` on a violated precondition.

Array convention: the memory of every array is the Julia array's; shapes are written
in reverse (C order). Julia ``f0 :: (X, Y, Z, C, B)`` is ``(B, C, Z, Y, X)`` here;
``phi :: (X, Y, Z, 3, B)`` is ``(B, 3, Z, Y, X)`` with component 0 the X coordinate,
1-based. ``push``'s output size is given in array order ``(Z, Y, X)``.

Dispatch:

* numpy arrays (and CPU torch tensors) go through Tier A, the symbol-compatible
  ``pull/pullg/pullh/push/pushg`` entry points, one (batch, channel) slice per call,
  exactly like the ``ccall`` loop of src/pushpull_cpu.jl:42-51 (those entry points stage
  the slice through GPU memory);
* torch tensors on the GPU go through Tier B (``pp_pull`` ...), one stream-ordered launch
  for all channels and batches.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from .settings import Settings

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None


def _req(cond, msg):
    if not cond:
        raise AssertionError(msg)


def _is_dev(a) -> bool:
    return torch is not None and isinstance(a, torch.Tensor) and a.is_cuda


def _stream(t):
    return ct.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ptr(t):
    return ct.c_void_p(t.data_ptr())


def _f32(a, name):
    if _is_dev(a):
        if a.dtype != torch.float32:
            raise TypeError(f"`{name}` must be Float32")
        return a.contiguous()
    if torch is not None and isinstance(a, torch.Tensor):
        a = a.detach().numpy()
    a = np.asarray(a)
    if a.dtype != np.float32:
        raise TypeError(f"`{name}` must be Float32")
    return np.ascontiguousarray(a)


def _desc(sett: Settings, d0, d1, nchan, nbatch) -> _lib.PPDesc:
    """d0, d1 in array order (Z, Y, X)."""
    d = _lib.PPDesc()
    for a in range(3):
        d.d0[a] = int(d0[2 - a])
        d.d1[a] = int(d1[2 - a])
        d.deg[a] = int(sett.deg[a])
    d.nchan = int(nchan)
    d.nbatch = int(nbatch)
    d.ext = int(sett.ext)
    if len(sett.bnd) > _lib.PP_MAX_CHAN_BND:
        raise ValueError(f"at most {_lib.PP_MAX_CHAN_BND} per-channel boundary triples")
    d.nbnd = len(sett.bnd)
    for c, t in enumerate(sett.bnd):
        for q in range(3):
            d.bnd[3 * c + q] = int(t[q])
    return d


def _dims5(shape, lo, hi, name):
    """Split a Julia-style image shape (reversed) into (B, C, spatial)."""
    _req(lo <= len(shape) <= hi, f"`{name}` must have between {lo} & {hi} dimensions")
    sp = tuple(shape[-3:])
    lead = tuple(shape[:-3])
    return lead, sp


def _phi_dims(phi):
    _req(4 <= phi.ndim <= 5, "`ϕ` must have 4 or 5 dimensions")
    _req(phi.shape[-4] == 3, "`ϕ` must have three channels")
    nb = phi.shape[0] if phi.ndim == 5 else 1
    return nb, tuple(phi.shape[-3:])


def _empty(shape, like):
    if _is_dev(like):
        return torch.empty(shape, dtype=torch.float32, device=like.device)
    return np.empty(shape, dtype=np.float32)


def _zeros(shape, like):
    if _is_dev(like):
        return torch.zeros(shape, dtype=torch.float32, device=like.device)
    return np.zeros(shape, dtype=np.float32)


def _hostptr(a: np.ndarray, off_elems: int = 0):
    return ct.c_void_p(a.ctypes.data + 4 * off_elems)


# ------------------------------------------------------------------------------------
def _pull_family(op, ncomp, f0, phi, sett):
    f0 = _f32(f0, "f₀")
    phi = _f32(phi, "ϕ")
    lead, d0 = _dims5(f0.shape, 3, 5, "f₀")
    nb, d1 = _phi_dims(phi)
    nb_f = f0.shape[0] if f0.ndim == 5 else 1
    _req(nb == nb_f, "`ϕ` & `f₀` must have the same batch size")
    nc = f0.shape[-4] if f0.ndim >= 4 else 1
    _req(len(sett.bnd) == nc or len(sett.bnd) == 1, "length(sett.bnd) must be 1 or the number of channels")
    comp = () if ncomp == 1 else ((3,) if ncomp == 3 else (3, 3))
    out = _empty(lead + comp + d1, f0)
    n0, n1 = int(np.prod(d0)), int(np.prod(d1))
    if _is_dev(f0):
        _req(phi.device == f0.device, "`ϕ` & `f₀` must be on the same device")
        d = _desc(sett, d0, d1, nc, nb)
        fn = {"pull": _lib.lib().pp_pull, "pullg": _lib.lib().pp_pull_grad, "pullh": _lib.lib().pp_pull_hess}[op]
        _lib.check(fn(ct.byref(d), _ptr(out), _ptr(phi), _ptr(f0), _stream(f0)), f"pp_{op}")
        return out
    # Tier A: one ccall per (batch, channel) slice, src/pushpull_cpu.jl:42-51
    fn = getattr(_lib.lib(), op)
    dp = np.ascontiguousarray([d + 1 for d in sett.deg], dtype=np.uint64)
    dd0 = np.ascontiguousarray(d0[::-1], dtype=np.uint64)
    for b in range(nb):
        for c in range(nc):
            bnd = np.ascontiguousarray(sett.bnd_for(c), dtype=np.int32)
            fn(_hostptr(out, ncomp * n1 * (nc * b + c)), _hostptr(phi, 3 * n1 * b), _hostptr(f0, n0 * (nc * b + c)),
               _hostptr(dd0), n1, _hostptr(bnd), _hostptr(dp), int(sett.ext))
            st = _lib.lib().pp_last_status()
            if st != 0:
                raise _lib.PPError(f"{op} failed ({st}): {_lib.last_error()}")
    return out


def _push_family(op, ncomp, f1, phi, dim, sett, accumulate_into=None):
    f1 = _f32(f1, "f₁")
    phi = _f32(phi, "ϕ")
    if ncomp == 1:
        lead, sp = _dims5(f1.shape, 3, 5, "f₁")
        nb_f = f1.shape[0] if f1.ndim == 5 else 1
        nc = f1.shape[-4] if f1.ndim >= 4 else 1
    else:
        _req(4 <= f1.ndim <= 6, "`∇f` must have between 4 & 6 dimensions")
        _req(f1.shape[-4] == 3, "`∇f` must have three components")
        lead, sp = tuple(f1.shape[:-4]), tuple(f1.shape[-3:])
        nb_f = f1.shape[0] if f1.ndim == 6 else 1
        nc = f1.shape[-5] if f1.ndim >= 5 else 1
    nb, d1 = _phi_dims(phi)
    _req(nb == nb_f, "`ϕ` & `f₀` must have the same batch size")
    _req(sp == d1, "`f₁` and `ϕ` must have the same volume dimensions")
    _req(len(sett.bnd) == nc or len(sett.bnd) == 1, "length(sett.bnd) must be 1 or the number of channels")
    d0 = tuple(int(v) for v in dim)
    _req(len(d0) == 3, "dim must have three entries")
    n0, n1 = int(np.prod(d0)), int(np.prod(d1))
    if _is_dev(f1):
        _req(phi.device == f1.device, "`ϕ` & `f₁` must be on the same device")
        out = accumulate_into if accumulate_into is not None else _empty(lead + d0, f1)
        d = _desc(sett, d0, d1, nc, nb)
        fn = {"push": _lib.lib().pp_push, "pushg": _lib.lib().pp_push_grad}[op]
        _lib.check(fn(ct.byref(d), _ptr(out), _ptr(phi), _ptr(f1), int(accumulate_into is not None), _stream(f1)),
                   f"pp_{op}")
        return out
    out = _zeros(lead + d0, f1)
    fn = getattr(_lib.lib(), op)
    dp = np.ascontiguousarray([d + 1 for d in sett.deg], dtype=np.uint64)
    dd0 = np.ascontiguousarray(d0[::-1], dtype=np.uint64)
    for b in range(nb):
        for c in range(nc):
            bnd = np.ascontiguousarray(sett.bnd_for(c), dtype=np.int32)
            fn(_hostptr(out, n0 * (nc * b + c)), _hostptr(phi, 3 * n1 * b), _hostptr(f1, ncomp * n1 * (nc * b + c)),
               _hostptr(dd0), n1, _hostptr(bnd), _hostptr(dp), int(sett.ext))
            st = _lib.lib().pp_last_status()
            if st != 0:
                raise _lib.PPError(f"{op} failed ({st}): {_lib.last_error()}")
    return out


def pull_raw(f0, phi, sett=None):
    return _pull_family("pull", 1, f0, phi, sett or Settings())


def pull_grad_raw(f0, phi, sett=None):
    return _pull_family("pullg", 3, f0, phi, sett or Settings())


def pull_hess(f0, phi, sett=None):
    """pull_hess (src/pushpull_cpu.jl:111-148): out (B, C, 3, 3, Z1, Y1, X1)."""
    return _pull_family("pullh", 9, f0, phi, sett or Settings())


def push_raw(f1, phi, dim, sett=None):
    return _push_family("push", 1, f1, phi, dim, sett or Settings())


def push_grad(g1, phi, dim, sett=None):
    """push_grad (src/pushpull_cpu.jl:220-259): the adjoint of pull_grad."""
    return _push_family("pushg", 3, g1, phi, dim, sett or Settings())


def _same_device(*named):
    dev = named[0][1].device
    for name, t in named[1:]:
        _req(t.device == dev, f"`{name}` must be on the device of `{named[0][0]}` ({dev}), not {t.device}")


def _out_arg(t, shape, like, name):
    """A caller-supplied output: float32, contiguous, the expected shape, on like's device."""
    _req(isinstance(t, torch.Tensor) and t.dtype == torch.float32, f"`{name}` must be a Float32 tensor")
    _req(tuple(t.shape) == tuple(shape), f"`{name}` must be {tuple(shape)}, got {tuple(t.shape)}")
    _req(t.is_contiguous(), f"`{name}` must be contiguous (it is written in place)")
    _same_device(("f₀", like), (name, t))
    return t


def _backward(kind, f0, phi, delta, sett):
    """Device-only channel contractions used by the pullbacks."""
    f0 = _f32(f0, "f₀")
    phi = _f32(phi, "ϕ")
    delta = _f32(delta, "δ")
    _req(_is_dev(f0) and _is_dev(phi) and _is_dev(delta), "pullback kernels take device tensors")
    _same_device(("f₀", f0), ("ϕ", phi), ("δ", delta))
    lead, d0 = _dims5(f0.shape, 3, 5, "f₀")
    nb, d1 = _phi_dims(phi)
    _req(nb == (f0.shape[0] if f0.ndim == 5 else 1), "`ϕ` & `f₀` must have the same batch size")
    nc = f0.shape[-4] if f0.ndim >= 4 else 1
    want = lead + d1 if kind == 1 else lead + (3,) + d1
    _req(tuple(delta.shape) == want,
         f"δ must have the shape of {'pull' if kind == 1 else 'pull_grad'}(f0, phi), {want}, got {tuple(delta.shape)}")
    out = torch.empty((nb, 3) + d1 if phi.ndim == 5 else (3,) + d1, dtype=torch.float32, device=f0.device)
    d = _desc(sett, d0, d1, nc, nb)
    fn = _lib.lib().pp_pull_backward if kind == 1 else _lib.lib().pp_pull_grad_backward
    _lib.check(fn(ct.byref(d), _ptr(out), _ptr(phi), _ptr(f0), _ptr(delta), _stream(f0)), "pullback")
    return out


def pull_backward(f0, phi, delta, sett=None):
    """sum_c delta_c * pull_grad(f0, phi)_c  -> shape of phi (src/pushpull_pb.jl:13-17)."""
    return _backward(1, f0, phi, delta, sett or Settings())


def pull_grad_backward(f0, phi, delta, sett=None):
    """sum_{a,c} delta_{a,c} * pull_hess(f0, phi)_{a,.,c} -> shape of phi (src/pushpull_pb.jl:26-35)."""
    return _backward(3, f0, phi, delta, sett or Settings())


def pull_vjp(f0, phi, delta, sett=None, df0=None, accumulate=False):
    """Both cotangents of pull (the rrule, src/pushpull_pb.jl:11-17) from one launch:
    df0 = push(delta, phi, size(f0)) and dphi = sum_c delta_c grad pull(f0)_c (device
    tensors; df0 is zeroed first unless accumulate). Returns (df0, dphi)."""
    sett = sett or Settings()
    f0 = _f32(f0, "f₀")
    phi = _f32(phi, "ϕ")
    delta = _f32(delta, "δ")
    _req(_is_dev(f0) and _is_dev(phi) and _is_dev(delta), "pull_vjp takes device tensors")
    _same_device(("f₀", f0), ("ϕ", phi), ("δ", delta))
    lead, d0 = _dims5(f0.shape, 3, 5, "f₀")
    nb, d1 = _phi_dims(phi)
    _req(nb == (f0.shape[0] if f0.ndim == 5 else 1), "`ϕ` & `f₀` must have the same batch size")
    nc = f0.shape[-4] if f0.ndim >= 4 else 1
    _req(tuple(delta.shape) == lead + d1, "δ must have the shape of pull(f0, phi)")
    dphi = torch.empty(tuple(phi.shape), dtype=torch.float32, device=f0.device)
    if df0 is None:
        df0 = torch.empty(tuple(f0.shape), dtype=torch.float32, device=f0.device)
    else:
        _out_arg(df0, f0.shape, f0, "df0")
    d = _desc(sett, d0, d1, nc, nb)
    _lib.check(_lib.lib().pp_pull_vjp(ct.byref(d), _ptr(dphi), _ptr(df0), _ptr(phi), _ptr(f0), _ptr(delta),
                                      int(accumulate), _stream(f0)), "pp_pull_vjp")
    return df0, dphi


def objective_grad(f0, phi, g, sett=None, df0=None, accumulate=False):
    """Fused L = 0.5*||pull(f0,phi) - g||^2 with dL/dphi and dL/df0 (device tensors).

    Returns (loss as a 0-d float64 device tensor, dphi, df0)."""
    sett = sett or Settings()
    f0 = _f32(f0, "f₀")
    phi = _f32(phi, "ϕ")
    g = _f32(g, "g")
    _req(_is_dev(f0) and _is_dev(phi) and _is_dev(g), "objective_grad takes device tensors")
    _same_device(("f₀", f0), ("ϕ", phi), ("g", g))
    lead, d0 = _dims5(f0.shape, 3, 5, "f₀")
    nb, d1 = _phi_dims(phi)
    _req(nb == (f0.shape[0] if f0.ndim == 5 else 1), "`ϕ` & `f₀` must have the same batch size")
    nc = f0.shape[-4] if f0.ndim >= 4 else 1
    _req(tuple(g.shape) == lead + d1, "g must have the shape of pull(f0, phi)")
    dphi = torch.empty(tuple(phi.shape), dtype=torch.float32, device=f0.device)
    if df0 is None:
        df0 = torch.empty(tuple(f0.shape), dtype=torch.float32, device=f0.device)
    else:
        _out_arg(df0, f0.shape, f0, "df0")
    loss = torch.empty((), dtype=torch.float64, device=f0.device)
    d = _desc(sett, d0, d1, nc, nb)
    _lib.check(_lib.lib().pp_objective_grad(ct.byref(d), _ptr(dphi), _ptr(df0), _ptr(loss), _ptr(phi), _ptr(f0),
                                            _ptr(g), int(accumulate), _stream(f0)), "pp_objective_grad")
    return loss, dphi, df0


def objective_grad_step(f0, phi, g, step, sett=None, phi_next=None, df0=None, accumulate=False):
    """objective_grad followed by the descent step phi_next = phi + step * dL/dphi in the
    same pass where the kernels allow (pp_objective_grad_step); phi_next may be phi itself
    (updated in place; it must then be a contiguous float32 device tensor). Returns (loss,
    dphi, df0, phi_next)."""
    sett = sett or Settings()
    f0 = _f32(f0, "f₀")
    phi_in = phi
    phi = _f32(phi, "ϕ")
    g = _f32(g, "g")
    _req(_is_dev(f0) and _is_dev(phi) and _is_dev(g), "objective_grad_step takes device tensors")
    _same_device(("f₀", f0), ("ϕ", phi), ("g", g))
    lead, d0 = _dims5(f0.shape, 3, 5, "f₀")
    nb, d1 = _phi_dims(phi)
    _req(nb == (f0.shape[0] if f0.ndim == 5 else 1), "`ϕ` & `f₀` must have the same batch size")
    nc = f0.shape[-4] if f0.ndim >= 4 else 1
    _req(tuple(g.shape) == lead + d1, "g must have the shape of pull(f0, phi)")
    if phi_next is None:
        phi_next = torch.empty_like(phi)
    elif phi_next is phi_in:
        _req(phi_in is phi, "an in-place step needs phi as a contiguous float32 device tensor")
    else:
        _out_arg(phi_next, phi.shape, phi, "phi_next")
    dphi = torch.empty(tuple(phi.shape), dtype=torch.float32, device=f0.device)
    if df0 is None:
        df0 = torch.empty(tuple(f0.shape), dtype=torch.float32, device=f0.device)
    else:
        _out_arg(df0, f0.shape, f0, "df0")
    loss = torch.empty((), dtype=torch.float64, device=f0.device)
    d = _desc(sett, d0, d1, nc, nb)
    _lib.check(_lib.lib().pp_objective_grad_step(ct.byref(d), _ptr(dphi), _ptr(df0), _ptr(loss), _ptr(phi),
                                                 _ptr(f0), _ptr(g), _ptr(phi_next), float(step), int(accumulate),
                                                 _stream(f0)), "pp_objective_grad_step")
    return loss, dphi, df0, phi_next
