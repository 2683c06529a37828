"""CPU oracle for the pull/push path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of bench.py
import this module, and only as the checker. The product path (``pushpull``) never
imports it.

Two libraries sit behind it:

* ``_build/libpp_oracle.so`` -- this project's C restatement (pp_oracle.c), in FP32
  (``prec="f32"``, bit-exact against the reference C) and FP64 (``prec="f64"``, masks
  and knots in FP32, arithmetic in double).
* ``_ref/libpushpull_ref.so`` -- the reference's own ``artifacts/C/pushpull.c``
  compiled from source by ``make ref`` (present only where that tree was available).

Array convention (identical memory to the Julia arrays of src/pushpull_cpu.jl):
images are C-ordered ``(B, C, Z, Y, X)`` (Julia ``(X, Y, Z, C, B)``); deformations are
``(B, 3, Z, Y, X)`` holding 1-based voxel coordinates, component 0 along X.
The per-(batch, channel) slice loop mirrors src/pushpull_cpu.jl:42-51 (pull),
:95-106 (pull_grad), :137-146 (pull_hess), :192-201 (push), :248-257 (push_grad).
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "_build", "libpp_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libpushpull_ref.so")

_P = ct.c_void_p
_SIG = [_P, _P, _P, _P, ct.c_ulonglong, _P, _P, ct.c_int]
# pull_affine / push_affine (artifacts/C/pushpull.h:11-14)
_SIG_AFF = [_P, _P, _P, ct.c_ulonglong, _P, _P, _P, _P, ct.c_int]

_lib = None
_ref = None


def lib():
    """The restatement library (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library missing: {ORACLE_SO} (run make -C oracle)")
        L = ct.CDLL(ORACLE_SO)
        for p in ("f32", "f64"):
            for op in ("pull", "pullg", "pullh", "push", "pushg"):
                fn = getattr(L, f"ppo_{op}_{p}")
                fn.argtypes = _SIG
                fn.restype = None
        for p in ("f32", "f64"):
            for op in ("pull_affine", "push_affine"):
                fn = getattr(L, f"ppo_{op}_{p}")
                fn.argtypes = _SIG_AFF
                fn.restype = None
        L.ppo_knot_export.argtypes = [ct.c_float, ct.c_int]
        L.ppo_knot_export.restype = ct.c_longlong
        L.ppo_fold_export.argtypes = [ct.c_int, ct.c_longlong, ct.c_ulonglong, _P, _P]
        L.ppo_fold_export.restype = None
        L.ppo_weights_f64.argtypes = [ct.c_int, ct.c_double, _P, _P, _P]
        L.ppo_weights_f32.argtypes = [ct.c_int, ct.c_float, _P, _P, _P]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference C compiled from source (``make -C oracle ref``)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference build missing: {REF_SO}")
        L = ct.CDLL(REF_SO)
        for op in ("pull", "pullg", "pullh", "push", "pushg"):
            fn = getattr(L, op)
            fn.argtypes = _SIG
            fn.restype = None
        for op in ("pull_affine", "push_affine"):
            fn = getattr(L, op)
            fn.argtypes = _SIG_AFF
            fn.restype = None
        _ref = L
    return _ref


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def _fn(op: str, prec: str):
    if prec == "ref":
        return getattr(ref(), op)
    return getattr(lib(), f"ppo_{op}_{prec}")


def _dtype(prec):
    return np.float64 if prec == "f64" else np.float32


def _as5(a, lead):
    """Promote a Julia-style 3..5-D array (reversed axes) to (B, C, Z, Y, X)."""
    a = np.asarray(a)
    while a.ndim < 5:
        a = a[None]
    return a


def _bnd_for(sett, c):
    b = sett.bnd[0] if len(sett.bnd) == 1 else sett.bnd[c]
    return np.ascontiguousarray(b, dtype=np.int32)


def _dp(sett):
    return np.ascontiguousarray([d + 1 for d in sett.deg], dtype=np.uint64)


def pull(f0, phi, sett, prec="f32"):
    """pull (src/pushpull_cpu.jl:17-53): f0 (B,C,Z0,Y0,X0), phi (B,3,Z1,Y1,X1)."""
    return _pull_family("pull", 1, f0, phi, sett, prec)


def pull_grad(f0, phi, sett, prec="f32"):
    """pull_grad (src/pushpull_cpu.jl:69-108): out (B,C,3,Z1,Y1,X1)."""
    return _pull_family("pullg", 3, f0, phi, sett, prec)


def pull_hess(f0, phi, sett, prec="f32"):
    """pull_hess (src/pushpull_cpu.jl:111-148): out (B,C,3,3,Z1,Y1,X1), column-major 3x3."""
    return _pull_family("pullh", 9, f0, phi, sett, prec)


def _pull_family(op, ncomp, f0, phi, sett, prec):
    dt = _dtype(prec)
    f0 = np.ascontiguousarray(_as5(f0, 5), dtype=dt)
    phi = np.ascontiguousarray(_as5(phi, 5), dtype=np.float32)
    B, C = f0.shape[:2]
    assert phi.shape[1] == 3 and phi.shape[0] == B
    assert len(sett.bnd) in (1, C)
    d0 = np.ascontiguousarray(f0.shape[:1:-1], dtype=np.uint64)  # (X0, Y0, Z0)
    d1 = phi.shape[2:]
    n1 = int(np.prod(d1))
    out = np.zeros((B, C, ncomp) + d1, dtype=dt)
    dp = _dp(sett)
    fn = _fn(op, prec)
    for b in range(B):
        for c in range(C):
            bnd = _bnd_for(sett, c)
            o = out[b, c]
            fn(_ptr(o), _ptr(phi[b]), _ptr(f0[b, c]), _ptr(d0), n1, _ptr(bnd), _ptr(dp), int(sett.ext))
    if ncomp == 1:
        return out[:, :, 0]
    if ncomp == 9:
        return out.reshape((B, C, 3, 3) + d1)
    return out


def push(f1, phi, dim, sett, prec="f32"):
    """push (src/pushpull_cpu.jl:165-203): f1 (B,C,Z1,Y1,X1) -> (B,C)+dim, dim=(Z0,Y0,X0)."""
    return _push_family("push", 1, f1, phi, dim, sett, prec)


def push_grad(g1, phi, dim, sett, prec="f32"):
    """push_grad (src/pushpull_cpu.jl:220-259) with the adjoint semantics of
    ext/pushpull_gpu.jl / pushpull.cu:145-163: g1 (B,C,3,Z1,Y1,X1)."""
    return _push_family("pushg", 3, g1, phi, dim, sett, prec)


def _push_family(op, ncomp, f1, phi, dim, sett, prec):
    dt = _dtype(prec)
    f1 = np.asarray(f1)
    if ncomp == 3:
        while f1.ndim < 6:
            f1 = f1[None]
    else:
        f1 = _as5(f1, 5)
    f1 = np.ascontiguousarray(f1, dtype=dt)
    phi = np.ascontiguousarray(_as5(phi, 5), dtype=np.float32)
    B, C = f1.shape[:2]
    assert phi.shape[0] == B and phi.shape[1] == 3
    assert tuple(f1.shape[-3:]) == tuple(phi.shape[2:])
    dim = tuple(int(v) for v in dim)
    d0 = np.ascontiguousarray(dim[::-1], dtype=np.uint64)
    n1 = int(np.prod(phi.shape[2:]))
    out = np.zeros((B, C) + dim, dtype=dt)
    dp = _dp(sett)
    fn = _fn(op, prec)
    for b in range(B):
        for c in range(C):
            bnd = _bnd_for(sett, c)
            fn(_ptr(out[b, c]), _ptr(phi[b]), _ptr(f1[b, c]), _ptr(d0), n1, _ptr(bnd), _ptr(dp), int(sett.ext))
    return out


def knot(x: float, dp: int) -> int:
    return int(lib().ppo_knot_export(ct.c_float(x), dp))


def fold(code: int, i: int, d: int):
    b = ct.c_ulonglong()
    s = ct.c_int()
    lib().ppo_fold_export(code, i, d, ct.byref(b), ct.byref(s))
    return int(b.value), int(s.value)


def weights(dp: int, t: float, prec="f64"):
    dt = _dtype(prec)
    w = np.zeros(5, dt)
    dw = np.zeros(5, dt)
    hw = np.zeros(5, dt)
    fn = lib().ppo_weights_f64 if prec == "f64" else lib().ppo_weights_f32
    fn(dp, t, _ptr(w), _ptr(dw), _ptr(hw))
    return w[:dp], dw[:dp], hw[:dp]


def backward_phi(f0, phi, delta, sett, prec="f64"):
    """dL/dphi for L with upstream gradient delta on pull(f0, phi): the channel
    contraction sum_c delta_c * pull_grad(f0, phi)_c of src/pushpull_pb.jl:13-17."""
    g = pull_grad(f0, phi, sett, prec)  # (B,C,3,...)
    d = np.asarray(delta, dtype=g.dtype)
    d = _as5(d, 5)
    return (d[:, :, None] * g).sum(axis=1)  # (B,3,...)


def adjust_affine(Aff):
    """src/pushpull_cpu.jl:357-362: keep rows 1-3 and set column 4 to the row sums minus 1
    (1-based -> 0-based voxel indices), all in Float32 with the row sum taken in order.
    Aff is indexed like the Julia matrix (row, column). Returns the 3x4 float32 matrix."""
    A = np.asarray(Aff, dtype=np.float32)[:3, :4].copy()
    for r in range(3):
        s = np.float32(0)
        for c in range(4):
            s = np.float32(s + A[r, c])
        A[r, 3] = np.float32(s - np.float32(1))
    return A


def _aff12(A):
    """3x4 (row, column) -> the 12 floats of the Julia column-major matrix."""
    return np.ascontiguousarray(np.asarray(A, np.float32).T.ravel())


def affine_pull(f0, A0, d1, sett, prec="f32"):
    """affine_pull with an already 0-based 3x4 matrix A0 (see adjust_affine):
    f0 (B,C,Z0,Y0,X0) -> (B,C)+d1, d1 in array order (Z1,Y1,X1). pushpull.c:176-209."""
    dt = _dtype(prec)
    f0 = np.ascontiguousarray(_as5(f0, 5), dtype=dt)
    B, C = f0.shape[:2]
    d1 = tuple(int(v) for v in d1)
    out = np.zeros((B, C) + d1, dtype=dt)
    d0u = np.ascontiguousarray(f0.shape[:1:-1], dtype=np.uint64)
    d1u = np.ascontiguousarray(d1[::-1], dtype=np.uint64)
    a12 = _aff12(A0)
    dp = _dp(sett)
    fn = ref().pull_affine if prec == "ref" else getattr(lib(), f"ppo_pull_affine_{prec}")
    n1 = int(np.prod(d1))
    for b in range(B):
        for c in range(C):
            bnd = _bnd_for(sett, c)
            fn(_ptr(out[b, c]), _ptr(f0[b, c]), _ptr(d0u), n1, _ptr(d1u), _ptr(a12), _ptr(bnd), _ptr(dp),
               int(sett.ext))
    return out


def affine_push(f1, A0, d0, sett, prec="f32"):
    """affine_push: f1 (B,C,Z1,Y1,X1) -> (B,C)+d0, pushpull.c:213-244."""
    dt = _dtype(prec)
    f1 = np.ascontiguousarray(_as5(f1, 5), dtype=dt)
    B, C = f1.shape[:2]
    d0 = tuple(int(v) for v in d0)
    out = np.zeros((B, C) + d0, dtype=dt)
    d0u = np.ascontiguousarray(d0[::-1], dtype=np.uint64)
    d1u = np.ascontiguousarray(f1.shape[:1:-1], dtype=np.uint64)
    a12 = _aff12(A0)
    dp = _dp(sett)
    fn = ref().push_affine if prec == "ref" else getattr(lib(), f"ppo_push_affine_{prec}")
    n1 = int(np.prod(f1.shape[2:]))
    for b in range(B):
        for c in range(C):
            bnd = _bnd_for(sett, c)
            fn(_ptr(out[b, c]), _ptr(f1[b, c]), _ptr(d0u), n1, _ptr(d1u), _ptr(a12), _ptr(bnd), _ptr(dp),
               int(sett.ext))
    return out


# ---- compositions (src/compositions.jl) ----------------------------------------------
SLIDING = [(2, 1, 1), (1, 2, 1), (1, 1, 2)]  # compositions.jl:33; column c = channel c's triple


class _Sett:
    """Settings as the Julia constructor leaves them (deg tuple, list of bnd triples, ext)."""

    def __init__(self, deg, bnd, ext):
        self.deg, self.bnd, self.ext = tuple(deg), list(bnd), int(ext)


def identity(dzyx):
    """id (src/pushpull_general.jl:81-91): 1-based grid, (3, Z, Y, X) float32."""
    Z, Y, X = dzyx
    g = np.empty((3, Z, Y, X), dtype=np.float32)
    g[0] = np.arange(1, X + 1, dtype=np.float32)[None, None, :]
    g[1] = np.arange(1, Y + 1, dtype=np.float32)[None, :, None]
    g[2] = np.arange(1, Z + 1, dtype=np.float32)[:, None, None]
    return g


def compose(phi, theta, prec="f32"):
    """compose! (src/compositions.jl:27-39): for dim = 1:3, psi[.., dim] = pull(phi[.., dim],
    theta, Settings((1,1,1), bnd[:, dim], true)) -- one pull per channel with its own
    boundary triple. phi (B,3,Z0,Y0,X0) displacement, theta (B,3,Z1,Y1,X1) 1-based."""
    return pull(_as5(phi, 5), _as5(theta, 5), _Sett((1, 1, 1), SLIDING, 1), prec)


def exp(v, nsq=8, prec="f32"):
    """Exp (src/compositions.jl:41-49) as written: u = v / 2^nsq (exact); nsq times
    u = compose(u, Id .+ u) with Id .+ u formed in Float32 (the deformation is Float32 in
    the reference); finally u .+ Id. v (B,3,Z,Y,X)."""
    dt = _dtype(prec)
    v = np.asarray(_as5(v, 5), dtype=np.float32)
    Id = identity(v.shape[2:])[None]
    u = (v / np.float32(2.0 ** nsq)).astype(dt)
    for _ in range(nsq):
        theta = (Id + u.astype(np.float32)).astype(np.float32) if prec != "f64" else (Id + u).astype(np.float32)
        u = compose(u, theta, prec)
    return (u + Id).astype(dt)


# ---- push with a count image, multigrid resampling (src/multigrid.jl) ----------------
def push_count(f1, phi, dim, sett, count_bnd=None, prec="f32"):
    """pushc (artifacts/C/pushpull.cu:128-142) as its callers use it: (push(f1), push(ones))
    through the same deformation; the count uses `count_bnd` (default: sett's first triple).
    f1 None pushes the identity of phi's grid (src/multigrid.jl:269-271)."""
    phi = np.asarray(_as5(phi, 5), dtype=np.float32)
    B = phi.shape[0]
    if f1 is None:
        f1 = np.broadcast_to(identity(phi.shape[2:])[None], (B, 3) + phi.shape[2:]).copy()
    cb = tuple(count_bnd) if count_bnd is not None else tuple(sett.bnd[0])
    f0 = push(f1, phi, dim, sett, prec)
    ones = np.ones((B, 1) + phi.shape[2:], np.float32)
    c0 = push(ones, phi, dim, _Sett(sett.deg, [cb], sett.ext), prec)
    return f0, c0


def invert_def_rhs(phi, prec="f32"):
    """invert_def's pushes (src/multigrid.jl:266-275): g = push(Id, phi, Settings(1, sliding,
    0)), h = push(ones, phi, Settings(1, 1, 0)), g .-= h .* Id (elementwise in Float32)."""
    phi = np.asarray(_as5(phi, 5), dtype=np.float32)
    d = phi.shape[2:]
    g, h = push_count(None, phi, d, _Sett((1, 1, 1), SLIDING, 0), (1, 1, 1), prec)
    Id = identity(d)[None]
    g = g - h * Id
    return g, h


def zoom_affine(d_from, d_to):
    """Mat of restrict / prolong (src/multigrid.jl:23-25, 45-47): zm = d_to ./ d_from and
    os = (1 - zm) / 2 per Julia axis, in Float64, converted to Float32 (sizes (Z, Y, X))."""
    M = np.zeros((4, 4))
    for q in range(3):
        zm = d_to[2 - q] / d_from[2 - q]
        M[q, q] = zm
        M[q, 3] = (1 - zm) / 2
    M[3, 3] = 1
    return M.astype(np.float32)


def restrict(f1, d0, deg=2, prec="f32"):
    """restrict (src/multigrid.jl:21-35): affine_push(f1) ./ affine_push(ones)."""
    f1 = np.asarray(_as5(f1, 5), dtype=np.float32)
    M = adjust_affine(zoom_affine(f1.shape[2:], d0))
    s = _Sett((deg,) * 3, [(1, 1, 1)], 1)
    f0 = affine_push(f1, M, d0, s, prec)
    c0 = affine_push(np.ones((f1.shape[0], 1) + f1.shape[2:], np.float32), M, d0, s, prec)
    return f0 / c0


def prolong(f0, d1, deg=2, prec="f32"):
    """prolong (src/multigrid.jl:43-52): affine_pull with the sliding triples."""
    f0 = np.asarray(_as5(f0, 5), dtype=np.float32)
    return affine_pull(f0, adjust_affine(zoom_affine(d1, f0.shape[2:])), d1, _Sett((deg,) * 3, SLIDING, 1), prec)
