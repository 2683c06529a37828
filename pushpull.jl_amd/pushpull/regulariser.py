"""The sparse regulariser of Gauss-Newton registration (src/operator.jl, src/operator_sparse.jl).

Operator construction runs on the host in NumPy, as the reference's does in Julia:

* ``registration_operator(vx, λ, vx0=vx)`` (operator.jl:58-112): the dense stencil L of
  the absolute, membrane, bending and divergence penalties.
* ``reduce2fit(L, d, bnd)`` (operator.jl:120-159): collapses L along axes of size 1.
* ``sparsify(L, d, nd=3)`` (operator.jl:206-296): the list format of the kernels.

The two operators run on the GPU through libpushpull_hip:

* ``vel2mom(v, kernel, bnd)`` / ``vel2mom_(u, v, kernel, bnd)`` (operator_sparse.jl:6-40):
  u = L v.
* ``relax_(g, h, kernel, bnd, nit, v)`` (operator_sparse.jl:43-62): ``nit`` Gauss-Seidel
  sweeps of (L + H) v = g, updating v.

Both are bit-identical to the reference's C code (artifacts/C/sparse_operator.c). Device
tensors go through ``pp_vel2mom`` / ``pp_relax`` on the current stream; NumPy arrays
through the host-pointer entry points ``vel2mom`` / ``relax`` (synchronous).

Array order is this package's: a velocity field is ``(C, Z, Y, X)`` (Julia
``(X, Y, Z, C)``); L is ``(D, Kz, Ky, Kx)`` (Julia ``(Kx, Ky, Kz, D)``); sizes ``d`` and
``stride`` are ``(Z, Y, X)``; a boundary table ``bnd`` holds one ``(x, y, z)`` triple of
codes per channel (Julia's ``bnd[:, c]``), by default ``[2 1 1; 1 2 1; 1 1 2]``.

Sums of list values: ``sparsify`` stores, before each list, the sum of its non-centre
values, added in list order in FP32 -- the order Julia's ``sum`` uses for fewer than 16
terms; for longer lists Julia may vectorise that sum, so the slot can differ from
Julia's in the last bit. The GPU operators take the kernel as given.
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .ops import _is_dev, _ptr, _req, _stream

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

DEFAULT_BND = ((2, 1, 1), (1, 2, 1), (1, 1, 2))
MAXD = 5  # largest stencil side (operator_sparse.jl:66,103-104)
MAXN = 8  # most channels (sparse_operator_dev.cu:8)


# ---------------------------------------------------------------- construction ----
def _acc(L, idx, rhs):
    """``L[idx] .+= rhs`` with Julia's promotion: an FP32 right-hand side adds in FP32, an
    FP64 one adds in FP64 and is rounded to FP32 on the store."""
    rhs = np.asarray(rhs)
    if rhs.dtype == np.float32:
        L[idx] = L[idx] + rhs
    else:
        L[idx] = (L[idx].astype(np.float64) + rhs.astype(np.float64)).astype(np.float32)


def _pow4(x, T):
    """``x^4``: Julia raises an FP32 to an integer power in FP64 (power by squaring) and
    rounds back."""
    if T is np.float32:
        y = np.float64(x) * np.float64(x)
        return np.float32(y * y)
    return np.float64(x) ** 4


def registration_operator(vx, lam, vx0=None) -> np.ndarray:
    """Dense differential operator of registration (src/operator.jl:58-112).

    ``lam`` = (absolute, membrane, bending, divergence) penalties. Returns L as
    ``(D, K, K, K)`` FP32 with K = 1, 3 or 5 and D = 3 (6 when lam[3] != 0, the last three
    channels being the (x,y), (x,z), (y,z) cross terms). The arithmetic runs in FP32 when
    vx, lam and vx0 are all FP32 (as multigrid passes them, src/multigrid.jl:175-178) and
    in FP64 otherwise (exact for integer arguments, as in Julia).
    """
    _req(len(vx) == 3, "Wrong sized vx.")
    _req(len(lam) == 4, "Wrong sized λ.")
    if vx0 is None:
        vx0 = vx
    f32 = all(np.asarray(a).dtype == np.float32 for a in (vx, lam, vx0))
    T = np.float32 if f32 else np.float64
    vx = [T(a) for a in np.asarray(vx)]
    vx0 = [T(a) for a in np.asarray(vx0)]
    lam = [T(a) for a in np.asarray(lam)]
    d123 = 3 if (lam[1] != 0 or lam[3] != 0) else 1
    d123 = 5 if lam[2] != 0 else d123
    d4 = 3 if lam[3] == 0 else 6
    L = np.zeros((d4, d123, d123, d123), dtype=np.float32)
    c = (d123 + 1) // 2 - 1  # 0-based centre
    D = np.array([-1, 2, -1], dtype=T)  # Laplacian
    r3 = slice(c - 1, c + 2)
    r5 = slice(c - 2, c + 3)

    def sq(a):
        return a * a

    if lam[0] != 0:
        for i in range(3):
            _acc(L, (i, c, c, c), lam[0] * sq(vx0[i]))
    if lam[1] != 0:  # membrane energy
        for i in range(3):
            _acc(L, (i, c, c, r3), D * (lam[1] * sq(vx0[i]) / sq(vx[0])))
            _acc(L, (i, c, r3, c), D * (lam[1] * sq(vx0[i]) / sq(vx[1])))
            _acc(L, (i, r3, c, c), D * (lam[1] * sq(vx0[i]) / sq(vx[2])))
    if lam[2] != 0:  # bending energy
        DD = np.array([1, -4, 6, -4, 1], dtype=T)
        DDt = np.outer(D, D)  # symmetric
        two = T(2)
        for i in range(3):
            _acc(L, (i, c, c, r5), DD * (lam[2] * sq(vx0[i]) / _pow4(vx[0], T)))
            _acc(L, (i, c, r5, c), DD * (lam[2] * sq(vx0[i]) / _pow4(vx[1], T)))
            _acc(L, (i, r5, c, c), DD * (lam[2] * sq(vx0[i]) / _pow4(vx[2], T)))
            _acc(L, (i, c, r3, r3), DDt * (two * lam[2] * sq(vx0[i]) / sq(vx[0] * vx[1])))
            _acc(L, (i, r3, c, r3), DDt * (two * lam[2] * sq(vx0[i]) / sq(vx[0] * vx[2])))
            _acc(L, (i, r3, r3, c), DDt * (two * lam[2] * sq(vx0[i]) / sq(vx[1] * vx[2])))
    if lam[3] != 0:  # squared divergence
        g = np.array([-0.5, 0.0, 0.5])
        GGt = -np.outer(g, g)  # FP64, as Julia's [-1/2, 0, 1/2]
        _acc(L, (0, c, c, r3), D * (lam[3] * sq(vx0[0]) / sq(vx[0])))
        _acc(L, (1, c, r3, c), D * (lam[3] * sq(vx0[1]) / sq(vx[1])))
        _acc(L, (2, r3, c, c), D * (lam[3] * sq(vx0[2]) / sq(vx[2])))
        # numpy (y, x) planes hold Julia's [x, y] matrices transposed; GGt is symmetric
        _acc(L, (3, c, r3, r3), GGt * np.float64(lam[3] * vx0[0] * vx0[1] / (vx[0] * vx[1])))
        _acc(L, (4, r3, c, r3), GGt * np.float64(lam[3] * vx0[0] * vx0[2] / (vx[0] * vx[2])))
        _acc(L, (5, r3, r3, c), GGt * np.float64(lam[3] * vx0[1] * vx0[2] / (vx[1] * vx[2])))
    return L


def reduce2fit(L, d, bnd=DEFAULT_BND) -> np.ndarray:
    """Collapse L along every axis where the volume has size 1 (src/operator.jl:120-159):
    the stencil is summed along that axis, with alternating signs for channels whose
    boundary code there is 2. ``d`` is ``(Z, Y, X)``."""
    L = np.asarray(L)
    dp = L.shape  # (D, Kz, Ky, Kx), fixed before any reduction as in the reference
    b = np.asarray(bnd)
    for ax in range(3):  # x, y, z
        if int(d[2 - ax]) != 1:
            continue
        n = dp[3 - ax]
        cen = (n + 1) // 2
        W = np.ones((dp[0], n), dtype=np.float64)
        for ch in range(dp[0]):
            if ch < b.shape[0] and int(b[ch][ax]) == 2:
                W[ch] = [1.0 if (t - cen) % 2 == 0 else -1.0 for t in range(1, n + 1)]
        axis = 3 - ax
        Lm = np.moveaxis(L.astype(np.float64), axis, 1)  # (D, n, ...)
        acc = np.zeros((Lm.shape[0],) + Lm.shape[2:], dtype=np.float64)
        for t in range(Lm.shape[1]):
            acc = acc + W[:, t].reshape((-1,) + (1,) * (acc.ndim - 1)) * Lm[:, t]
        L = np.moveaxis(acc[:, None], 1, axis).astype(np.float32)
    return L


@dataclass(eq=False)
class SparseKernel:
    """The `sparsify` tuple (src/operator.jl:162-205) in this package's array order.

    ``offset[j, i]`` / ``length[j, i]``: the list of output channel i and input channel j
    (Julia ``offset[i, j]``); its entries are ``values[offset : offset + length]`` and
    ``patch_indices`` likewise (0-based, the first entry being the centre), and
    ``values[offset - 1]`` is the sum of its non-centre values. ``indices`` holds the same
    positions as offsets into a (X, Y, Z, C) array, relative to the centre voxel.
    """
    stride: tuple
    d: tuple
    nchan: int
    offset: np.ndarray
    length: np.ndarray
    values: np.ndarray
    indices: np.ndarray
    patch_indices: np.ndarray
    _handles: dict = field(default_factory=dict, repr=False)

    def handle(self, device_index: int) -> ct.c_void_p:
        """The kernel uploaded to one GPU (created on first use)."""
        h = self._handles.get(device_index)
        if h is None:
            h = ct.c_void_p()
            with torch.cuda.device(device_index):
                _lib.check(_lib.lib().pp_sparse_create(
                    ct.byref(h), (ct.c_int64 * 3)(*[int(s) for s in self.d[::-1]]), int(self.nchan),
                    (ct.c_int32 * 3)(*[int(s) for s in self.stride[::-1]]),
                    _hptr(self.offset), _hptr(self.length), _hptr(self.values), _hptr(self.patch_indices),
                    int(self.values.size)), "pp_sparse_create")
            self._handles[device_index] = h
        return h

    def __del__(self):
        try:
            for h in self._handles.values():
                _lib.lib().pp_sparse_destroy(h)
        except Exception:  # interpreter shutdown
            pass


def _hptr(a: np.ndarray):
    return a.ctypes.data_as(ct.c_void_p)


def sparsify(L, d, nd: int = 3) -> SparseKernel:
    """List form of the dense stencil L for volumes of size d (src/operator.jl:206-296)."""
    L = np.asarray(L)
    D, kz, ky, kx = L.shape
    dp = (kx, ky, kz)
    _req(all(s % 2 == 1 for s in dp), "First three dimensions of `L` must be odd")
    _req(D in (3, 6), "Incorrectly sized `L`")
    dx, dy, dz = int(d[2]), int(d[1]), int(d[0])
    cv = ((kx + 1) // 2, (ky + 1) // 2, (kz + 1) // 2)  # 1-based centre

    def image_offset(i, j, k, c1):
        return i - cv[0] + dx * (j - cv[1] + dy * (k - cv[2] + dz * (c1 - 1)))

    def patch_offset(i, j, k):
        return i - 1 + kx * (j - 1 + ky * (k - 1))

    def val(i, j, k, c):  # Julia L[i, j, k, c], 1-based
        return L[c - 1, k - 1, j - 1, i - 1]

    lists = {(a, b): [] for b in range(1, nd + 1) for a in range(1, nd + 1)}  # (c1, c2) = (out, in)

    def scan(c, emit):
        for k in range(1, kz + 1):
            for j in range(1, ky + 1):
                for i in range(1, kx + 1):
                    if (i, j, k) != cv and val(i, j, k, c) != 0:
                        emit(i, j, k, np.float32(val(i, j, k, c)))

    for c in range(1, nd + 1):
        lists[c, c].append((np.float32(val(*cv, c)), image_offset(*cv, c), patch_offset(*cv)))
        scan(c, lambda i, j, k, w, c=c: lists[c, c].append((w, image_offset(i, j, k, c), patch_offset(i, j, k))))
    pairs = [(c1, c2) for c1 in range(1, nd + 1) for c2 in range(c1 + 1, nd + 1)]
    if D * 2 == nd * (nd + 1):
        for n, (c1, c2) in enumerate(pairs):
            c = nd + 1 + n
            w0 = np.float32(val(*cv, c))
            lists[c1, c2].append((w0, image_offset(*cv, c2), patch_offset(*cv)))
            lists[c2, c1].append((w0, image_offset(*cv, c1), patch_offset(*cv)))

            def both(i, j, k, w, c1=c1, c2=c2):
                lists[c1, c2].append((w, image_offset(i, j, k, c2), patch_offset(i, j, k)))
                lists[c2, c1].append((w, image_offset(i, j, k, c1), patch_offset(i, j, k)))

            scan(c, both)
    else:
        for c1, c2 in pairs:
            lists[c1, c2].append((np.float32(0.0), image_offset(*cv, c2), patch_offset(*cv)))
            lists[c2, c1].append((np.float32(0.0), image_offset(*cv, c1), patch_offset(*cv)))

    # rearrange (operator.jl:211-224): column-major over (i, j), a sum slot before each list
    total = sum(len(v) for v in lists.values()) + nd * nd
    values = np.zeros(total, dtype=np.float32)
    indices = np.zeros(total, dtype=np.int32)
    patch = np.zeros(total, dtype=np.int32)
    offset = np.zeros((nd, nd), dtype=np.int32)
    length = np.zeros((nd, nd), dtype=np.int32)
    to = 0
    for j in range(1, nd + 1):
        for i in range(1, nd + 1):
            ent = lists[i, j]
            s = np.float32(0.0)
            for n, (w, _, _) in enumerate(ent[1:]):
                s = w if n == 0 else np.float32(s + w)
            values[to] = s
            for n, (w, io, po) in enumerate(ent):
                values[to + 1 + n] = w
                indices[to + 1 + n] = io
                patch[to + 1 + n] = po
            offset[j - 1, i - 1] = to + 1
            length[j - 1, i - 1] = len(ent)
            to += len(ent) + 1
    return SparseKernel(stride=(kz, ky, kx), d=(int(d[0]), int(d[1]), int(d[2])), nchan=nd, offset=offset,
                        length=length, values=values, indices=indices, patch_indices=patch)


# ---------------------------------------------------------------- operators ----
def _bnd(bnd, nc):
    b = np.asarray(bnd, dtype=np.int64)
    if b.shape != (nc, 3):
        raise AssertionError("inappropriate boundary conditions")
    return (ct.c_int32 * (3 * nc))(*[int(x) for x in b.reshape(-1)])


def _vol(a, kernel, name, nchan=None):
    """Check one volume against the kernel; returns (is_device, array)."""
    if _is_dev(a):
        _req(a.dtype == torch.float32, f"`{name}` must be Float32")
        _req(a.is_contiguous(), f"`{name}` must be contiguous")
    else:
        _req(isinstance(a, np.ndarray) and a.dtype == np.float32 and a.flags.c_contiguous,
             f"`{name}` must be a contiguous Float32 array")
    _req(a.ndim in (3, 4), "too many dimensions" if a.ndim > 4 else "too few dimensions")
    sp = tuple(int(s) for s in a.shape[-3:])
    _req(sp == tuple(kernel.d), "incompatible dimensions")
    c = 1 if a.ndim == 3 else int(a.shape[0])
    if nchan is not None:
        _req(c == nchan, f"`{name}` must have {nchan} channels")
    return _is_dev(a), a


def vel2mom_(u, v, kernel: SparseKernel, bnd=DEFAULT_BND):
    """vel2mom! (src/operator_sparse.jl:16-40): u = L v, in place. As in the reference,
    voxels within (stride-1)/2 of a face add L v to the u they are given and the others
    overwrite it."""
    nc = kernel.nchan
    dev, v = _vol(v, kernel, "v", nc)
    dev_u, u = _vol(u, kernel, "u", nc)
    _req(dev == dev_u, "`u` and `v` must both be device tensors or both host arrays")
    b = _bnd(bnd, nc)
    if dev:
        _req(u.device == v.device, "`u` and `v` must be on the same device")
        h = kernel.handle(v.device.index)
        with torch.cuda.device(v.device):
            _lib.check(_lib.lib().pp_vel2mom(h, _ptr(u), _ptr(v), b, 1, _stream(v)), "pp_vel2mom")
        return u
    d = (ct.c_ulonglong * 5)(*[int(s) for s in kernel.d[::-1]], nc, 1)
    dp = (ct.c_ulonglong * 3)(*[int(s) for s in kernel.stride[::-1]])
    _lib.lib().vel2mom(_hptr(u), _hptr(v), d, _hptr(kernel.offset), _hptr(kernel.length), _hptr(kernel.values),
                       _hptr(kernel.indices), _hptr(kernel.patch_indices), dp, b)
    _lib.check(_lib.lib().pp_last_status(), "vel2mom")
    return u


def vel2mom(v, kernel: SparseKernel, bnd=DEFAULT_BND):
    """u = L v (src/operator_sparse.jl:6-12)."""
    if _is_dev(v):
        u = torch.zeros_like(v)
    else:
        u = np.zeros_like(np.asarray(v))
    return vel2mom_(u, v, kernel, bnd)


def _hcode(g, h, nc):
    """The Hessian layout (src/operator_sparse.jl:90-101)."""
    if h is None or (h.numel() if _is_dev(h) else np.asarray(h).size) == 0:
        return 0
    _req(h.ndim == g.ndim, "incompatible dimensionality")
    ch = 1 if h.ndim == 3 else int(h.shape[0])
    if ch == nc:
        return 1
    _req(2 * ch == nc * (nc + 1), "incompatible hessian dimensions")
    return 2


def relax_(g, h, kernel: SparseKernel, bnd, nit: int, v):
    """relax! (src/operator_sparse.jl:43-62): ``nit`` Gauss-Seidel sweeps of
    (L + H) v = g, updating and returning v. ``h`` may be None (no H), hold ``nchan``
    diagonal channels, or ``nchan*(nchan+1)/2`` (the diagonal, then the pairs (j, i>j))."""
    nc = kernel.nchan
    b = np.asarray(bnd)
    if b.ndim != 2 or b.shape != (nc, 3) or not np.isin(b, (0, 1, 2)).all():
        raise AssertionError("inappropriate boundary conditions")
    if any(s > MAXD for s in kernel.stride):
        raise AssertionError("regulariser dimensions too big")
    dev, g = _vol(g, kernel, "g", nc)
    dev_v, v = _vol(v, kernel, "v", nc)
    _req(dev == dev_v, "`g` and `v` must both be device tensors or both host arrays")
    hc = _hcode(g, h, nc)
    if hc:
        dev_h, h = _vol(h, kernel, "h")
        _req(dev_h == dev, "`h` must live where `g` does")
    bb = _bnd(bnd, nc)
    if dev:
        _req(g.device == v.device and (not hc or h.device == v.device), "arrays must be on one device")
        hk = kernel.handle(v.device.index)
        with torch.cuda.device(v.device):
            _lib.check(_lib.lib().pp_relax(hk, _ptr(v), _ptr(g), _ptr(h) if hc else None, hc, bb, int(nit),
                                           _stream(v)), "pp_relax")
        return v
    d = (ct.c_ulonglong * 5)(*[int(s) for s in kernel.d[::-1]], nc, hc)
    dp = (ct.c_ulonglong * 3)(*[int(s) for s in kernel.stride[::-1]])
    hp = _hptr(h) if hc else None
    for _ in range(int(nit)):  # one sweep per call, as src/operator_sparse.jl:50-60
        _lib.lib().relax(_hptr(v), d, _hptr(g), hp, _hptr(kernel.offset), _hptr(kernel.length),
                         _hptr(kernel.values), _hptr(kernel.indices), _hptr(kernel.patch_indices), dp, bb)
        _lib.check(_lib.lib().pp_last_status(), "relax")
    return v
