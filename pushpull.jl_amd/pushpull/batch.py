"""Batches of uneven volumes in one persistent launch, and HIP graph capture.

The reference only batches same-size volumes, one native call per (channel, batch) slice
(src/pushpull_cpu.jl:42-51). ``BatchPlan`` takes a list of volumes with their own sizes,
channel counts, degrees and deformations and runs one op over all of them in one launch
(``pp_batch_plan_create`` / ``pp_batch_run``). ``Graph`` captures any sequence of calls on a
side stream into a HIP graph (``pp_graph_*``) and replays it. ``run_static`` runs one op
over a list of volumes without any device allocation (``pp_batch_run_static``).

Per volume (C-order shapes, as everywhere in this package):
  f    image (C, Z0, Y0, X0) for pull-type ops; values (C, Z1, Y1, X1) for push,
       (C, 3, Z1, Y1, X1) for push_grad
  phi  (3, Z1, Y1, X1), 1-based
  aux  upstream gradient (pull_backward) or target (objective), (C, Z1, Y1, X1)
  deg  degree 0..4, one for every axis or a triple (X, Y, Z) as Settings takes
  bnd  a boundary triple for all channels, or a list of C triples, one per channel (at
       most 10; along an axis the channels' codes equal or all in {1, 2})
  ext  extrapolation flag (default 1)
  dim  push, push_grad: the (Z, Y, X) output grid
  out, out2, loss
       outputs (allocated when absent): out = pull / push / dphi; out2 = dL/df (objective)
       or push(aux) (pull_vjp); loss = the objective's 0.5 sum r^2 (float64)
  phi_next, step
       objective only: phi_next (phi's shape; may be phi itself) receives
       phi + step * dL/dphi, a descent step fused into the gradient pass
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import _lib

OPS = {"pull": 0, "push": 3, "push_grad": 4, "pull_backward": 5, "objective": 7, "pull_vjp": 8}
BATCH_CHAN_BND = 10  # PP_BATCH_CHAN_BND: channels of a volume with boundary codes of their own


class PPVolume(ct.Structure):
    _fields_ = [
        ("f", ct.c_void_p), ("phi", ct.c_void_p), ("aux", ct.c_void_p), ("out", ct.c_void_p),
        ("out2", ct.c_void_p), ("loss", ct.c_void_p),
        ("d0", ct.c_int64 * 3), ("d1", ct.c_int64 * 3),
        ("nchan", ct.c_int32), ("deg", ct.c_int32 * 3), ("ext", ct.c_int32), ("bnd", ct.c_int32 * 3),
        ("nbnd", ct.c_int32), ("chan_bnd", ct.c_int32 * (3 * BATCH_CHAN_BND)),
        ("phi_next", ct.c_void_p), ("step", ct.c_float),
    ]


def _sigs():
    L = _lib.lib()
    L.pp_batch_plan_create.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.POINTER(ct.c_void_p)]
    L.pp_batch_plan_create.restype = ct.c_int
    L.pp_batch_run.argtypes = [ct.c_void_p, ct.c_void_p]
    L.pp_batch_run.restype = ct.c_int
    L.pp_batch_plan_destroy.argtypes = [ct.c_void_p]
    L.pp_batch_plan_destroy.restype = ct.c_int
    L.pp_batch_run_static.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.c_void_p]
    L.pp_batch_run_static.restype = ct.c_int
    for n, a in (("pp_graph_begin", [ct.c_void_p]), ("pp_graph_end", [ct.c_void_p, ct.POINTER(ct.c_void_p)]),
                 ("pp_graph_launch", [ct.c_void_p, ct.c_void_p]), ("pp_graph_destroy", [ct.c_void_p])):
        fn = getattr(L, n)
        fn.argtypes = a
        fn.restype = ct.c_int
    return L


def _boundaries(b, C, where):
    """`bnd` of a volume as a list of triples: one for all channels, or one per channel."""
    try:
        nested = len(b) > 0 and hasattr(b[0], "__len__")
        codes = [tuple(int(x) for x in t) for t in b] if nested else [tuple(int(x) for x in b)]
    except TypeError:
        codes = [()]
    if any(len(t) != 3 or any(x not in (0, 1, 2) for x in t) for t in codes):
        raise ValueError(f"{where}: `bnd` must be three codes in 0..2, or one such triple per channel, got {b}")
    if len(codes) > 1 and len(codes) != C:
        raise ValueError(f"{where}: `bnd` holds {len(codes)} triples for {C} channels")
    if len(codes) > BATCH_CHAN_BND:
        raise ValueError(f"{where}: at most {BATCH_CHAN_BND} channels take boundary codes of their own")
    return codes


def _p(t):
    return ct.c_void_p(t.data_ptr()) if t is not None else ct.c_void_p(0)


def _shape3(v, key):
    t = v.get(key)
    if t is not None and not isinstance(t, torch.Tensor):
        raise TypeError(f"`{key}` must be a torch tensor")
    return t


def _volume_records(op: str, volumes: list[dict]):
    """Validate a list of volumes and describe them as a pp_volume array.

    pp_batch_plan_create and pp_batch_run_static see only pointers and sizes, so every
    shape, dtype, device and layout is checked here: a mismatch would make the kernels
    read or write outside a tensor. Shapes first, then dtypes, then devices and layout.
    Returns (records, ctypes array); the records keep every tensor the array points at
    alive (inputs made contiguous, outputs allocated where not given)."""
    if op not in OPS:
        raise ValueError(f"unknown batch op {op!r}; one of {sorted(OPS)}")
    if not volumes:
        raise ValueError("a batch needs at least one volume")
    arr = (PPVolume * len(volumes))()
    recs = []
    for i, v in enumerate(volumes):
        f, phi, aux = _shape3(v, "f"), _shape3(v, "phi"), _shape3(v, "aux")
        out, out2, loss = _shape3(v, "out"), _shape3(v, "out2"), _shape3(v, "loss")
        phi_next = _shape3(v, "phi_next")
        where = f"volume {i}"
        if f is None or phi is None:
            raise ValueError(f"{where}: `f` and `phi` are required")
        if phi.ndim != 4 or phi.shape[0] != 3:
            raise ValueError(f"{where}: `phi` must be (3, Z, Y, X), got {tuple(phi.shape)}")
        d1 = tuple(phi.shape[1:])
        deg = v["deg"]
        deg = tuple(int(x) for x in deg) if isinstance(deg, (tuple, list)) else (int(deg),) * 3
        if len(deg) != 3 or any(not 0 <= x <= 4 for x in deg):
            raise ValueError(f"{where}: `deg` must be a degree or three degrees in 0..4, got {v['deg']}")
        if op in ("push", "push_grad"):
            if "dim" not in v:
                raise ValueError(f"{where}: push needs `dim`, the (Z, Y, X) output grid")
            d0 = tuple(int(x) for x in v["dim"])
            if len(d0) != 3 or min(d0) < 1:
                raise ValueError(f"{where}: `dim` must be three positive sizes, got {d0}")
            lead = f.shape[:1] if op == "push" else f.shape[:2]
            want_f = (lead + d1) if op == "push" else (f.shape[:1] + (3,) + d1)
            if f.ndim != len(want_f) or tuple(f.shape) != tuple(want_f) or f.shape[0] < 1:
                shape = "(C, Z1, Y1, X1)" if op == "push" else "(C, 3, Z1, Y1, X1)"
                raise ValueError(f"{where}: `f` must be {shape} on phi's grid {d1}, got {tuple(f.shape)}")
        else:
            if f.ndim != 4 or f.shape[0] < 1:
                raise ValueError(f"{where}: `f` must be (C, Z0, Y0, X0), got {tuple(f.shape)}")
            d0 = tuple(f.shape[1:])
        C = int(f.shape[0])
        bnd = _boundaries(v.get("bnd", (1, 1, 1)), C, where)
        if op in ("pull_backward", "objective", "pull_vjp"):
            if aux is None:
                raise ValueError(f"{where}: `aux` is required for {op}")
            if tuple(aux.shape) != (C,) + d1:
                raise ValueError(f"{where}: `aux` must be {(C,) + d1}, got {tuple(aux.shape)}")
        out_shape = {"pull": (C,) + d1, "push": (C,) + d0, "push_grad": (C,) + d0}.get(op, (3,) + d1)
        if out is not None and tuple(out.shape) != out_shape:
            raise ValueError(f"{where}: `out` must be {out_shape}, got {tuple(out.shape)}")
        if op in ("objective", "pull_vjp") and out2 is not None and tuple(out2.shape) != (C,) + d0:
            raise ValueError(f"{where}: `out2` must be {(C,) + d0}, got {tuple(out2.shape)}")
        if op == "objective" and loss is not None and tuple(loss.shape) not in ((), (1,)):
            raise ValueError(f"{where}: `loss` must hold one value, got shape {tuple(loss.shape)}")
        if phi_next is not None:
            if op != "objective":
                raise ValueError(f"{where}: `phi_next` (a fused descent step) is only for the objective")
            if tuple(phi_next.shape) != tuple(phi.shape):
                raise ValueError(f"{where}: `phi_next` must be {tuple(phi.shape)}, got {tuple(phi_next.shape)}")
        # dtypes
        for key, t, dt in (("f", f, torch.float32), ("phi", phi, torch.float32), ("aux", aux, torch.float32),
                           ("out", out, torch.float32), ("out2", out2, torch.float32),
                           ("loss", loss, torch.float64), ("phi_next", phi_next, torch.float32)):
            if t is not None and t.dtype != dt:
                raise TypeError(f"{where}: `{key}` must be {dt}, got {t.dtype}")
        # devices and layout
        if not f.is_cuda:
            raise TypeError(f"{where}: batched volumes are device tensors")
        for key, t in (("phi", phi), ("aux", aux), ("out", out), ("out2", out2), ("loss", loss),
                       ("phi_next", phi_next)):
            if t is not None and t.device != f.device:
                raise ValueError(f"{where}: `{key}` is on {t.device}, `f` on {f.device}")
        for key, t in (("out", out), ("out2", out2), ("loss", loss), ("phi_next", phi_next)):
            if t is not None and not t.is_contiguous():
                raise ValueError(f"{where}: `{key}` must be contiguous (it is written in place)")
        f, phi = f.contiguous(), phi.contiguous()
        aux = aux.contiguous() if aux is not None else None
        if out is None:
            out = torch.empty(out_shape, dtype=torch.float32, device=f.device)
        if op in ("objective", "pull_vjp") and out2 is None:
            out2 = torch.empty((C,) + d0, dtype=torch.float32, device=f.device)
        if op == "objective" and loss is None:
            loss = torch.zeros((), dtype=torch.float64, device=f.device)
        if op not in ("objective", "pull_vjp"):
            out2 = None
        if op != "objective":
            loss = None
        rec = dict(f=f, phi=phi, aux=aux, out=out, out2=out2, loss=loss, phi_next=phi_next)
        recs.append(rec)
        s = arr[i]
        s.f, s.phi, s.aux, s.out, s.out2, s.loss = (_p(rec[k]) for k in ("f", "phi", "aux", "out", "out2", "loss"))
        s.phi_next, s.step = _p(phi_next), float(v.get("step", 0.0))
        for q in range(3):
            s.d0[q] = int(d0[2 - q])
            s.d1[q] = int(d1[2 - q])
            s.bnd[q] = bnd[0][q]
            s.deg[q] = deg[q]
        s.nbnd = len(bnd) if len(bnd) > 1 else 0
        if len(bnd) > 1:
            for c, t in enumerate(bnd):
                for q in range(3):
                    s.chan_bnd[3 * c + q] = t[q]
        s.nchan, s.ext = C, int(v.get("ext", 1))
    return recs, arr


class BatchPlan:
    """One op over a list of volumes (dicts with keys f, phi, deg, bnd, ext, aux, dim and
    optionally the outputs out, out2, loss)."""

    def __init__(self, op: str, volumes: list[dict]):
        self.handle = None
        self.vols, self._arr = _volume_records(op, volumes)
        self.L = _sigs()
        self.op = op
        h = ct.c_void_p()
        _lib.check(self.L.pp_batch_plan_create(OPS[op], ct.cast(self._arr, ct.c_void_p), len(volumes), ct.byref(h)),
                   "pp_batch_plan_create")
        self.handle = h

    def run(self, stream=None):
        s = stream if stream is not None else torch.cuda.current_stream()
        _lib.check(self.L.pp_batch_run(self.handle, ct.c_void_p(s.cuda_stream)), "pp_batch_run")
        return self.vols

    def close(self):
        if self.handle:
            _lib.check(self.L.pp_batch_plan_destroy(self.handle), "pp_batch_plan_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# volumes per pp_batch_run_static call (the size of its fixed device table)
STATIC_MAX_VOLUMES = 256


def run_static(op: str, volumes: list[dict], stream=None):
    """One op over a list of volumes through pp_batch_run_static: the results of a
    ``BatchPlan`` of the same volumes, without the plan's device allocation (the library
    keeps the descriptor table in a fixed per-device array of 256 volumes). Pass every
    output (out, out2, loss) to keep the call free of allocations, e.g. under graph
    capture. Calls on one stream are ordered; calls that overlap on one GPU race on the
    shared table. Returns the per-volume records (keep them until the stream has run)."""
    recs, arr = _volume_records(op, volumes)
    L = _sigs()
    s = stream if stream is not None else torch.cuda.current_stream()
    _lib.check(L.pp_batch_run_static(OPS[op], ct.cast(arr, ct.c_void_p), len(volumes), ct.c_void_p(s.cuda_stream)),
               "pp_batch_run_static")
    return recs


class Graph:
    """Capture ``fn(stream)`` into a HIP graph. ``fn`` issues library calls on ``stream``
    (it is also torch's current stream while ``fn`` runs) into buffers allocated before
    the capture: nothing may be allocated while capturing."""

    def __init__(self, fn, stream: "torch.cuda.Stream | None" = None):
        self.L = _sigs()
        self.stream = stream or torch.cuda.Stream()
        torch.cuda.synchronize()
        s = ct.c_void_p(self.stream.cuda_stream)
        _lib.check(self.L.pp_graph_begin(s), "pp_graph_begin")
        try:
            with torch.cuda.stream(self.stream):
                fn(self.stream)
        finally:
            x = ct.c_void_p()
            _lib.check(self.L.pp_graph_end(s, ct.byref(x)), "pp_graph_end")
        self.exec = x

    def launch(self):
        _lib.check(self.L.pp_graph_launch(self.exec, ct.c_void_p(self.stream.cuda_stream)), "pp_graph_launch")

    def close(self):
        if self.exec:
            _lib.check(self.L.pp_graph_destroy(self.exec), "pp_graph_destroy")
            self.exec = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
