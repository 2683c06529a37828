"""Batches of uneven volumes in one persistent launch, and HIP graph capture.

The reference only batches same-size volumes, one native call per (channel, batch) slice
(src/pushpull_cpu.jl:42-51). ``BatchPlan`` takes a list of volumes with their own sizes,
channel counts, degrees and deformations and runs one op over all of them in one launch
(``pp_batch_plan_create`` / ``pp_batch_run``). ``Graph`` captures any sequence of calls on a
side stream into a HIP graph (``pp_graph_*``) and replays it.

Per volume (C-order shapes, as everywhere in this package):
  f    image (C, Z0, Y0, X0) for pull-type ops; values (C, Z1, Y1, X1) for push,
       (C, 3, Z1, Y1, X1) for push_grad
  phi  (3, Z1, Y1, X1), 1-based
  aux  upstream gradient (pull_backward) or target (objective), (C, Z1, Y1, X1)
  deg  isotropic degree; bnd a boundary triple shared by the volume's channels; ext
"""
from __future__ import annotations

import ctypes as ct

import torch

from . import _lib

OPS = {"pull": 0, "push": 3, "push_grad": 4, "pull_backward": 5, "objective": 7, "pull_vjp": 8}


class PPVolume(ct.Structure):
    _fields_ = [
        ("f", ct.c_void_p), ("phi", ct.c_void_p), ("aux", ct.c_void_p), ("out", ct.c_void_p),
        ("out2", ct.c_void_p), ("loss", ct.c_void_p),
        ("d0", ct.c_int64 * 3), ("d1", ct.c_int64 * 3),
        ("nchan", ct.c_int32), ("deg", ct.c_int32), ("ext", ct.c_int32), ("bnd", ct.c_int32 * 3),
    ]


def _sigs():
    L = _lib.lib()
    L.pp_batch_plan_create.argtypes = [ct.c_int, ct.c_void_p, ct.c_int, ct.POINTER(ct.c_void_p)]
    L.pp_batch_plan_create.restype = ct.c_int
    L.pp_batch_run.argtypes = [ct.c_void_p, ct.c_void_p]
    L.pp_batch_run.restype = ct.c_int
    L.pp_batch_plan_destroy.argtypes = [ct.c_void_p]
    L.pp_batch_plan_destroy.restype = ct.c_int
    for n, a in (("pp_graph_begin", [ct.c_void_p]), ("pp_graph_end", [ct.c_void_p, ct.POINTER(ct.c_void_p)]),
                 ("pp_graph_launch", [ct.c_void_p, ct.c_void_p]), ("pp_graph_destroy", [ct.c_void_p])):
        fn = getattr(L, n)
        fn.argtypes = a
        fn.restype = ct.c_int
    return L


def _p(t):
    return ct.c_void_p(t.data_ptr()) if t is not None else ct.c_void_p(0)


class BatchPlan:
    """One op over a list of volumes (dicts with keys f, phi, deg, bnd, ext, aux)."""

    def __init__(self, op: str, volumes: list[dict]):
        self.L = _sigs()
        self.op = op
        self.vols = []
        arr = (PPVolume * len(volumes))()
        for i, v in enumerate(volumes):
            f = v["f"].contiguous()
            phi = v["phi"].contiguous()
            if f.dtype != torch.float32 or phi.dtype != torch.float32 or not f.is_cuda:
                raise TypeError("batched volumes are float32 device tensors")
            if phi.ndim != 4 or phi.shape[0] != 3:
                raise AssertionError("`ϕ` must be (3, Z, Y, X)")
            d1 = tuple(phi.shape[1:])
            C = f.shape[0]
            aux = v.get("aux")
            aux = aux.contiguous() if aux is not None else None
            if op in ("pull", "pull_backward", "objective", "pull_vjp"):
                d0 = tuple(f.shape[-3:])
            else:
                d0 = tuple(v["dim"])
            if op == "pull":
                shape = (C,) + d1
            elif op in ("push", "push_grad"):
                shape = (C,) + d0
            else:
                shape = (3,) + d1
            out = v.get("out")
            if out is None:
                out = torch.empty(shape, dtype=torch.float32, device=f.device)
            elif tuple(out.shape) != shape or not out.is_contiguous():
                raise AssertionError(f"`out` must be a contiguous {shape} tensor")
            out2 = loss = None
            if op in ("objective", "pull_vjp"):
                out2 = v.get("out2")
                out2 = torch.empty_like(f) if out2 is None else out2
            if op == "objective":
                loss = v.get("loss")
                loss = torch.zeros((), dtype=torch.float64, device=f.device) if loss is None else loss
            rec = dict(f=f, phi=phi, aux=aux, out=out, out2=out2, loss=loss)
            self.vols.append(rec)
            s = arr[i]
            s.f, s.phi, s.aux, s.out, s.out2, s.loss = (_p(rec[k]) for k in ("f", "phi", "aux", "out", "out2", "loss"))
            for q in range(3):
                s.d0[q] = int(d0[2 - q])
                s.d1[q] = int(d1[2 - q])
                s.bnd[q] = int(v.get("bnd", (1, 1, 1))[q])
            s.nchan, s.deg, s.ext = int(C), int(v["deg"]), int(v.get("ext", 1))
        self._arr = arr
        h = ct.c_void_p()
        _lib.check(self.L.pp_batch_plan_create(OPS[op], ct.cast(arr, ct.c_void_p), len(volumes), ct.byref(h)),
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
