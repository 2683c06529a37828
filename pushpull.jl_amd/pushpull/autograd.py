"""Differentiable pull / pull_grad / push: torch.autograd.Function analogues of the
ChainRulesCore rrules in src/pushpull_pb.jl:11-51.

  pull(f0, phi)      -> df0 = push(delta, phi, size(f0));  dphi = sum_c delta_c grad pull(f0)_c
  pull_grad(f0, phi) -> df0 = push_grad(delta, ...);       dphi = sum_{a,c} delta_{a,c} H_{a,.,c}
  push(f1, phi, d)   -> df1 = pull(delta, phi);            dphi = sum_c f1_c grad pull(delta)_c

The phi pullbacks are single fused kernels (pp_pull_backward / pp_pull_grad_backward):
the 3C-channel pull_grad the Julia rule materialises before contracting
(src/pushpull_pb.jl:11-20) is never written to memory. When both gradients of pull are
wanted, one launch computes them (pp_pull_vjp: each tile of f0 is staged once, phi and
delta are read once). For pull_grad and push, when both gradients are wanted on the GPU,
the pullback runs on a second stream while the other gradient runs on the current one
(the two are independent).
"""
from __future__ import annotations

import torch

from . import ops
from .settings import Settings


def _spatial(t):
    return tuple(t.shape[-3:])


_side_streams = {}


def _side_stream(device):
    """A second stream per device (default priority) for the φ pullback."""
    s = _side_streams.get(device.index)
    if s is None:
        s = _side_streams[device.index] = torch.cuda.Stream(device=device)
    return s


def _both(first, second, tensors):
    """Run first() on the current stream and second() on a side stream at the same time;
    the two are independent ops of one rrule (they write different arrays). Returns both
    results, the second recorded on the current stream, which waits for the side stream."""
    if not (tensors[0].is_cuda and torch.cuda.is_available()):
        return first(), second()
    dev = tensors[0].device
    cur = torch.cuda.current_stream(dev)
    side = _side_stream(dev)
    side.wait_stream(cur)
    with torch.cuda.stream(side):
        b = second()
    a = first()
    cur.wait_stream(side)
    b.record_stream(cur)
    for t in tensors:  # inputs read on the side stream
        t.record_stream(side)
    return a, b


class _Pull(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, phi, sett):
        ctx.sett = sett
        ctx.save_for_backward(f0, phi)
        return ops.pull_raw(f0, phi, sett)

    @staticmethod
    def backward(ctx, delta):
        f0, phi = ctx.saved_tensors
        delta = delta.contiguous()
        df0 = dphi = None

        def d_f0():
            return ops.push_raw(delta, phi, _spatial(f0), ctx.sett).reshape(f0.shape)

        def d_phi():
            return ops.pull_backward(f0.reshape(_as5(f0)), phi.reshape(_as5(phi)),
                                     delta.reshape(_as5(delta)), ctx.sett).reshape(phi.shape)

        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            if f0.is_cuda:  # both cotangents from one launch (pp_pull_vjp)
                df0, dphi = ops.pull_vjp(f0.reshape(_as5(f0)), phi.reshape(_as5(phi)),
                                         delta.reshape(_as5(delta)), ctx.sett)
                df0, dphi = df0.reshape(f0.shape), dphi.reshape(phi.shape)
            else:
                df0, dphi = _both(d_f0, d_phi, (f0, phi, delta))
        elif ctx.needs_input_grad[0]:
            df0 = d_f0()
        elif ctx.needs_input_grad[1]:
            dphi = d_phi()
        return df0, dphi, None


class _PullGrad(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f0, phi, sett):
        ctx.sett = sett
        ctx.save_for_backward(f0, phi)
        return ops.pull_grad_raw(f0, phi, sett)

    @staticmethod
    def backward(ctx, delta):
        f0, phi = ctx.saved_tensors
        delta = delta.contiguous()
        df0 = dphi = None

        def d_f0():
            return ops.push_grad(delta, phi, _spatial(f0), ctx.sett).reshape(f0.shape)

        def d_phi():
            f5 = f0.reshape(_as5(f0))
            d6 = delta.reshape((f5.shape[0], f5.shape[1], 3) + _spatial(delta))
            return ops.pull_grad_backward(f5, phi.reshape(_as5(phi)), d6, ctx.sett).reshape(phi.shape)

        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            df0, dphi = _both(d_f0, d_phi, (f0, phi, delta))
        elif ctx.needs_input_grad[0]:
            df0 = d_f0()
        elif ctx.needs_input_grad[1]:
            dphi = d_phi()
        return df0, dphi, None


class _Push(torch.autograd.Function):
    @staticmethod
    def forward(ctx, f1, phi, dim, sett):
        ctx.sett = sett
        ctx.save_for_backward(f1, phi)
        return ops.push_raw(f1, phi, dim, sett)

    @staticmethod
    def backward(ctx, delta):
        f1, phi = ctx.saved_tensors
        delta = delta.contiguous()
        df1 = dphi = None

        def d_f1():
            return ops.pull_raw(delta, phi, ctx.sett).reshape(f1.shape)

        def d_phi():
            return ops.pull_backward(delta.reshape(_as5(delta)), phi.reshape(_as5(phi)),
                                     f1.reshape(_as5(f1)), ctx.sett).reshape(phi.shape)

        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            df1, dphi = _both(d_f1, d_phi, (f1, phi, delta))
        elif ctx.needs_input_grad[0]:
            df1 = d_f1()
        elif ctx.needs_input_grad[1]:
            dphi = d_phi()
        return df1, dphi, None, None


def _as5(t):
    s = tuple(t.shape)
    while len(s) < 5:
        s = (1,) + s
    return s


def _wants_grad(*ts):
    """True when the call must record an autograd node. The rrules run on the GPU only: a
    CPU tensor that requires grad is refused rather than silently dropping the graph."""
    if not torch.is_grad_enabled() or not any(isinstance(t, torch.Tensor) and t.requires_grad for t in ts):
        return False
    if not all(t.is_cuda for t in ts if isinstance(t, torch.Tensor)):
        raise TypeError("differentiable pull/pull_grad/push take device tensors; "
                        "a CPU tensor that requires grad is not supported (detach it or move it to the GPU)")
    return True


def pull(f0, phi, sett=None):
    """pull(f₀, ϕ, sett) (src/pushpull_cpu.jl:17, ext/pushpull_gpu.jl:15); differentiable
    for device tensors (src/pushpull_pb.jl:11-20)."""
    sett = sett or Settings()
    if _wants_grad(f0, phi):
        return _Pull.apply(f0, phi, sett)
    return ops.pull_raw(f0, phi, sett)


def pull_grad(f0, phi, sett=None):
    """pull_grad(f₀, ϕ, sett) (src/pushpull_cpu.jl:69); differentiable (src/pushpull_pb.jl:26-35)."""
    sett = sett or Settings()
    if _wants_grad(f0, phi):
        return _PullGrad.apply(f0, phi, sett)
    return ops.pull_grad_raw(f0, phi, sett)


def push(f1, phi, dim, sett=None):
    """push(f₁, ϕ, d, sett) (src/pushpull_cpu.jl:165); differentiable (src/pushpull_pb.jl:41-51).
    ``dim`` is the output volume size in array order (Z, Y, X)."""
    sett = sett or Settings()
    if _wants_grad(f1, phi):
        return _Push.apply(f1, phi, tuple(dim), sett)
    return ops.push_raw(f1, phi, dim, sett)
