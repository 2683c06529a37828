"""Differentiable pull / pull_grad / push: torch.autograd.Function analogues of the
ChainRulesCore rrules in src/pushpull_pb.jl:11-51.

  pull(f0, phi)      -> df0 = push(delta, phi, size(f0));  dphi = sum_c delta_c grad pull(f0)_c
  pull_grad(f0, phi) -> df0 = push_grad(delta, ...);       dphi = sum_{a,c} delta_{a,c} H_{a,.,c}
  push(f1, phi, d)   -> df1 = pull(delta, phi);            dphi = sum_c f1_c grad pull(delta)_c

The phi pullbacks are single fused kernels (pp_pull_backward / pp_pull_grad_backward):
the 3C-channel pull_grad the Julia rule materialises before contracting
(src/pushpull_pb.jl:113-117) is never written to memory.
"""
from __future__ import annotations

import torch

from . import ops
from .settings import Settings


def _spatial(t):
    return tuple(t.shape[-3:])


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
        if ctx.needs_input_grad[0]:
            df0 = ops.push_raw(delta, phi, _spatial(f0), ctx.sett).reshape(f0.shape)
        if ctx.needs_input_grad[1]:
            dphi = ops.pull_backward(f0.reshape(_as5(f0)), phi.reshape(_as5(phi)),
                                     delta.reshape(_as5(delta)), ctx.sett).reshape(phi.shape)
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
        if ctx.needs_input_grad[0]:
            df0 = ops.push_grad(delta, phi, _spatial(f0), ctx.sett).reshape(f0.shape)
        if ctx.needs_input_grad[1]:
            f5 = f0.reshape(_as5(f0))
            d6 = delta.reshape((f5.shape[0], f5.shape[1], 3) + _spatial(delta))
            dphi = ops.pull_grad_backward(f5, phi.reshape(_as5(phi)), d6, ctx.sett).reshape(phi.shape)
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
        if ctx.needs_input_grad[0]:
            df1 = ops.pull_raw(delta, phi, ctx.sett).reshape(f1.shape)
        if ctx.needs_input_grad[1]:
            dphi = ops.pull_backward(delta.reshape(_as5(delta)), phi.reshape(_as5(phi)),
                                     f1.reshape(_as5(f1)), ctx.sett).reshape(phi.shape)
        return df1, dphi, None, None


def _as5(t):
    s = tuple(t.shape)
    while len(s) < 5:
        s = (1,) + s
    return s


def _wants_grad(*ts):
    return torch.is_grad_enabled() and any(isinstance(t, torch.Tensor) and t.requires_grad for t in ts)


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
