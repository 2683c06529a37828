"""PushPull (MI355X): the B-spline pull/push sampling path of PushPull.jl on gfx950.

Same names and argument order as the Julia module (src/PushPull.jl:19):
``pull, push, pull_grad, Settings, id`` plus the unexported ``push_grad`` and
``pull_hess``. Compute runs only in libpushpull_hip.so (hand-written HIP kernels);
there is no CPU fallback.
"""
from .settings import Settings, id  # noqa: A001
from .autograd import pull, pull_grad, push
from .ops import push_grad, pull_hess, pull_backward, pull_grad_backward, pull_vjp, objective_grad
from .ops import objective_grad_step
from .affine import affine_pull, affine_push, adjust_affine
from .compositions import compose, compose_, Exp
from .multigrid import push_count, affine_push_count, invert_def_rhs, restrict, prolong
from .regulariser import registration_operator, reduce2fit, sparsify, vel2mom, vel2mom_, relax_, SparseKernel
from . import _lib

__all__ = [
    "Settings", "id", "pull", "push", "pull_grad", "push_grad", "pull_hess",
    "pull_backward", "pull_grad_backward", "pull_vjp", "objective_grad", "objective_grad_step", "affine_pull", "affine_push", "adjust_affine",
    "compose", "compose_", "Exp", "push_count", "affine_push_count", "invert_def_rhs", "restrict", "prolong",
    "registration_operator", "reduce2fit", "sparsify", "vel2mom", "vel2mom_", "relax_", "SparseKernel",
]


def library_path() -> str:
    return _lib.LIB_PATH
