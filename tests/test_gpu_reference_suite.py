"""The reference's own gradient suite, test/runtests.jl:34-95, re-run on the GPU through the
torch.autograd mirror of its rrules (src/pushpull_pb.jl:11-51) (needs a GPU).

Geometry, settings and checks as there:
  * an 8 x 7 x 1 volume (Julia (X, Y, Z)), C = 2 channels, phi = id + randn (per voxel,
    not smooth), Settings(deg, (0, 1, 2), 1): circulant x, reflect y, antisymmetric z on
    the one-voxel z axis;
  * test_grad (test/utils.jl:3-25): sum((g_ad - g_fd)^2) / sum(g_ad^2) < 1e-3, g_fd by
    central differences with eps = 0.01 applied to one Float32 entry at a time, the loss
    evaluated by the same (GPU) functions in Float32;
  * degree 2 and 3: d/dphi and d/df of sum((pull(f1, phi) - f2)^2) and of
    sum((push(f1, phi, d) - f2)^2); degree 3: d/dphi of sum((pull_grad(f1, phi) - g2)^2)
    (the pull_grad rrule: pull_hess contracted with the upstream gradient);
  * a batch of 2 (Julia (X, Y, Z, C, B)) at degree 3, the four pull/push gradients;
  * the adjoint test |<f1, pull(f2, phi)> - <push(f1, phi, d), f2>| < 1e-5.
The reference draws its data from Julia's RNG; here it is seeded NumPy normal data of the
same shapes, over a few seeds.
"""
import numpy as np
import pytest

from helpers import report

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-3     # test/runtests.jl:44
ADJ_TOL = 1e-5  # test/runtests.jl:93
EPS = 0.01     # test/utils.jl:9
D = (1, 7, 8)  # Julia d = (8, 7, 1), in array order (Z, Y, X)
C = 2


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(DEV)


def _data(seed, batch=None):
    rng = np.random.default_rng(seed)
    lead = () if batch is None else (batch,)
    f1 = rng.standard_normal(lead + (C,) + D).astype(np.float32)
    f2 = rng.standard_normal(lead + (C,) + D).astype(np.float32)
    phi = rng.standard_normal(lead + (3,) + D).astype(np.float32)
    phi += pp.id(dim=D[::-1]).reshape((1,) * len(lead) + (3,) + D)
    return f1, f2, phi


def fd_ratio(fun, theta0):
    """test/utils.jl:3-25: autograd against central differences, one entry at a time."""
    th = _t(theta0).requires_grad_(True)
    fun(th).backward()
    g0 = th.grad.detach().double().cpu().numpy().ravel()
    th = _t(theta0)
    flat = th.view(-1)
    gfd = np.empty(flat.numel())
    with torch.no_grad():
        for i in range(flat.numel()):
            v = np.float32(flat[i].item())
            vp = np.float32(v + np.float32(EPS))  # theta[i] .+= eps in Float32
            flat[i] = float(vp)
            ep = float(fun(th))
            flat[i] = float(np.float32(vp - np.float32(2 * EPS)))  # theta[i] .-= 2eps
            em = float(fun(th))
            gfd[i] = (ep - em) / (2 * EPS)
            flat[i] = float(v)
    return float(((g0 - gfd) ** 2).sum() / (g0 ** 2).sum())


def _pull_push_cases(f1, f2, phi, s):
    tf2 = _t(f2)
    dim = D
    return {
        "pull/phi": (lambda th: ((pp.pull(_t(f1), th, s) - tf2) ** 2).sum(), phi),
        "pull/f": (lambda th: ((pp.pull(th, _t(phi), s) - tf2) ** 2).sum(), f1),
        "push/phi": (lambda th: ((pp.push(_t(f1), th, dim, s) - tf2) ** 2).sum(), phi),
        "push/f": (lambda th: ((pp.push(th, _t(phi), dim, s) - tf2) ** 2).sum(), f1),
    }


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("deg", [2, 3])
def test_runtests_pull_push_gradients(deg, seed):
    # test/runtests.jl:62-73
    f1, f2, phi = _data(seed)
    s = pp.Settings((deg,) * 3, (0, 1, 2), 1)
    for name, (fun, theta) in _pull_push_cases(f1, f2, phi, s).items():
        err = fd_ratio(fun, theta)
        report(f"runtests deg{deg} seed{seed} {name}", err)
        assert err < TOL, (deg, seed, name, err)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_runtests_pull_grad_gradient(seed):
    # test/runtests.jl:75-76: g2 = pull_grad(f2, phi); d/dphi of sum((pull_grad(f1, phi) - g2)^2)
    f1, f2, phi = _data(seed)
    s = pp.Settings((3, 3, 3), (0, 1, 2), 1)
    g2 = pp.pull_grad(_t(f2), _t(phi), s)
    err = fd_ratio(lambda th: ((pp.pull_grad(_t(f1), th, s) - g2) ** 2).sum(), phi)
    report(f"runtests pull_grad/phi seed{seed}", err)
    assert err < TOL, (seed, err)


@pytest.mark.parametrize("seed", [0, 1])
def test_runtests_batch_of_two(seed):
    # test/runtests.jl:79-91: batch size 2, degree 3
    f1, f2, phi = _data(seed, batch=2)
    s = pp.Settings((3, 3, 3), (0, 1, 2), 1)
    for name, (fun, theta) in _pull_push_cases(f1, f2, phi, s).items():
        err = fd_ratio(fun, theta)
        report(f"runtests batch2 seed{seed} {name}", err)
        assert err < TOL, (seed, name, err)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_runtests_adjoint(seed):
    # test/runtests.jl:93-95: |sum(f1 .* pull(f2, phi)) - sum(push(f1, phi, d) .* f2)| < 1e-5
    f1, f2, phi = _data(seed, batch=2)
    s = pp.Settings((3, 3, 3), (0, 1, 2), 1)
    a = float((_t(f1) * pp.pull(_t(f2), _t(phi), s)).sum())
    b = float((pp.push(_t(f1), _t(phi), D, s) * _t(f2)).sum())
    report(f"runtests adjoint seed{seed}", abs(a - b))
    assert abs(a - b) < ADJ_TOL, (a, b)
