"""The tiled scatters' fixed-point LDS cells at the limit of their range (needs a GPU).

Value scatters (push, and the df0 half of pull_vjp) accumulate 32-bit cells whose scale
is fixed per tile and channel from wmax^3 * sum|v| (wmax the largest 1-D B-spline
weight), so that no cell can leave int32. The case that reaches the bound: every sample
of the deformation on one knot, all with the same value. The centre cell then receives
wmax^3 * v from every sample of every tile, and every other cell gets its exact share
(at an integer position the 1-D weights are those of the spline at its knots). Closed
forms pin the result: the centre cell holds N * v * wmax^3 and the image sums to N * v
(partition of unity; N samples, interior position). A wrapped cell would miss both by
~2^32 quanta.

Also: random values of both signs on a deformation compressed to a few voxels (up to
2^10 contributions per cell, cancelling), against the FP64 oracle at the usual 1e-5 bar.
"""
import numpy as np
import pytest

import oracle
from helpers import S, check_close, report

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
# largest 1-D weight of the degree-n B-spline (its value at a knot)
WMAX = {1: 1.0, 2: 0.75, 3: 2.0 / 3.0, 4: 115.0 / 192.0}


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture
def tiled():
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    old_t = pp._lib.get_option(pp._lib.PP_OPT_USE_TUNED)
    pp._lib.set_option(pp._lib.PP_OPT_TILING, 2)
    pp._lib.set_option(pp._lib.PP_OPT_USE_TUNED, 0)
    yield
    pp._lib.set_option(pp._lib.PP_OPT_TILING, old)
    pp._lib.set_option(pp._lib.PP_OPT_USE_TUNED, old_t)


@pytest.mark.parametrize("op", ["push", "vjp"])
@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_all_samples_on_one_knot(deg, op, tiled):
    d1 = (32, 32, 32)  # 32 tiles of 1024 samples
    d0 = (20, 22, 24)
    pos = (11.0, 9.0, 13.0)  # (x, y, z), 1-based knots well inside the image
    phi = np.empty((1, 3) + d1, np.float32)
    for a in range(3):
        phi[0, a] = pos[a]
    v = 7.25
    f1 = np.full((1, 1) + d1, v, np.float32)
    s, so = pp.Settings(deg, 1, 1), S(deg, 1, 1)
    if op == "push":
        got = H(pp.push(T(f1), T(phi), d0, s))
    else:
        f0 = np.zeros((1, 1) + d0, np.float32)
        got, _ = pp.pull_vjp(T(f0), T(phi), T(f1), s)
        got = H(got)
    n = float(np.prod(d1))
    centre = got[0, 0, int(pos[2]) - 1, int(pos[1]) - 1, int(pos[0]) - 1]
    want = n * v * WMAX[deg] ** 3
    # (1e-5: the tiles' partial sums meet in float atomics; a wrapped cell misses by ~100%)
    assert abs(centre - want) <= 1e-5 * want, (centre, want)
    total = float(got.astype(np.float64).sum())
    assert abs(total - n * v) <= 1e-5 * n * v, (total, n * v)
    e = check_close(got, oracle.push(f1, phi, d0, so, prec="f64"), f"{op} one knot deg {deg}", scale=v)
    report(f"scatter cells one knot deg{deg} {op} vs f64", e)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_compressed_deformation_random_signs(deg, tiled):
    rng = np.random.default_rng(500 + deg)
    d1 = (32, 32, 16)
    d0 = (18, 19, 20)
    phi = np.empty((1, 3) + d1, np.float32)
    for a, c in enumerate((9.3, 10.6, 8.2)):
        phi[0, a] = c + rng.uniform(-1.5, 1.5, d1).astype(np.float32)
    f1 = rng.standard_normal((1, 2) + d1).astype(np.float32)
    f1[0, 1, 3, 4, 5] = 1.0e4  # one large value sets its tile's scale
    s, so = pp.Settings(deg, 1, 1), S(deg, 1, 1)
    got = H(pp.push(T(f1), T(phi), d0, s))
    e = check_close(got, oracle.push(f1, phi, d0, so, prec="f64"), f"compressed push deg {deg}", scale=1.0)
    report(f"scatter cells compressed deg{deg} push vs f64", e)
