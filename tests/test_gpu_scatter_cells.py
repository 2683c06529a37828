"""The tiled scatters' fixed-point LDS cells (needs a GPU).

Value scatters (push, the df0 half of pull_vjp, the dL/df0 half of the objective) and
push_grad accumulate 64-bit fixed-point cells whose scale is fixed per tile and channel
from the tile's largest |value| (a quantum of at most 2^-49 of it), so that no cell can
overflow and every cell is exact to that quantum before its one rounding to float.

* Per-cell error: log-normal magnitudes spanning 1e-4..1e4 with random signs and a 1e6
  spike in a few tiles, degrees 1-4, against the FP64 oracle cell by cell:
  |got - ref64| <= k * 2^-24 * push64(|v|) + 64 * 2^-50 * M, where push64(|v|) is the
  cell's mass (the error model of the reference's float atomics: each addition into a
  cell rounds at 2^-24 of a partial sum no larger than it), k = 8 or, where larger, 1.5x
  the reference's own FP32 worst case on the same data, and M the largest |v| within
  reach of the cell (the fixed-point quantum of its tiles, half of it per contribution).
  The worst error relative to the mass, over the cells where the first term dominates,
  is reported beside the reference's own sequential FP32 accumulation on the same data.
* The range limit: every sample of the deformation on one knot, all with the same value.
  The centre cell then receives wmax^3 * v from every sample of every tile (wmax the
  largest 1-D B-spline weight), every other cell its exact share; closed forms pin the
  centre cell (N * v * wmax^3) and the image sum (N * v, partition of unity).
* Random values of both signs on a deformation compressed to a few voxels (up to 2^10
  contributions per cell, cancelling), against the FP64 oracle at the usual 1e-5 bar.
"""
import numpy as np
import pytest

import oracle
from helpers import S, check_close, identity, report, smooth_field

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


def _lognormal_with_spikes(rng, d1):
    mag = np.exp(rng.normal(0.0, np.log(1e4) / 3.0, d1)).clip(1e-4, 1e4)
    v = (np.where(rng.random(d1) < 0.5, -1.0, 1.0) * mag).astype(np.float32)
    # a 1e6 spike in three tiles (16 x 8 x 8 tiles of the phi grid)
    for z, y, x in ((3, 4, 5), (20, 13, 40), (35, 40, 17)):
        v[z, y, x] = 1.0e6
    return v


@pytest.mark.parametrize("op", ["push", "vjp", "objective"])
@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_per_cell_error_lognormal_with_spikes(deg, op, tiled):
    rng = np.random.default_rng(700 + deg)
    d = (40, 48, 64)  # (Z, Y, X): the image and the phi grid
    phi = (identity(d) + smooth_field(rng, d, 3.0, 4.0))[None].astype(np.float32)
    v = _lognormal_with_spikes(rng, d)[None, None]
    s, so = pp.Settings(deg, 1, 1), S(deg, 1, 1)
    zeros = np.zeros((1, 1) + d, np.float32)
    if op == "push":
        got = H(pp.push(T(v), T(phi), d, s))
    elif op == "vjp":
        got, _ = pp.pull_vjp(T(zeros), T(phi), T(v), s)
        got = H(got)
    else:  # r = pull(0, phi) - g = v
        _, _, got = pp.objective_grad(T(zeros), T(phi), T(-v), s)
        got = H(got)
    ref = oracle.push(v, phi, d, so, prec="f64")
    mass = oracle.push(np.abs(v), phi, d, so, prec="f64")
    # largest |v| within reach of each cell: its 16 x 8 x 8 tile and the neighbouring ones
    # (displacement <= 4 plus the footprint stays below one tile)
    from scipy.ndimage import maximum_filter

    tmax = np.abs(v[0, 0]).reshape(d[0] // 8, 8, d[1] // 8, 8, d[2] // 16, 16).max(axis=(1, 3, 5))
    tmax = maximum_filter(tmax, size=3, mode="nearest")
    reach = np.repeat(np.repeat(np.repeat(tmax, 8, 0), 8, 1), 16, 2)[None, None]
    err = np.abs(got.astype(np.float64) - ref)
    inv = 2.0 ** 24 / np.where(mass > 0, mass, 1.0)
    # the reference's own sequential FP32 accumulation on the same data (pushpull.c, through
    # its bit-exact restatement): its error in units of 2^-24 * mass reaches 5-13 at degrees
    # 1-3 (its n float additions into a cell), and ~90 at degree 4, where the FP32
    # evaluation of the quartic weights dominates
    ratio_ref = np.abs(oracle.push(v, phi, d, so, prec="f32").astype(np.float64) - ref) * inv
    # (x1.5: the weights' FP32 rounding differs between the two evaluation orders)
    k = max(8.0, 1.5 * float(ratio_ref.max()))
    bound = k * 2.0 ** -24 * mass + 64.0 * 2.0 ** -50 * reach
    assert (err <= bound).all(), (int((err > bound).sum()), float((err / bound).max()), k)
    # relative to the cell's own mass, where the quantum floor does not dominate
    rel = mass >= 2.0 ** -20 * reach
    report(f"scatter cells per-cell {op} deg{deg} max err / (2^-24 mass)", float((err * inv)[rel].max()))
    report(f"scatter cells per-cell {op} deg{deg} reference FP32 max err / (2^-24 mass)", float(ratio_ref[rel].max()))
    report(f"scatter cells per-cell {op} deg{deg} cells below 2^-20 of reach (absolute floor)", float((~rel).mean()))
