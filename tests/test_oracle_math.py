"""Pins the oracle by what the mathematics fixes, independently of the reference build:

* B-spline weights against a brute-force Cox-de Boor recursion in FP64 (Unser 1999),
  the derivative identities B'_n(u) = B_{n-1}(u+1/2) - B_{n-1}(u-1/2) and
  B''_n(u) = B_{n-2}(u+1) - 2 B_{n-2}(u) + B_{n-2}(u-1), partition of unity and
  linear reproduction;
* boundary folding against explicit periodic / mirrored index lists;
* pull known answers: constants, identity deformation (sampled B-spline kernels),
  integer circular shifts, linear ramps;
* pull_grad / pull_hess against central differences of pull, in FP64;
* adjointness of (pull, push) and (pull_grad, push_grad) in FP64;
* the knot conversion saturating like x86-64 for coordinates beyond the int64 range.
"""
import numpy as np
import pytest

import oracle
from helpers import S, identity


def bspline(n, u):
    """Centred B-spline of degree n by the Cox-de Boor recursion (FP64)."""
    u = np.asarray(u, np.float64)
    if n == 0:
        return np.where(np.abs(u) < 0.5, 1.0, np.where(np.abs(u) == 0.5, 0.5, 0.0))
    h = (n + 1) / 2.0
    return ((u + h) * bspline(n - 1, u + 0.5) + (h - u) * bspline(n - 1, u - 0.5)) / n


def dbspline(n, u):
    return bspline(n - 1, u + 0.5) - bspline(n - 1, u - 0.5)


def hbspline(n, u):
    return bspline(n - 2, u + 1) - 2 * bspline(n - 2, u) + bspline(n - 2, u - 1)


def _canonical_t(dp, rng, n=200):
    lo, hi = dp / 2 - 1, dp / 2
    t = rng.uniform(lo, hi, size=n)
    return t[(t > lo + 1e-9) & (t < hi - 1e-9)]


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_weights_match_cox_de_boor(deg):
    rng = np.random.default_rng(deg)
    dp = deg + 1
    for t in _canonical_t(dp, rng):
        w, dw, hw = oracle.weights(dp, float(t))
        k = np.arange(dp)
        np.testing.assert_allclose(w, bspline(deg, t - k), atol=1e-14)
        if deg >= 1:
            np.testing.assert_allclose(dw, dbspline(deg, t - k), atol=1e-13)
        if deg >= 2:
            np.testing.assert_allclose(hw, hbspline(deg, t - k), atol=1e-12)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_weights_partition_and_linear_reproduction(deg):
    rng = np.random.default_rng(10 + deg)
    dp = deg + 1
    k = np.arange(dp)
    for t in _canonical_t(dp, rng):
        w, dw, hw = oracle.weights(dp, float(t))
        assert abs(w.sum() - 1) < 1e-14
        assert abs((k * w).sum() - t) < 1e-13          # sum_k (o+k) w_k = x
        assert abs(dw.sum()) < 1e-13
        assert abs((k * dw).sum() - 1) < 1e-13         # d/dx of the above
        if deg >= 2:
            assert abs(hw.sum()) < 1e-12
            assert abs((k * hw).sum()) < 1e-12


def test_weights_degree0_and_deg1_derivative_constants():
    w, dw, hw = oracle.weights(1, 0.3)
    assert w.tolist() == [1.0] and dw.tolist() == [0.0] and hw.tolist() == [0.0]
    w, dw, hw = oracle.weights(2, 0.25)
    assert w.tolist() == [0.75, 0.25] and dw.tolist() == [-1.0, 1.0] and hw.tolist() == [0.0, 0.0]
    _, _, hw = oracle.weights(3, 1.2)
    assert hw.tolist() == [1.0, -2.0, 1.0]


@pytest.mark.parametrize("d", [1, 2, 3, 5, 8])
def test_fold_against_explicit_lists(d):
    period = list(range(d))
    mirror = list(range(d)) + list(range(d - 1, -1, -1))
    sign = [1] * d + [-1] * d
    for i in range(-5 * d - 3, 5 * d + 3):
        assert oracle.fold(0, i, d) == (period[i % d], 1)
        assert oracle.fold(1, i, d) == (mirror[i % (2 * d)], 1)
        assert oracle.fold(2, i, d) == (mirror[i % (2 * d)], sign[i % (2 * d)])
        assert oracle.fold(7, i, d) == oracle.fold(1, i, d)  # unknown codes reflect


def test_fold_int64_extremes():
    imin = -(2 ** 63)  # the knot of a coordinate beyond the int64 range
    for d in (1, 7, 13, 256):
        b, s = oracle.fold(0, imin, d)
        assert b == imin % d and s == 1  # Python's % is the mathematical modulus
        b, s = oracle.fold(2, imin + 1, d)
        m = (imin + 1) % (2 * d)
        assert (b, s) == ((m, 1) if m < d else (2 * d - 1 - m, -1))


def test_knot_matches_ceil_and_saturates():
    for x, dp in [(0.0, 2), (0.3, 4), (-0.3, 4), (2.5, 3), (1e6 + 0.5, 2), (-7.25, 5)]:
        assert oracle.knot(x, dp) == int(np.ceil(np.float32(x) - np.float32(0.5 * dp)))
    assert oracle.knot(1e19, 2) == -(2 ** 63)
    assert oracle.knot(-2e19, 4) == -(2 ** 63)
    assert oracle.knot(9.2e18, 2) == int(np.float32(9.2e18))


# ---- pull known answers (FP64 oracle) ------------------------------------------------
def _rand_phi(rng, d1, d0, spread):
    phi = identity(d1)
    phi += rng.uniform(-spread, spread, size=phi.shape).astype(np.float32)
    return phi


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bnd", [0, 1])
def test_pull_of_constant_is_constant(deg, bnd):
    rng = np.random.default_rng(deg * 3 + bnd)
    d = (6, 7, 5)
    phi = _rand_phi(rng, d, d, 9.0)
    out = oracle.pull(np.full(d, 2.5, np.float32), phi, S(deg, bnd, 1), prec="f64")
    np.testing.assert_allclose(out, 2.5, rtol=1e-13)


@pytest.mark.parametrize("deg,kern", [
    (0, [1.0]), (1, [1.0]), (2, [1 / 8, 3 / 4, 1 / 8]), (3, [1 / 6, 2 / 3, 1 / 6]),
    (4, [1 / 384, 19 / 96, 115 / 192, 19 / 96, 1 / 384]),
])
def test_pull_at_identity_is_sampled_bspline_convolution(deg, kern):
    rng = np.random.default_rng(deg)
    d = (7, 6, 9)
    f = rng.standard_normal(d).astype(np.float32)
    out = oracle.pull(f, identity(d), S(deg, 0, 1), prec="f64")[0, 0]
    ref = f.astype(np.float64)
    r = len(kern) // 2
    for ax in range(3):
        acc = np.zeros_like(ref)
        for k, w in enumerate(kern):
            acc += w * np.roll(ref, k - r, axis=ax)
        ref = acc
    np.testing.assert_allclose(out, ref, atol=1e-12)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_pull_integer_shift_circulant_is_roll(deg):
    rng = np.random.default_rng(30 + deg)
    d = (5, 6, 7)
    f = rng.standard_normal(d).astype(np.float32)
    shift = (2, -3, 4)  # (z, y, x)
    phi = identity(d)
    phi[0] += shift[2]
    phi[1] += shift[1]
    phi[2] += shift[0]
    a = oracle.pull(f, phi, S(deg, 0, 1), prec="f64")[0, 0]
    b = oracle.pull(f, identity(d), S(deg, 0, 1), prec="f64")[0, 0]
    np.testing.assert_allclose(a, np.roll(b, (-shift[0], -shift[1], -shift[2]), axis=(0, 1, 2)), atol=1e-12)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_pull_reproduces_linear_ramp_inside(deg):
    d = (12, 12, 12)
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) for n in d), indexing="ij")
    f = (0.5 * x - 0.25 * y + 2.0 * z + 1.0).astype(np.float32)
    rng = np.random.default_rng(deg)
    phi = (4.0 + rng.uniform(0, 4, size=(3, 4, 4, 4))).astype(np.float32)  # interior samples
    out = oracle.pull(f, phi, S(deg, 1, 1), prec="f64")[0, 0]
    xs = phi.astype(np.float64) - 1
    np.testing.assert_allclose(out, 0.5 * xs[0] - 0.25 * xs[1] + 2.0 * xs[2] + 1.0, atol=1e-10)


# ---- derivatives by central differences (FP64) -------------------------------------
def _fd_points(rng, n, d):
    """Points on a 2^-8 lattice so +-2^-12 perturbations are exact in float32 and never
    cross a knot (knots sit on integers or half-integers)."""
    base = rng.integers(3, min(d) - 3, size=(3, n)).astype(np.float64)
    frac = rng.integers(4, 60, size=(3, n)) / 128.0
    return (base + frac).astype(np.float32).reshape(3, n, 1, 1)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
def test_pull_grad_is_derivative_of_pull(deg):
    rng = np.random.default_rng(50 + deg)
    d = (11, 12, 13)
    f = rng.standard_normal(d).astype(np.float32)
    s = S(deg, (0, 1, 2), 1)
    phi = _fd_points(rng, 25, d)
    g = oracle.pull_grad(f, phi, s, prec="f64")[0, 0]
    h = 2.0 ** -12
    for a in range(3):
        pp_, pm = phi.copy(), phi.copy()
        pp_[a] += np.float32(h)
        pm[a] -= np.float32(h)
        fd = (oracle.pull(f, pp_, s, prec="f64") - oracle.pull(f, pm, s, prec="f64"))[0, 0] / (2 * h)
        np.testing.assert_allclose(g[a], fd, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("deg", [2, 3, 4])
def test_pull_hess_is_derivative_of_pull_grad(deg):
    rng = np.random.default_rng(60 + deg)
    d = (11, 12, 13)
    f = rng.standard_normal(d).astype(np.float32)
    s = S(deg, (1, 2, 0), 1)
    phi = _fd_points(rng, 25, d)
    H = oracle.pull_hess(f, phi, s, prec="f64")[0, 0]  # (3 (b), 3 (a), ...)
    h = 2.0 ** -12
    for b in range(3):
        pp_, pm = phi.copy(), phi.copy()
        pp_[b] += np.float32(h)
        pm[b] -= np.float32(h)
        fd = (oracle.pull_grad(f, pp_, s, prec="f64") - oracle.pull_grad(f, pm, s, prec="f64"))[0, 0] / (2 * h)
        for a in range(3):
            np.testing.assert_allclose(H[b, a], fd[a], rtol=1e-5, atol=1e-6)


# ---- adjointness (FP64) ------------------------------------------------------------
@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bnd", [(0, 0, 0), (1, 1, 1), (2, 2, 2), (0, 1, 2)])
def test_push_is_adjoint_of_pull(deg, bnd):
    rng = np.random.default_rng(deg * 7 + sum(bnd))
    d0, d1 = (6, 5, 7), (5, 6, 4)
    f = rng.standard_normal(d0)
    g = rng.standard_normal(d1)
    phi = _rand_phi(rng, d1, d0, 6.0)
    s = S(deg, bnd, 1)
    lhs = np.sum(oracle.pull(f, phi, s, prec="f64") * g)
    rhs = np.sum(f * oracle.push(g, phi, d0, s, prec="f64"))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs) + 1)


@pytest.mark.parametrize("deg", [1, 2, 3, 4])
@pytest.mark.parametrize("bnd", [(0, 0, 0), (2, 1, 0)])
def test_push_grad_is_adjoint_of_pull_grad(deg, bnd):
    rng = np.random.default_rng(100 + deg)
    d0, d1 = (6, 5, 7), (5, 6, 4)
    f = rng.standard_normal(d0)
    G = rng.standard_normal((3,) + d1)
    phi = _rand_phi(rng, d1, d0, 6.0)
    s = S(deg, bnd, 1)
    lhs = np.sum(oracle.pull_grad(f, phi, s, prec="f64")[0, 0] * G)
    rhs = np.sum(f * oracle.push_grad(G, phi, d0, s, prec="f64")[0, 0])
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs) + 1)


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bnd", [0, 1])
def test_push_of_ones_at_identity_is_ones(deg, bnd):
    d = (6, 7, 5)
    out = oracle.push(np.ones(d), identity(d), d, S(deg, bnd, 1), prec="f64")
    np.testing.assert_allclose(out, 1.0, atol=1e-13)


def test_invalid_samples_are_nan_and_skipped():
    d = (4, 4, 4)
    f = np.ones(d, np.float32)
    phi = identity(d)
    phi[0, 0, 0, 0] = np.nan
    phi[1, 1, 1, 1] = 10.0       # outside with ext = 0
    out = oracle.pull(f, phi, S(1, 1, 0))[0, 0]
    assert np.isnan(out[0, 0, 0]) and np.isnan(out[1, 1, 1]) and np.isfinite(out).sum() == 62
    g = np.ones(d, np.float32)
    g[2, 2, 2] = np.inf           # non-finite pushed values are skipped
    p = oracle.push(g, identity(d), d, S(1, 1, 0), prec="f64")[0, 0]
    assert np.isfinite(p).all() and abs(p.sum() - 63) < 1e-12


# ---- affine (src/pushpull_cpu.jl:268-362, pushpull.c:176-244) ----------------------
def test_adjust_affine_known_answers():
    import pushpull as pp

    I = np.eye(4, dtype=np.float32)
    np.testing.assert_array_equal(pp.adjust_affine(I), np.eye(4, dtype=np.float32)[:3])
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = (2.5, -1.0, 4.0)           # 1-based shift -> same 0-based shift
    np.testing.assert_array_equal(pp.adjust_affine(T)[:, 3], np.float32([2.5, -1.0, 4.0]))
    S2 = np.diag(np.float32([2, 3, 0.5, 1]))  # x1 = 2 i1  ->  x0 = 2 i0 + 1
    np.testing.assert_array_equal(pp.adjust_affine(S2)[:, 3], np.float32([1.0, 2.0, -0.5]))
    rng = np.random.default_rng(0)
    for _ in range(20):
        A = rng.standard_normal((3, 4)).astype(np.float32)
        np.testing.assert_array_equal(pp.adjust_affine(A), oracle.adjust_affine(A))


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_affine_identity_and_shift(deg):
    import pushpull as pp

    rng = np.random.default_rng(deg)
    d = (6, 7, 8)
    f = rng.standard_normal(d).astype(np.float32)
    A0 = pp.adjust_affine(np.eye(4, dtype=np.float32))
    a = oracle.affine_pull(f, A0, d, S(deg, 0, 1), prec="f64")[0, 0]
    b = oracle.pull(f, identity(d), S(deg, 0, 1), prec="f64")[0, 0]
    np.testing.assert_allclose(a, b, atol=1e-13)
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = (2, -1, 3)  # (x, y, z) shift, circulant -> roll
    a = oracle.affine_pull(f, pp.adjust_affine(T), d, S(deg, 0, 1), prec="f64")[0, 0]
    np.testing.assert_allclose(a, np.roll(b, (-3, 1, -2), axis=(0, 1, 2)), atol=1e-12)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_affine_push_is_adjoint(deg):
    import pushpull as pp

    rng = np.random.default_rng(10 + deg)
    d0, d1 = (7, 6, 9), (5, 8, 6)
    A = np.eye(4)
    A[:3, :3] += rng.uniform(-0.3, 0.3, size=(3, 3))
    A[:3, 3] = rng.uniform(-2, 2, size=3)
    A0 = pp.adjust_affine(A.astype(np.float32))
    f = rng.standard_normal(d0)
    g = rng.standard_normal(d1)
    s = S(deg, (0, 1, 2), 1)
    lhs = np.sum(oracle.affine_pull(f, A0, d1, s, prec="f64") * g)
    rhs = np.sum(f * oracle.affine_push(g, A0, d0, s, prec="f64"))
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))
