"""Anisotropic degrees (a degree per axis, src/pushpull_general.jl:27-30) on the LDS-tiled
kernels (needs a GPU).

The tiled kernels are compiled per degree; a triple whose degrees differ runs on the
kernel of the largest, each axis's taps placed in that kernel's window with the window's
other taps weighted 0. Checked here for every op against the FP64 oracle (which evaluates
each axis with its own degree), on the tiled path and on the per-sample path, and for the
case where that placement could differ from the reference: a non-finite image value that
only the window's zero-weighted taps reach must not reach the output (the reference never
reads it), while one inside an axis's own support gives the reference's NaN.
"""
import numpy as np
import pytest

import oracle
from helpers import S, check_close, identity, report, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
DEGS = [(1, 3, 0), (3, 1, 2), (0, 0, 2), (4, 2, 1), (2, 2, 1)]


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(params=["tiled", "per-sample"])
def path(request):
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    old_t = pp._lib.get_option(pp._lib.PP_OPT_USE_TUNED)
    pp._lib.set_option(pp._lib.PP_OPT_TILING, 2 if request.param == "tiled" else 0)
    pp._lib.set_option(pp._lib.PP_OPT_USE_TUNED, 0)
    yield request.param
    pp._lib.set_option(pp._lib.PP_OPT_TILING, old)
    pp._lib.set_option(pp._lib.PP_OPT_USE_TUNED, old_t)


def _case(rng, d, C, maxdisp):
    f = np.stack([smooth_field(rng, d, 2.0, 1.0)[0] for _ in range(C)])[None].astype(np.float32)
    phi = (identity(d) + smooth_field(rng, d, 4.0, maxdisp))[None].astype(np.float32)
    g = np.stack([smooth_field(rng, d, 2.0, 1.0)[0] for _ in range(C)])[None].astype(np.float32)
    return f, phi, g


@pytest.mark.parametrize("deg", DEGS)
@pytest.mark.parametrize("bnd", [0, 1, 2])
def test_aniso_every_op_vs_f64(deg, bnd, path):
    rng = np.random.default_rng(900 + 10 * sum(deg) + bnd)
    d = (36, 40, 44)  # (Z, Y, X)
    f, phi, g = _case(rng, d, 2, 5.0)
    g3 = rng.standard_normal((1, 2, 3) + d).astype(np.float32)
    s, so = pp.Settings(deg, bnd, 1), S(deg, bnd, 1)
    tag = f"aniso deg{deg} bnd{bnd} {path}"
    sc = float(np.abs(g).max())
    e = {}
    e["pull"] = check_close(H(pp.pull(T(f), T(phi), s)), oracle.pull(f, phi, so, prec="f64"), f"{tag} pull")
    e["pull_grad"] = check_close(H(pp.pull_grad(T(f), T(phi), s)), oracle.pull_grad(f, phi, so, prec="f64"),
                                 f"{tag} pull_grad", scale=float(np.abs(f).max()))
    e["pull_hess"] = check_close(H(pp.pull_hess(T(f), T(phi), s)), oracle.pull_hess(f, phi, so, prec="f64"),
                                 f"{tag} pull_hess", scale=float(np.abs(f).max()))
    e["push"] = check_close(H(pp.push(T(g), T(phi), d, s)), oracle.push(g, phi, d, so, prec="f64"), f"{tag} push",
                            scale=sc)
    e["push_grad"] = check_close(H(pp.push_grad(T(g3), T(phi), d, s)), oracle.push_grad(g3, phi, d, so, prec="f64"),
                                 f"{tag} push_grad", scale=float(np.abs(g3).max()))
    e["backward"] = check_close(H(pp.pull_backward(T(f), T(phi), T(g), s)), oracle.backward_phi(f, phi, g, so),
                                f"{tag} backward")
    df0, dphi = pp.pull_vjp(T(f), T(phi), T(g), s)
    e["vjp df0"] = check_close(H(df0), oracle.push(g, phi, d, so, prec="f64"), f"{tag} vjp df0", scale=sc)
    e["vjp dphi"] = check_close(H(dphi), oracle.backward_phi(f, phi, g, so), f"{tag} vjp dphi")
    for k, v in e.items():
        report(f"{tag} {k} vs f64", v)


@pytest.mark.parametrize("deg", [(1, 3, 0), (3, 1, 2)])
def test_aniso_non_finite_image_values(deg, path):
    """NaN and Inf image values: the output's NaN mask and finite values match the FP64
    oracle (reference semantics) -- a value reached only by a zero-weighted window tap
    stays out of the output."""
    rng = np.random.default_rng(950 + sum(deg))
    d = (24, 28, 32)
    f, phi, _ = _case(rng, d, 1, 3.0)
    flat = f.reshape(-1)
    idx = rng.choice(flat.size, size=12, replace=False)
    flat[idx[:6]] = np.nan
    flat[idx[6:]] = np.inf
    s, so = pp.Settings(deg, 1, 1), S(deg, 1, 1)
    got = H(pp.pull(T(f), T(phi), s))
    ref = oracle.pull(f, phi, so, prec="f64")
    bad = ~np.isfinite(ref)
    assert 0 < bad.sum() < bad.size // 4
    assert np.array_equal(np.isnan(got), np.isnan(ref)), (int(np.isnan(got).sum()), int(np.isnan(ref).sum()))
    assert np.array_equal(np.isinf(got), np.isinf(ref))
    m = np.isfinite(ref)
    err = float(np.linalg.norm(got[m].astype(np.float64) - ref[m]) / np.linalg.norm(ref[m]))
    assert err < 1e-5, err
    report(f"aniso non-finite image deg{deg} {path} pull vs f64 (finite entries)", err)
    gotg = H(pp.pull_grad(T(f), T(phi), s))
    refg = oracle.pull_grad(f, phi, so, prec="f64")
    assert np.array_equal(np.isnan(gotg), np.isnan(refg))


@pytest.mark.parametrize("cap", [2048, 6770, 13596])
@pytest.mark.parametrize("op", ["pull", "push", "pull_vjp", "objective"])
def test_aniso_box_budgets(op, cap):
    """Each LDS box budget on the anisotropic kernels (a small one clips boxes and sends
    samples through the per-axis global path) against the FP64 oracle."""
    rng = np.random.default_rng(970)
    d = (40, 48, 56)
    f, phi, g = _case(rng, d, 1, 6.0)
    deg = (3, 1, 2)
    s, so = pp.Settings(deg, 1, 1), S(deg, 1, 1)
    L = pp._lib
    keys = (L.PP_OPT_TILING, L.PP_OPT_BOX_CAP, L.PP_OPT_USE_TUNED)
    old = [L.get_option(k) for k in keys]
    try:
        L.set_option(L.PP_OPT_TILING, 2)
        L.set_option(L.PP_OPT_BOX_CAP, cap)
        L.set_option(L.PP_OPT_USE_TUNED, 0)
        sc = float(np.abs(g).max())
        tag = f"aniso budget {cap} {op}"
        if op == "pull":
            e = check_close(H(pp.pull(T(f), T(phi), s)), oracle.pull(f, phi, so, prec="f64"), tag)
        elif op == "push":
            e = check_close(H(pp.push(T(g), T(phi), d, s)), oracle.push(g, phi, d, so, prec="f64"), tag, scale=sc)
        elif op == "pull_vjp":
            df, dphi = pp.pull_vjp(T(f), T(phi), T(g), s)
            check_close(H(df), oracle.push(g, phi, d, so, prec="f64"), tag + " df", scale=sc)
            e = check_close(H(dphi), oracle.backward_phi(f, phi, g, so), tag + " dphi")
        else:
            _, dphi, df = pp.objective_grad(T(f), T(phi), T(g), s)
            r = oracle.pull(f, phi, so, prec="f64") - g
            check_close(H(df), oracle.push(r, phi, d, so, prec="f64"), tag + " df", scale=float(np.abs(r).max()))
            e = check_close(H(dphi), oracle.backward_phi(f, phi, r, so), tag + " dphi")
        report(f"{tag} vs f64", e)
    finally:
        for k, v in zip(keys, old):
            L.set_option(k, v)


def test_aniso_tuner_records_its_own_key(tmp_path):
    """The autotuner times tiled candidates for a degree triple too, under a key of its own
    (not the isotropic degree's), and the tuned call matches the oracle."""
    from pushpull import tune as pt

    rng = np.random.default_rng(980)
    d = (40, 48, 56)
    f, phi, g = _case(rng, d, 1, 5.0)
    s, so = pp.Settings((1, 3, 0), 1, 1), S((1, 3, 0), 1, 1)
    pt.clear()
    v = pt.tune("push", T(g), T(phi), s, dim=d)
    cands = pt.results()
    assert len(cands) == 13 and any(c["variant"] > 0 for c in cands), cands
    path = str(tmp_path / "tune.txt")
    pt.save(path)
    lines = [l.split() for l in open(path) if not l.startswith("#")]
    assert len(lines) == 1 and int(lines[0][3]) == v
    assert int(lines[0][1]) not in (1, 2, 3, 4, 5)  # not an isotropic degree's key
    got = H(pp.push(T(g), T(phi), d, s))
    check_close(got, oracle.push(g, phi, d, so, prec="f64"), "aniso tuned push", scale=float(np.abs(g).max()))
    pt.clear()
