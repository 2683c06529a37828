"""HIP kernels against the oracle and the reference's golden outputs (needs a GPU).

Bar: NaN masks identical; finite entries within 1e-5 norm-wise relative error of the
FP32 reference (float atomics in push make the summation order nondeterministic, which
that tolerance covers), and within 1e-5 of the FP64 oracle.
"""
import numpy as np
import pytest

import oracle
from helpers import S, case_settings, identity, rel_err, report, same_nan_mask, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-5


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(params=["tiled", "per-sample"])
def path(request):
    """Run a test through the LDS-tiled kernels and through the one-work-item-per-sample
    kernels (PP_OPT_TILING); the tiled kernels cover degrees 1-3 of pull, push,
    push_grad, the pull pullback and the fused objective."""
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    old_t = pp._lib.get_option(pp._lib.PP_OPT_USE_TUNED)
    pp._lib.set_option(pp._lib.PP_OPT_TILING, 2 if request.param == "tiled" else 0)
    pp._lib.set_option(pp._lib.PP_OPT_USE_TUNED, 0)
    yield request.param
    pp._lib.set_option(pp._lib.PP_OPT_TILING, old)
    pp._lib.set_option(pp._lib.PP_OPT_USE_TUNED, old_t)


@pytest.fixture(params=[0, 1, 2])
def tile_shape(request):
    old = pp._lib.get_option(pp._lib.PP_OPT_TILE_SHAPE)
    pp._lib.set_option(pp._lib.PP_OPT_TILE_SHAPE, request.param)
    yield request.param
    pp._lib.set_option(pp._lib.PP_OPT_TILE_SHAPE, old)


def _sett(c):
    return pp.Settings(tuple(int(v) for v in c["deg"]), tuple(int(v) for v in c["bnd"]), int(c["ext"]))


def _check(got, ref, what, tol=TOL, scale=None):
    """NaN masks equal; ||got - ref|| <= tol*||ref|| over finite entries, plus (when
    `scale`, the largest input magnitude, is given) an absolute floor at FP32 rounding
    level, 1e-6*sqrt(N)*scale, for outputs that are mathematically zero (e.g. degree-1
    Hessians on one-voxel grids, where both sides hold rounding noise only)."""
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    assert same_nan_mask(got, ref), f"{what}: NaN mask differs"
    m = np.isfinite(ref)
    num = float(np.linalg.norm(got[m].astype(np.float64) - ref[m]))
    den = float(np.linalg.norm(ref[m].astype(np.float64)))
    floor = 1e-6 * np.sqrt(max(1, int(m.sum()))) * scale if scale is not None else 0.0
    assert num <= tol * den + floor, f"{what}: rel err {num / den if den else num:.3g}"
    return num / den if den else num


def _scale(a):
    a = np.asarray(a)
    f = a[np.isfinite(a)]
    return float(np.abs(f).max()) if f.size else 0.0


def _check_all(results):
    """results: list of (what, got, ref[, scale]); report every failing case at once."""
    bad = []
    for r in results:
        what, got, ref = r[:3]
        try:
            _check(got, ref, what, scale=r[3] if len(r) > 3 else None)
        except AssertionError as ex:
            bad.append(str(ex))
    assert not bad, f"{len(bad)} failing: " + "; ".join(bad[:12])


@pytest.mark.parametrize("op", ["pull", "pullg", "pullh", "push", "pushg"])
def test_golden_tier_b(golden, op, path):
    results = []
    for i, c in enumerate(golden):
        s = _sett(c)
        d0 = c["f0"].shape
        f0, phi = T(c["f0"]), T(c["phi"])
        if op == "pull":
            got, ref = H(pp.pull(f0, phi, s)), c["pull"]
        elif op == "pullg":
            got, ref = H(pp.pull_grad(f0, phi, s)), c["pullg"]
        elif op == "pullh":
            got, ref = H(pp.pull_hess(f0, phi, s)).reshape(c["pullh"].shape), c["pullh"]
        elif op == "push":
            got, ref = H(pp.push(T(c["f1"]), phi, d0, s)), c["push"]
        else:  # no usable reference output (pushpull.c:144): the oracle's adjoint
            got = H(pp.push_grad(T(c["g1"]), phi, d0, s))
            ref = oracle.push_grad(c["g1"], c["phi"], d0, case_settings(c), prec="f64")[0, 0]
        src = {"push": c["f1"], "pushg": c["g1"]}.get(op, c["f0"])
        results.append((f"case {i} deg {c['deg'].tolist()} {op}", got, ref, _scale(src)))
    _check_all(results)


def test_golden_tier_a(golden):
    """Host arrays go through the symbol-compatible entry points (pull, pullg, ...)."""
    for i, c in enumerate(golden[::3]):
        s = _sett(c)
        d0 = c["f0"].shape
        sf, s1, sg = _scale(c["f0"]), _scale(c["f1"]), _scale(c["g1"])
        _check(pp.pull(c["f0"], c["phi"], s), c["pull"], f"A{i} pull", scale=sf)
        _check(pp.pull_grad(c["f0"], c["phi"], s), c["pullg"], f"A{i} pullg", scale=sf)
        _check(pp.pull_hess(c["f0"], c["phi"], s).reshape(c["pullh"].shape), c["pullh"], f"A{i} pullh", scale=sf)
        _check(pp.push(c["f1"], c["phi"], d0, s), c["push"], f"A{i} push", scale=s1)
        ref = oracle.push_grad(c["g1"], c["phi"], d0, case_settings(c), prec="f64")[0, 0]
        _check(pp.push_grad(c["g1"], c["phi"], d0, s), ref, f"A{i} pushg", scale=sg)


def test_tier_a_push_accumulates_into_caller_array():
    import ctypes as ct

    rng = np.random.default_rng(3)
    d0 = (6, 5, 4)
    phi = identity((5, 5, 5)) + rng.standard_normal((3, 5, 5, 5)).astype(np.float32)
    f1 = rng.standard_normal((5, 5, 5)).astype(np.float32)
    base = rng.standard_normal(d0).astype(np.float32)
    out = base.copy()
    L = pp._lib.lib()
    dd0 = np.array(d0[::-1], np.uint64)
    bnd = np.array([1, 1, 1], np.int32)
    dp = np.array([3, 3, 3], np.uint64)
    P = lambda a: ct.c_void_p(a.ctypes.data)  # noqa: E731
    L.push(P(out), P(phi), P(f1), P(dd0), f1.size, P(bnd), P(dp), 1)
    ref = base.astype(np.float64) + oracle.push(f1, phi, d0, S(2, 1, 1), prec="f64")[0, 0]
    assert rel_err(out, ref) < TOL


def _volume_case(rng, d, sigma_f, maxdisp, out_frac=0.0):
    f = smooth_field(rng, d, sigma_f, 1.0)[0]
    f = (f / f.std()).astype(np.float32)
    phi = identity(d) + smooth_field(rng, d, 4.0, maxdisp)
    if out_frac:
        flat = phi.reshape(3, -1)
        n = flat.shape[1]
        idx = rng.choice(n, size=int(out_frac * n), replace=False)
        ax = rng.integers(0, 3, size=idx.size)
        dims = np.array(d[::-1], np.float32)[ax]
        lo = rng.uniform(-0.5, 0.0, size=idx.size) * dims
        hi = dims + 1 + rng.uniform(0.0, 0.5, size=idx.size) * dims
        flat[ax, idx] = np.where(rng.random(idx.size) < 0.5, lo, hi).astype(np.float32)
    return f, phi


@pytest.mark.parametrize("deg", [1, 2, 3])
@pytest.mark.parametrize("bnd", [0, 1])
def test_config2_pull_push_adjoint(deg, bnd, path):
    """192^3, smoothed noise, displacement up to 8 voxels, 5% sent outside the FOV."""
    rng = np.random.default_rng(100 + 10 * deg + bnd)
    d = (192, 192, 192)
    f, phi = _volume_case(rng, d, 2.0, 8.0, out_frac=0.05)
    g = rng.standard_normal(d).astype(np.float32)
    s = pp.Settings(deg, bnd, 1)
    so = S(deg, bnd, 1)
    tf, tphi, tg = T(f), T(phi), T(g)
    f1 = H(pp.pull(tf, tphi, s))
    r32 = oracle.pull(f, phi, so)[0, 0]
    r64 = oracle.pull(f, phi, so, prec="f64")[0, 0]
    tag = f"config2 deg{deg} bnd{bnd} {path}"
    report(f"{tag} pull vs f32 reference", _check(f1, r32, "pull vs f32 ref"))
    report(f"{tag} pull vs f64", _check(f1, r64, "pull vs f64"))
    f0 = H(pp.push(tg, tphi, d, s))
    p64 = oracle.push(g, phi, d, so, prec="f64")[0, 0]
    report(f"{tag} push vs f64", _check(f0, p64, "push vs f64"))
    lhs = float(torch.sum(T(f1).double() * tg.double()))
    rhs = float(torch.sum(tf.double() * T(f0).double()))
    report(f"{tag} adjoint <pull f, g> vs <f, push g>", abs(lhs - rhs) / (abs(lhs) + abs(rhs)))
    assert abs(lhs - rhs) / (abs(lhs) + abs(rhs)) < TOL
    # FP64 oracle adjointness (pins push as the exact adjoint of pull)
    lhs64 = float(np.sum(r64 * g.astype(np.float64)))
    rhs64 = float(np.sum(f.astype(np.float64) * p64))
    assert abs(lhs64 - rhs64) / (abs(lhs64) + abs(rhs64)) < 1e-12


@pytest.mark.parametrize("deg", [1, 2, 3])
@pytest.mark.parametrize("bnd", [0, 1])
def test_config2_ext0_nan_and_skip(deg, bnd, path):
    """192^3 as above with extrapolate=false: the 5% of samples sent outside the field of
    view pull to NaN (VOXOK, artifacts/C/pushpull.c:6-9, 120) and push skips them
    (pushpull.c:114); the NaN masks must match the reference exactly, the finite values
    its FP32 path and the FP64 oracle, and the adjoint identity holds over the valid
    samples."""
    rng = np.random.default_rng(400 + 10 * deg + bnd)
    d = (192, 192, 192)
    f, phi = _volume_case(rng, d, 2.0, 8.0, out_frac=0.05)
    g = rng.standard_normal(d).astype(np.float32)
    s = pp.Settings(deg, bnd, 0)
    so = S(deg, bnd, 0)
    tf, tphi, tg = T(f), T(phi), T(g)
    f1 = H(pp.pull(tf, tphi, s))
    r32 = oracle.pull(f, phi, so)[0, 0]
    r64 = oracle.pull(f, phi, so, prec="f64")[0, 0]
    nan_frac = float(np.isnan(r64).mean())
    assert 0.04 < nan_frac < 0.2, nan_frac  # the outside samples (and those past the -0.01 / d-0.99 limits)
    tag = f"config2 ext0 deg{deg} bnd{bnd} {path}"
    report(f"{tag} pull vs f32 reference", _check(f1, r32, "pull vs f32 ref"))
    report(f"{tag} pull vs f64", _check(f1, r64, "pull vs f64"))
    f0 = H(pp.push(tg, tphi, d, s))
    p64 = oracle.push(g, phi, d, so, prec="f64")[0, 0]
    report(f"{tag} push vs f64", _check(f0, p64, "push vs f64"))
    valid = np.isfinite(f1)
    lhs = float(np.sum(f1[valid].astype(np.float64) * g[valid].astype(np.float64)))
    rhs = float(np.sum(f.astype(np.float64) * f0.astype(np.float64)))
    report(f"{tag} adjoint over valid samples", abs(lhs - rhs) / (abs(lhs) + abs(rhs)))
    assert abs(lhs - rhs) / (abs(lhs) + abs(rhs)) < TOL


def test_config1_nan_mask_ext0(path):
    """64^3, U[0,1), trilinear, reflect, ext=0: the NaN mask must match exactly."""
    rng = np.random.default_rng(1)
    d = (64, 64, 64)
    f = rng.random(d, dtype=np.float32)
    phi = identity(d) + smooth_field(rng, d, 4.0, 3.0)
    s = pp.Settings(1, 1, 0)
    got = H(pp.pull(T(f), T(phi), s))
    ref = oracle.pull(f, phi, S(1, 1, 0))[0, 0]
    assert np.isnan(ref).sum() > 0
    report(f"config1 pull ext0 vs f32 reference {path}", _check(got, ref, "config 1 pull"))
    report(f"config1 pull ext0 vs f64 {path}", _check(got, oracle.pull(f, phi, S(1, 1, 0), prec="f64")[0, 0], "config 1 f64"))


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_pull_grad_backward_and_hess(deg, path):
    rng = np.random.default_rng(40 + deg)
    d = (40, 36, 44)
    f, phi = _volume_case(rng, d, 2.0, 5.0, out_frac=0.02)
    fc = np.stack([f, np.roll(f, 3, axis=0)])[None]  # (1, 2, ...)
    delta = rng.standard_normal((1, 2) + d).astype(np.float32)
    delta3 = rng.standard_normal((1, 2, 3) + d).astype(np.float32)
    s = pp.Settings(deg, [(1, 1, 1), (0, 2, 1)], 0)
    so = S(deg, [(1, 1, 1), (0, 2, 1)], 0)
    tf, tphi = T(fc), T(phi[None])
    _check(H(pp.pull_grad(tf, tphi, s)), oracle.pull_grad(fc, phi[None], so, prec="f64"), "pull_grad")
    _check(H(pp.pull_hess(tf, tphi, s)), oracle.pull_hess(fc, phi[None], so, prec="f64"), "pull_hess")
    _check(H(pp.pull_backward(tf, tphi, T(delta), s)), oracle.backward_phi(fc, phi[None], delta, so), "pull_backward")
    H64 = oracle.pull_hess(fc, phi[None], so, prec="f64")  # (B, C, 3 (b), 3 (a), ...)
    # sum_{a,c} delta[a,c] * H[a, b]  (src/pushpull_pb.jl:26-35)
    ref = np.zeros((1, 3) + d)
    for c in range(2):
        for a in range(3):
            for b in range(3):
                ref[0, b] += delta3[0, c, a] * H64[0, c, b, a]
    _check(H(pp.pull_grad_backward(tf, tphi, T(delta3), s)), ref, "pull_grad_backward")
    g3 = delta3
    _check(H(pp.push_grad(T(g3), tphi, d, s)), oracle.push_grad(g3, phi[None], d, so, prec="f64"), "push_grad")


def test_push_grad_is_adjoint_of_pull_grad(path):
    rng = np.random.default_rng(5)
    d = (48, 40, 32)
    f, phi = _volume_case(rng, d, 2.0, 6.0, out_frac=0.05)
    g = rng.standard_normal((3,) + d).astype(np.float32)
    for deg in (1, 2, 3, 4):
        s = pp.Settings(deg, (0, 1, 2), 1)
        lhs = float(torch.sum(pp.pull_grad(T(f), T(phi), s).double() * T(g).double()))
        rhs = float(torch.sum(T(f).double() * pp.push_grad(T(g), T(phi), d, s).double()))
        assert abs(lhs - rhs) / (abs(lhs) + abs(rhs)) < TOL, deg


def test_batches_channels_per_channel_bounds(path, tile_shape):
    rng = np.random.default_rng(9)
    d0, d1 = (20, 18, 22), (16, 19, 21)
    B, C = 2, 3
    f0 = rng.standard_normal((B, C) + d0).astype(np.float32)
    phi = np.stack([identity(d1) + rng.uniform(-3, 3, (3,) + d1).astype(np.float32) * 2 for _ in range(B)])
    f1 = rng.standard_normal((B, C) + d1).astype(np.float32)
    bnds = [(0, 1, 2), (1, 1, 1), (2, 0, 1)]
    for deg in [(2, 2, 2), (1, 3, 0)]:
        s = pp.Settings(deg, bnds, 1)
        so = S(deg, bnds, 1)
        _check(H(pp.pull(T(f0), T(phi), s)), oracle.pull(f0, phi, so, prec="f64"), f"pull {deg}")
        _check(H(pp.pull_grad(T(f0), T(phi), s)), oracle.pull_grad(f0, phi, so, prec="f64"), f"pullg {deg}")
        _check(H(pp.push(T(f1), T(phi), d0, s)), oracle.push(f1, phi, d0, so, prec="f64"), f"push {deg}")
        _check(H(pp.pull_backward(T(f0), T(phi), T(f1), s)), oracle.backward_phi(f0, phi, f1, so), f"bwd {deg}")


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_objective_grad_fused_matches_composition(deg, path, tile_shape):
    rng = np.random.default_rng(70 + deg)
    d = (36, 40, 44)
    f, phi = _volume_case(rng, d, 2.0, 6.0)
    g = smooth_field(rng, d, 2.0, 1.0)[:2][None]
    fc = np.stack([f, -f])[None]
    s = pp.Settings(deg, 1, 1)
    so = S(deg, 1, 1)
    loss, dphi, df0 = pp.objective_grad(T(fc), T(phi[None]), T(g), s)
    r = oracle.pull(fc, phi[None], so, prec="f64") - g
    _check(H(dphi), oracle.backward_phi(fc, phi[None], r, so), "objgrad dphi")
    _check(H(df0), oracle.push(r, phi[None], d, so, prec="f64"), "objgrad df0")
    assert abs(float(loss) - 0.5 * np.sum(r ** 2)) / (0.5 * np.sum(r ** 2)) < TOL


def test_errors_raise():
    s = pp.Settings(2, 1, 1)
    f = torch.zeros((4, 4, 4), device=DEV)
    phi = torch.zeros((3, 5, 5, 5), device=DEV)
    with pytest.raises(AssertionError):
        pp.pull(f, torch.zeros((2, 5, 5, 5), device=DEV), s)
    with pytest.raises(AssertionError):
        pp.push(torch.zeros((4, 4, 4), device=DEV), phi, (4, 4, 4), s)
    with pytest.raises(TypeError):
        pp.pull(f.double(), phi, s)


def test_golden_affine(golden_affine, path):
    """affine_pull / affine_push against the reference (pushpull.c:176-244), Tier B
    (device tensors) and Tier A (host arrays through pull_affine / push_affine)."""
    results = []
    for i, c in enumerate(golden_affine):
        s = _sett(c)
        d1, d0 = tuple(int(v) for v in c["d1"]), c["f0"].shape
        sf, s1 = _scale(c["f0"]), _scale(c["f1"])
        results.append((f"affine {i} pull B", H(pp.affine_pull(T(c["f0"]), c["aff"], d1, s)), c["pull_affine"], sf))
        results.append((f"affine {i} push B", H(pp.affine_push(T(c["f1"]), c["aff"], d0, s)), c["push_affine"], s1))
        results.append((f"affine {i} pull A", pp.affine_pull(c["f0"], c["aff"], d1, s), c["pull_affine"], sf))
        results.append((f"affine {i} push A", pp.affine_push(c["f1"], c["aff"], d0, s), c["push_affine"], s1))
    _check_all(results)


@pytest.mark.parametrize("deg", [1, 3])
def test_affine_at_size(deg, path):
    rng = np.random.default_rng(deg)
    d0, d1 = (96, 112, 128), (100, 90, 120)
    f = smooth_field(rng, d0, 2.0, 1.0)[0]
    A = np.eye(4)
    A[:3, :3] += rng.uniform(-0.15, 0.15, size=(3, 3))
    A[:3, 3] = rng.uniform(-5, 5, size=3)
    s = pp.Settings(deg, (1, 0, 2), 1)
    so = S(deg, (1, 0, 2), 1)
    A0 = pp.adjust_affine(A.astype(np.float32))
    _check(H(pp.affine_pull(T(f), A, d1, s)), oracle.affine_pull(f, A0, d1, so, prec="f64")[0, 0], "affine pull")
    g = rng.standard_normal(d1).astype(np.float32)
    _check(H(pp.affine_push(T(g), A, d0, s)), oracle.affine_push(g, A0, d0, so, prec="f64")[0, 0], "affine push")
