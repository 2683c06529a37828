"""pp_pull_vjp (both cotangents of pull from one launch) against the FP64 oracle's push and
phi pullback, and against the separate GPU entry points (needs a GPU).

Bar: NaN masks identical; finite entries within 1e-5 norm-wise relative error of the FP64
oracle; the fused and the separate GPU calls agree to the same bar.
"""
import numpy as np
import pytest

import oracle
from helpers import S, check_close, identity, report, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


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


def _case(rng, d, C, maxdisp, out_frac=0.0):
    f = np.stack([smooth_field(rng, d, 2.0, 1.0)[0] for _ in range(C)])[None].astype(np.float32)
    phi = (identity(d) + smooth_field(rng, d, 4.0, maxdisp))[None].astype(np.float32)
    if out_frac:
        flat = phi.reshape(3, -1)
        idx = rng.choice(flat.shape[1], size=int(out_frac * flat.shape[1]), replace=False)
        flat[rng.integers(0, 3, size=idx.size), idx] = rng.choice([-20.0, 70.0], size=idx.size).astype(np.float32)
    delta = np.stack([smooth_field(rng, d, 2.0, 1.0)[0] for _ in range(C)])[None].astype(np.float32)
    return f, phi, delta


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("bnd", [0, 1, 2])
def test_pull_vjp_matches_oracle(deg, bnd, path):
    rng = np.random.default_rng(100 + 7 * deg + bnd)
    d = (30, 34, 38)
    f, phi, delta = _case(rng, d, 2, 5.0, out_frac=0.03)
    s, so = pp.Settings(deg, bnd, 1), S(deg, bnd, 1)
    df0, dphi = pp.pull_vjp(T(f), T(phi), T(delta), s)
    sc = float(np.abs(delta).max())
    e0 = check_close(H(df0), oracle.push(delta, phi, d, so, prec="f64"), f"vjp df0 deg {deg} bnd {bnd}", scale=sc)
    e1 = check_close(H(dphi), oracle.backward_phi(f, phi, delta, so), f"vjp dphi deg {deg} bnd {bnd}")
    report(f"vjp deg{deg} bnd{bnd} {path} df0 vs f64", e0)
    report(f"vjp deg{deg} bnd{bnd} {path} dphi vs f64", e1)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_pull_vjp_matches_separate_calls(deg, path):
    """Fused = pp_push + pp_pull_backward on the same inputs, with ext=0 (invalid samples
    give NaN in dphi and are skipped by push), non-finite entries of delta (skipped by
    push, NaN in dphi), per-channel boundary triples and a batch of two."""
    rng = np.random.default_rng(300 + deg)
    d = (24, 28, 33)
    f, phi, delta = _case(rng, d, 3, 6.0)
    f = np.concatenate([f, -0.5 * f])
    phi = np.concatenate([phi, phi[:, :, ::-1].copy()])
    delta = np.concatenate([delta, delta[:, :, :, ::-1].copy()])
    delta[0, 1, 3, 4, 5] = np.nan
    delta[1, 2, 7, 8, 9] = np.inf
    s = pp.Settings(deg, [(1, 1, 1), (0, 2, 1), (2, 2, 2)], 0)
    df0, dphi = pp.pull_vjp(T(f), T(phi), T(delta), s)
    ref_df0 = pp.push(T(delta), T(phi), d, s)
    ref_dphi = pp.pull_backward(T(f), T(phi), T(delta), s)
    check_close(H(df0), H(ref_df0), f"vjp df0 vs push deg {deg}", scale=1.0)
    check_close(H(dphi), H(ref_dphi), f"vjp dphi vs pull_backward deg {deg}")
    assert np.isnan(H(dphi)).any() and np.isfinite(H(df0)).all()


def test_pull_vjp_accumulates():
    rng = np.random.default_rng(5)
    d = (20, 21, 22)
    f, phi, delta = _case(rng, d, 1, 4.0)
    s = pp.Settings(3, 1, 1)
    base = torch.full(f.shape, 2.0, device=DEV)
    df0, _ = pp.pull_vjp(T(f), T(phi), T(delta), s, df0=base.clone(), accumulate=True)
    ref, _ = pp.pull_vjp(T(f), T(phi), T(delta), s)
    check_close(H(df0), H(ref) + 2.0, "vjp accumulate", scale=2.0)


def test_autograd_pull_uses_vjp_and_matches_oracle():
    """torch.autograd through pull with both inputs requiring gradients (the fused
    path) against the FP64 oracle's push and pullback."""
    rng = np.random.default_rng(9)
    d = (26, 25, 27)
    f, phi, delta = _case(rng, d, 2, 4.0)
    s, so = pp.Settings(2, 1, 1), S(2, 1, 1)
    tf = T(f).requires_grad_(True)
    tp = T(phi).requires_grad_(True)
    pp.pull(tf, tp, s).backward(T(delta))
    check_close(H(tf.grad), oracle.push(delta, phi, d, so, prec="f64"), "autograd df0", scale=1.0)
    check_close(H(tp.grad), oracle.backward_phi(f, phi, delta, so), "autograd dphi")


@pytest.mark.parametrize("tiling", [0, 1, 2])
@pytest.mark.parametrize("deg", [1, 2])
def test_pull_channels_differing_only_in_sign(tiling, deg):
    """Channels whose boundary codes differ only between reflect (1) and antisymmetric (2)
    run as one launch with per-channel signs: the per-sample signed pull (trilinear, where
    per-sample kernels are forced) and the tiled kernels (the automatic choice), against the
    FP64 oracle, with out-of-FOV samples folding."""
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    pp._lib.set_option(pp._lib.PP_OPT_TILING, tiling)
    try:
        rng = np.random.default_rng(40 + deg + 10 * tiling)
        d = (21, 26, 19)
        f, phi, _ = _case(rng, d, 3, 6.0, out_frac=0.05)
        bnd = [(2, 1, 1), (1, 2, 1), (1, 1, 2)]
        s, so = pp.Settings(deg, bnd, 1), S(deg, bnd, 1)
        got = H(pp.pull(T(f), T(phi), s))
        e = check_close(got, oracle.pull(f, phi, so, prec="f64"), f"signed pull deg {deg} tiling {tiling}")
        report(f"signed pull deg{deg} tiling{tiling} vs f64", e)
    finally:
        pp._lib.set_option(pp._lib.PP_OPT_TILING, old)


@pytest.mark.parametrize("deg", [1, 3, (1, 3, 0)])
def test_every_op_with_channels_differing_only_in_sign(deg):
    """Every op on channels whose boundary codes differ only between reflect (1) and
    antisymmetric (2) -- the sliding boundaries of a displacement field: one launch with
    per-channel signs on the tiled kernels of an isotropic degree (an anisotropic one runs
    once per channel), with samples outside the field of view taking the kernels' global
    paths, against the FP64 oracle."""
    rng = np.random.default_rng(500 + (deg if isinstance(deg, int) else 9))
    d = (22, 25, 20)
    f, phi, delta = _case(rng, d, 3, 6.0, out_frac=0.04)
    bnd = [(2, 1, 1), (1, 2, 1), (1, 1, 2)]
    s, so = pp.Settings(deg, bnd, 1), S(deg, bnd, 1)
    g3 = rng.standard_normal((1, 3, 3) + d).astype(np.float32)
    tag = f"signed channels deg {deg}"
    sc = float(np.abs(delta).max())
    e = {}
    e["pull_grad"] = check_close(H(pp.pull_grad(T(f), T(phi), s)), oracle.pull_grad(f, phi, so, prec="f64"),
                                 f"{tag} pull_grad", scale=float(np.abs(f).max()))
    e["pull_hess"] = check_close(H(pp.pull_hess(T(f), T(phi), s)), oracle.pull_hess(f, phi, so, prec="f64"),
                                 f"{tag} pull_hess", scale=float(np.abs(f).max()))
    e["push"] = check_close(H(pp.push(T(delta), T(phi), d, s)), oracle.push(delta, phi, d, so, prec="f64"),
                            f"{tag} push", scale=sc)
    e["push_grad"] = check_close(H(pp.push_grad(T(g3), T(phi), d, s)), oracle.push_grad(g3, phi, d, so, prec="f64"),
                                 f"{tag} push_grad", scale=float(np.abs(g3).max()))
    e["backward"] = check_close(H(pp.pull_backward(T(f), T(phi), T(delta), s)),
                                oracle.backward_phi(f, phi, delta, so), f"{tag} backward")
    # pull_grad's pullback: dphi_b = sum_{a,c} delta_{a,c} H_{a,b,c}
    hs = oracle.pull_hess(f, phi, so, prec="f64")  # (B, C, 3, 3, ...)
    ref = np.einsum("bca...,bcad...->bd...", g3.astype(np.float64), hs)
    e["grad_backward"] = check_close(H(pp.pull_grad_backward(T(f), T(phi), T(g3), s)), ref, f"{tag} grad_backward")
    df0, dphi = pp.pull_vjp(T(f), T(phi), T(delta), s)
    e["vjp df0"] = check_close(H(df0), oracle.push(delta, phi, d, so, prec="f64"), f"{tag} vjp df0", scale=sc)
    e["vjp dphi"] = check_close(H(dphi), oracle.backward_phi(f, phi, delta, so), f"{tag} vjp dphi")
    loss, dphi, df = pp.objective_grad(T(f), T(phi), T(delta), s)
    r = oracle.pull(f, phi, so, prec="f64") - delta
    e["objective df"] = check_close(H(df), oracle.push(r, phi, d, so, prec="f64"), f"{tag} objective df",
                                    scale=float(np.abs(r).max()))
    e["objective dphi"] = check_close(H(dphi), oracle.backward_phi(f, phi, r, so), f"{tag} objective dphi")
    assert abs(float(loss) - 0.5 * np.sum(r[np.isfinite(r)] ** 2)) <= 1e-5 * 0.5 * np.sum(r[np.isfinite(r)] ** 2)
    for k, v in e.items():
        report(f"{tag} {k} vs f64", v)


@pytest.mark.parametrize("case", ["deg1", "deg3", "sliding", "per-axis", "deg0", "mixed-codes", "per-sample",
                                  "accumulate"])
def test_objective_grad_step_equals_objective_then_axpy(case):
    """pp_objective_grad_step against pp_objective_grad followed by pp_axpy, bit for bit (NaN
    where dL/dphi is NaN): the fused tiled launch (shared codes, codes differing in sign only,
    a degree per axis) and the fallback (degree 0, codes that differ otherwise, per-sample
    kernels forced), into a separate buffer and in place."""
    import ctypes as ct

    from pushpull import _lib

    rng = np.random.default_rng(700 + len(case))
    d = (22, 26, 30)
    f, phi, g = _case(rng, d, 3, 5.0, out_frac=0.03)
    deg, bnd, ext = {"deg1": (1, 1, 1), "deg3": (3, 1, 0), "sliding": (1, [(2, 1, 1), (1, 2, 1), (1, 1, 2)], 1),
                     "per-axis": ((1, 3, 0), 1, 1), "deg0": (0, 1, 1), "mixed-codes": (2, [(1, 1, 1), (0, 2, 1), (1, 1, 1)], 1),
                     "per-sample": (3, 1, 1), "accumulate": (2, 1, 1)}[case]
    s = pp.Settings(deg, bnd, ext)
    step = -0.05
    acc = case == "accumulate"
    L = _lib.lib()
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    if case == "per-sample":
        pp._lib.set_option(pp._lib.PP_OPT_TILING, 0)
    try:
        base = torch.full(f.shape, 0.25, device=DEV)
        loss_r, dphi_r, df_r = pp.objective_grad(T(f), T(phi), T(g), s, df0=base.clone() if acc else None,
                                                 accumulate=acc)
        ref = T(phi)
        _lib.check(L.pp_axpy(ct.c_void_p(ref.data_ptr()), ct.c_void_p(dphi_r.data_ptr()), step, ref.numel(),
                             ct.c_void_p(torch.cuda.current_stream().cuda_stream)), "axpy")
        ph = T(phi)
        loss, dphi, df, nxt = pp.objective_grad_step(T(f), ph, T(g), step, s, df0=base.clone() if acc else None,
                                                      accumulate=acc)
        ph2 = T(phi)
        _, _, _, same_buf = pp.objective_grad_step(T(f), ph2, T(g), step, s, phi_next=ph2)
        torch.cuda.synchronize()
    finally:
        pp._lib.set_option(pp._lib.PP_OPT_TILING, old)

    def same(a, b):
        na, nb = torch.isnan(a), torch.isnan(b)
        return torch.equal(na, nb) and torch.equal(a[~na], b[~nb])

    assert same(dphi, dphi_r), case
    assert same(nxt, ref), f"{case}: step into a separate buffer differs from pp_axpy"
    assert same_buf is ph2 and same(ph2, ref), f"{case}: step in place differs from pp_axpy"
    check_close(H(df), H(df_r), f"{case} df", scale=1.0)
    assert abs(float(loss) - float(loss_r)) <= 1e-6 * abs(float(loss_r)), case
    assert torch.equal(ph, T(phi)), f"{case}: phi changed although phi_next was a separate buffer"
