"""Autotuner on the GPU: every variant runs, the choice is recorded and used, and the
result of the tuned call still matches the oracle."""
import numpy as np
import pytest

import oracle
from helpers import S, identity, rel_err, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")
from pushpull import tune as pt  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("op", ["pull", "pull_grad", "push", "push_grad", "pull_backward", "objective", "pull_vjp"])
def test_tune_picks_a_variant_and_results_stay_exact(op, tmp_path):
    rng = np.random.default_rng(3)
    d = (48, 40, 56)
    f = smooth_field(rng, d, 2.0, 1.0)[0]
    phi = identity(d) + smooth_field(rng, d, 4.0, 5.0)
    g = rng.standard_normal(d).astype(np.float32)
    g3 = rng.standard_normal((3,) + d).astype(np.float32)
    dev = "cuda:0"
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    s = pp.Settings(3, 1, 1)
    pt.clear()
    src = {"push": T(g), "push_grad": T(g3)}.get(op, T(f))
    v = pt.tune(op, src, T(phi), s, aux=T(g), dim=d)
    assert 0 <= v <= 3
    path = str(tmp_path / "tune.txt")
    pt.save(path)
    lines = [l for l in open(path) if not l.startswith("#")]
    assert len(lines) == 1 and int(lines[0].split()[3]) == v
    cands = pt.results()
    best = min(cands, key=lambda c: c["ms"])
    assert best["variant"] == v and int(lines[0].split()[4]) == best["cap"], (cands, lines)
    # per sample, then 3 tile shapes x (the default budget + the other budgets of 6..3 WG/CU)
    caps = {c["cap"] for c in cands}
    assert len(cands) in (13, 16) and 0 in caps and caps <= {0, 6770, 8136, 10184, 13596}, cands
    pt.clear()
    pt.load(path)
    so = S(3, 1, 1)
    if op == "pull":
        got = pp.pull(T(f), T(phi), s).cpu().numpy()
        ref = oracle.pull(f, phi, so, prec="f64")[0, 0]
    elif op == "pull_grad":
        got = pp.pull_grad(T(f), T(phi), s).cpu().numpy()
        ref = oracle.pull_grad(f, phi, so, prec="f64")[0, 0]
    elif op == "pull_vjp":
        df, got = pp.pull_vjp(T(f), T(phi), T(g), s)
        got = got.cpu().numpy()
        ref = oracle.backward_phi(f, phi, g, so)[0]
        assert rel_err(df.cpu().numpy(), oracle.push(g, phi, d, so, prec="f64")[0, 0]) < 1e-5
    elif op == "push":
        got = pp.push(T(g), T(phi), d, s).cpu().numpy()
        ref = oracle.push(g, phi, d, so, prec="f64")[0, 0]
    elif op == "push_grad":
        got = pp.push_grad(T(g3), T(phi), d, s).cpu().numpy()
        ref = oracle.push_grad(g3, phi, d, so, prec="f64")[0, 0]
    elif op == "pull_backward":
        got = pp.pull_backward(T(f), T(phi), T(g), s).cpu().numpy()
        ref = oracle.backward_phi(f, phi, g, so)[0]
    else:
        loss, got, _ = pp.objective_grad(T(f), T(phi), T(g), s)
        got = got.cpu().numpy()
        ref = oracle.backward_phi(f, phi, oracle.pull(f, phi, so, prec="f64") - g, so)[0]
    assert rel_err(got, ref) < 1e-5
    pt.clear()


@pytest.mark.parametrize("shape", [0, 1, 2])
@pytest.mark.parametrize("cap", [6770, 8136, 10184, 13596])
@pytest.mark.parametrize("op", ["push", "pull_vjp", "objective"])
def test_every_box_budget_matches_oracle(op, cap, shape):
    """Each candidate the tuner may choose (tile shape x LDS box budget) against the FP64
    oracle: a smaller budget clips more boxes and sends more taps through global atomics."""
    rng = np.random.default_rng(11)
    d = (40, 48, 56)
    f = smooth_field(rng, d, 2.0, 1.0)[0]
    phi = identity(d) + smooth_field(rng, d, 4.0, 6.0)
    g = rng.standard_normal(d).astype(np.float32)
    dev = "cuda:0"
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    L = pp._lib
    old = [L.get_option(k) for k in (L.PP_OPT_TILING, L.PP_OPT_TILE_SHAPE, L.PP_OPT_BOX_CAP, L.PP_OPT_USE_TUNED)]
    try:
        L.set_option(L.PP_OPT_TILING, 2)
        L.set_option(L.PP_OPT_TILE_SHAPE, shape)
        L.set_option(L.PP_OPT_BOX_CAP, cap)
        L.set_option(L.PP_OPT_USE_TUNED, 0)
        s, so = pp.Settings(3, 1, 1), S(3, 1, 1)
        if op == "push":
            got = pp.push(T(g), T(phi), d, s).cpu().numpy()
            ref = oracle.push(g, phi, d, so, prec="f64")[0, 0]
        elif op == "pull_vjp":
            df, got = pp.pull_vjp(T(f), T(phi), T(g), s)
            got = got.cpu().numpy()
            ref = oracle.backward_phi(f, phi, g, so)[0]
            assert rel_err(df.cpu().numpy(), oracle.push(g, phi, d, so, prec="f64")[0, 0]) < 1e-5
        else:
            _, got, df = pp.objective_grad(T(f), T(phi), T(g), s)
            got = got.cpu().numpy()
            r = oracle.pull(f, phi, so, prec="f64") - g
            ref = oracle.backward_phi(f, phi, r, so)[0]
            assert rel_err(df.cpu().numpy(), oracle.push(r, phi, d, so, prec="f64")[0, 0]) < 1e-5
        assert rel_err(got, ref) < 1e-5
    finally:
        for k, v in zip((L.PP_OPT_TILING, L.PP_OPT_TILE_SHAPE, L.PP_OPT_BOX_CAP, L.PP_OPT_USE_TUNED), old):
            L.set_option(k, v)
