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
    assert len(lines) == 1 and int(lines[0].split()[-1]) == v
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
