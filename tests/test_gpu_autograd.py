"""torch.autograd through pull, pull_grad and push (the rrules of src/pushpull_pb.jl:11-51)
on the GPU (needs a GPU).

With both gradients wanted, the backward runs the φ pullback on a second stream beside the
other gradient; the results must equal the single-gradient backward passes (which run on
one stream) exactly, and the FP64 oracle to 1e-5 norm-wise: d/df0 of <pull(f0, φ), δ> is
push(δ, φ) and d/dφ is Σ_c δ_c ∇pull(f0)_c; for pull_grad, push_grad and the Hessian
contraction; for push, pull(δ) and Σ_c f1_c ∇pull(δ)_c.
"""
import numpy as np
import pytest

import oracle
from helpers import S, identity, rel_err, same_nan_mask, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
TOL = 1e-5


def T(a, grad=False):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t.requires_grad_(grad)


def H(t):
    return t.detach().cpu().numpy()


def _inputs(rng, d0=(12, 13, 14), d1=(11, 12, 13)):
    f0 = smooth_field(rng, d0, 2.0, 3.0)[0][None, None].astype(np.float32)  # (1, 1, Z, Y, X)
    sc = np.array(d0)[::-1, None, None, None] / np.array(d1)[::-1, None, None, None]
    phi = (identity(d1) * sc + smooth_field(rng, d1, 2.0, 2.0)).astype(np.float32)
    return f0, phi


def _close(got, ref, what):
    assert same_nan_mask(got, ref), what
    assert rel_err(got, ref) <= TOL, (what, rel_err(got, ref))


@pytest.mark.parametrize("deg", [1, 3])
def test_pull_gradients(deg):
    rng = np.random.default_rng(60 + deg)
    f0, phi = _inputs(rng)
    s, so = pp.Settings(deg, 1, 1), S((deg,) * 3, (1, 1, 1), 1)
    delta = rng.standard_normal((1, 1) + phi.shape[1:]).astype(np.float32)
    tf, tp = T(f0, True), T(phi, True)
    (pp.pull(tf, tp, s) * T(delta)).sum().backward()
    _close(H(tf.grad), oracle.push(delta, phi, f0.shape[-3:], so, "f64"), "df0")
    _close(H(tp.grad), oracle.backward_phi(f0, phi[None], delta, so)[0], "dphi")
    # one gradient at a time (no side stream) gives the same bits
    tf1, tp1 = T(f0, True), T(phi)
    (pp.pull(tf1, tp1, s) * T(delta)).sum().backward()
    tf2, tp2 = T(f0), T(phi, True)
    (pp.pull(tf2, tp2, s) * T(delta)).sum().backward()
    assert np.array_equal(H(tp2.grad), H(tp.grad))
    # push accumulates with float atomics: its order is free, so compare to tolerance
    _close(H(tf1.grad), H(tf.grad), "df0 single vs both")


def test_pull_grad_gradients():
    rng = np.random.default_rng(70)
    f0, phi = _inputs(rng)
    s, so = pp.Settings(2, 1, 1), S((2, 2, 2), (1, 1, 1), 1)
    delta = rng.standard_normal((1, 1, 3) + phi.shape[1:]).astype(np.float32)
    tf, tp = T(f0, True), T(phi, True)
    (pp.pull_grad(tf, tp, s) * T(delta)).sum().backward()
    _close(H(tf.grad), oracle.push_grad(delta, phi, f0.shape[-3:], so, "f64"), "df0")
    hess = oracle.pull_hess(f0, phi, so, "f64")  # (B, C, 3, 3, ...)
    want = (delta[:, :, :, None] * hess).sum(axis=(1, 2))[0]
    _close(H(tp.grad), want, "dphi")


def test_push_gradients():
    rng = np.random.default_rng(80)
    f0, phi = _inputs(rng)
    s, so = pp.Settings(3, 1, 1), S((3, 3, 3), (1, 1, 1), 1)
    f1 = rng.standard_normal((1, 1) + phi.shape[1:]).astype(np.float32)
    delta = rng.standard_normal(f0.shape).astype(np.float32)
    tf1, tp = T(f1, True), T(phi, True)
    (pp.push(tf1, tp, f0.shape[-3:], s) * T(delta)).sum().backward()
    _close(H(tf1.grad), oracle.pull(delta, phi, so, "f64"), "df1")
    _close(H(tp.grad), oracle.backward_phi(delta, phi[None], f1, so)[0], "dphi")


def test_gradients_match_finite_differences():
    # the rrules against calculus rather than the oracle's formulas: central differences of
    # L = <pull(f0, φ), δ> in FP64 on the host oracle, at a few entries of f0 and φ
    rng = np.random.default_rng(90)
    f0, phi = _inputs(rng)
    s, so = pp.Settings(3, 1, 1), S((3, 3, 3), (1, 1, 1), 1)
    delta = rng.standard_normal((1, 1) + phi.shape[1:]).astype(np.float32)
    tf, tp = T(f0, True), T(phi, True)
    (pp.pull(tf, tp, s) * T(delta)).sum().backward()
    gf, gp = H(tf.grad), H(tp.grad)

    def loss(f, p):
        return float((oracle.pull(f, p, so, "f64") * delta).sum())

    for idx in [(0, 0, 6, 6, 7), (0, 0, 3, 9, 2), (0, 0, 10, 1, 12)]:
        e = np.zeros_like(f0, dtype=np.float64)
        e[idx] = 1.0
        fd = (loss(f0 + e, phi) - loss(f0 - e, phi)) / 2  # L is linear in f0
        assert abs(fd - gf[idx]) <= 1e-4 * max(1.0, abs(fd)), (idx, fd, gf[idx])
    h = 1e-2
    for idx in [(0, 5, 6, 6), (1, 4, 7, 3), (2, 8, 2, 9)]:
        e = np.zeros_like(phi)
        e[idx] = h
        fd = (loss(f0, phi + e) - loss(f0, phi - e)) / (2 * h)
        assert abs(fd - gp[idx]) <= 2e-3 * max(1.0, abs(fd)), (idx, fd, gp[idx])
