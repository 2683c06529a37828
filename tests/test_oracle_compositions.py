"""Pins the oracle's compose / Exp (src/compositions.jl:9-49) without a Julia run.

compose is three pulls with per-channel boundary triples; the pulls themselves are pinned
bit-exactly against the reference C (test_oracle_golden.py). What is pinned here is the
composition layer:

* compose at the identity returns phi exactly (degree 1 at integer coordinates);
* the sliding matrix's orientation: a deformation pointing outside the field of view along
  axis a flips the sign of channel a only (antisymmetric along its own axis, reflect along
  the others) -- swapping axes or channels fails this;
* Exp(0) is the identity exactly;
* Exp of a constant velocity c is Id + c / 2^nsq away from the boundary: the reference's
  update u = compose(u, Id + u) keeps u o (Id + u) only (not u + u o (Id + u)), so the
  displacement stays c / 2^nsq -- the test fails for the group-law variant (Id + c);
* where the reference C build is present, compose through it equals the restatement.
"""
import numpy as np
import pytest

import oracle


def test_compose_identity_is_exact():
    rng = np.random.default_rng(3)
    for d in [(6, 7, 8), (1, 5, 9), (4, 1, 3)]:
        phi = rng.standard_normal((2, 3) + d).astype(np.float32)
        Id = np.stack([oracle.identity(d)] * 2)
        assert np.array_equal(oracle.compose(phi, Id), phi)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sliding_boundary_orientation(axis):
    d = (9, 10, 11)  # (Z, Y, X)
    ones = np.ones((1, 3) + d, np.float32)
    theta = oracle.identity(d)[None].copy()
    theta[0, axis] -= 3.0  # sample 3 voxels beyond the low edge along `axis`
    psi = oracle.compose(ones, theta)[0]
    ax = 2 - axis  # array axis of X/Y/Z within (Z, Y, X)
    idx = np.arange(d[ax]).reshape([-1 if k == ax else 1 for k in range(3)])
    outside = np.broadcast_to(idx < 3, d)
    for c in range(3):
        want = np.where(outside, -1.0, 1.0) if c == axis else np.ones(d)
        assert np.array_equal(psi[c], want.astype(np.float32)), (axis, c)


def test_exp_of_zero_is_identity():
    d = (5, 6, 7)
    assert np.array_equal(oracle.exp(np.zeros((1, 3) + d, np.float32))[0], oracle.identity(d))


@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_exp_of_constant_velocity(prec):
    d = (16, 17, 18)
    c = np.array([0.5, -0.25, 1.0], np.float32)
    v = np.broadcast_to(c[:, None, None, None], (3,) + d).astype(np.float32)[None]
    nsq = 8
    e = oracle.exp(v, nsq, prec)[0]
    Id = oracle.identity(d)
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    want = Id + (c / 2 ** nsq)[:, None, None, None]
    assert np.allclose(e[inner], want[inner], rtol=0, atol=1e-6)
    assert not np.allclose(e[inner], (Id + c[:, None, None, None])[inner], atol=1e-3)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) absent")
def test_compose_restatement_matches_reference_build():
    rng = np.random.default_rng(5)
    d0, d1 = (7, 8, 9), (6, 9, 5)
    phi = rng.standard_normal((2, 3) + d0).astype(np.float32)
    theta = np.stack([oracle.identity(d1) + rng.uniform(-4, 4, (3,) + d1).astype(np.float32) for _ in range(2)])
    a = oracle.compose(phi, theta, "f32")
    b = oracle.compose(phi, theta, "ref")
    assert np.array_equal(a, b, equal_nan=True)
