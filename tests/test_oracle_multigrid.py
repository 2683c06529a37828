"""Pins the oracle's push-with-count, invert_def pushes and restrict / prolong
(src/multigrid.jl:21-52, 261-285; pushc at artifacts/C/pushpull.cu:128-142).

The pushes themselves are pinned bit-exactly against the reference C elsewhere; here the
known answers of the callers:

* the count (push of ones) at the identity deformation is one everywhere (partition of
  unity; reflect folding preserves sums), for every degree;
* invert_def's right-hand side at the identity is exactly (g, h) = (0, 1), and for an
  integer translation t it is (-t, 1) away from the edges -- which pins the signs of
  g .-= h .* Id and the identity convention (1-based);
* restrict of a constant image is that constant (the count normalises);
* prolong of a constant displacement keeps the constant inside the volume;
* restrict / prolong through the reference build equal the restatement.
"""
import numpy as np
import pytest

import oracle
from helpers import S


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_count_at_identity_is_one(deg):
    d = (7, 8, 9)
    phi = oracle.identity(d)[None]
    f1 = np.random.default_rng(deg).standard_normal((1, 2) + d).astype(np.float32)
    f0, c0 = oracle.push_count(f1, phi, d, S(deg, 1, 1), prec="f64")
    assert np.allclose(c0, 1.0, rtol=0, atol=1e-12)


def test_invert_def_rhs_identity_and_translation():
    d = (10, 11, 12)
    g, h = oracle.invert_def_rhs(oracle.identity(d)[None])
    assert np.array_equal(h, np.ones_like(h)) and np.array_equal(g, np.zeros_like(g))
    t = np.array([2.0, -1.0, 3.0], np.float32)
    phi = oracle.identity(d)[None] + t[None, :, None, None, None]
    g, h = oracle.invert_def_rhs(phi)
    inner = (0, slice(None), slice(4, -4), slice(4, -4), slice(4, -4))
    assert np.array_equal(h[0, 0][inner[2:]], np.ones((2, 3, 4), np.float32))
    assert np.array_equal(g[inner], np.broadcast_to(-t[:, None, None, None], g[inner].shape).astype(np.float32))


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_restrict_of_constant(deg):
    f1 = np.full((1, 2, 12, 10, 9), 3.5, np.float32)
    f0 = oracle.restrict(f1, (6, 5, 5), deg)
    assert np.allclose(f0, 3.5, rtol=4e-6, atol=0)  # FP32 sums of ~7 weighted ones


def test_prolong_of_constant_displacement():
    c = np.array([0.5, -1.5, 2.0], np.float32)
    f0 = np.broadcast_to(c[None, :, None, None, None], (1, 3, 6, 5, 7)).astype(np.float32)
    f1 = oracle.prolong(f0, (12, 10, 14), 2)
    inner = f1[0, :, 3:-3, 3:-3, 3:-3]
    assert np.allclose(inner, c[:, None, None, None], rtol=1e-6, atol=0)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) absent")
def test_restrict_prolong_restatement_matches_reference_build():
    rng = np.random.default_rng(8)
    f1 = rng.standard_normal((1, 2, 11, 9, 10)).astype(np.float32)
    assert np.array_equal(oracle.restrict(f1, (6, 5, 5), 2, "f32"), oracle.restrict(f1, (6, 5, 5), 2, "ref"))
    f0 = rng.standard_normal((1, 3, 5, 6, 4)).astype(np.float32)
    assert np.array_equal(oracle.prolong(f0, (10, 12, 8), 2, "f32"), oracle.prolong(f0, (10, 12, 8), 2, "ref"))
