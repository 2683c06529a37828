"""Pins the CPU oracle against the reference C (compiled from source).

* the FP32 restatement reproduces every committed golden output bit for bit
  (pull, pullg, pullh, push), including NaN masks, knots at thresholds, folds over
  several periods, coordinates beyond 2^20 and beyond the int64 range;
* the FP64 oracle agrees with the FP32 reference to FP32 rounding;
* the reference's CPU pushg returns zeros for finite input (its inverted test at
  artifacts/C/pushpull.c:144), which is why push_grad is pinned by adjointness instead;
* where oracle/_ref is present, a fresh random sweep compares restatement and
  reference live.
"""
import numpy as np
import pytest

import oracle
from helpers import S, case_settings, rel_err, same_nan_mask


def _eq(a, b):
    return np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)


def test_golden_count(golden):
    assert len(golden) >= 100


@pytest.mark.parametrize("op", ["pull", "pullg", "pullh", "push"])
def test_f32_restatement_bit_exact(golden, op):
    bad = []
    for i, c in enumerate(golden):
        s = case_settings(c)
        d0 = c["f0"].shape
        if op == "pull":
            got = oracle.pull(c["f0"], c["phi"], s)[0, 0]
        elif op == "pullg":
            got = oracle.pull_grad(c["f0"], c["phi"], s)[0, 0]
        elif op == "pullh":
            got = oracle.pull_hess(c["f0"], c["phi"], s)[0, 0].reshape(c["pullh"].shape)
        else:
            got = oracle.push(c["f1"], c["phi"], d0, s)[0, 0]
        if not _eq(got, c[op]):
            bad.append(i)
    assert not bad, f"{op}: cases differ from the reference: {bad[:10]}"


@pytest.mark.parametrize("op", ["pull", "pullg", "pullh", "push"])
def test_f64_oracle_close_to_reference(golden, op):
    worst = 0.0
    for c in golden:
        s = case_settings(c)
        d0 = c["f0"].shape
        if op == "pull":
            got = oracle.pull(c["f0"], c["phi"], s, prec="f64")[0, 0]
        elif op == "pullg":
            got = oracle.pull_grad(c["f0"], c["phi"], s, prec="f64")[0, 0]
        elif op == "pullh":
            got = oracle.pull_hess(c["f0"], c["phi"], s, prec="f64")[0, 0].reshape(c["pullh"].shape)
        else:
            got = oracle.push(c["f1"], c["phi"], d0, s, prec="f64")[0, 0]
        assert same_nan_mask(got, c[op])
        # extreme coordinates (|x| >= 2^20) make FP32 weights inexact by design; skip those
        phi = c["phi"]
        if np.any(np.abs(phi[np.isfinite(phi)]) > 2 ** 20):
            continue
        worst = max(worst, rel_err(c[op], got))
    assert worst < 2e-6, worst


def test_reference_pushg_is_a_noop_for_finite_input(golden):
    for c in golden:
        if np.all(np.isfinite(c["g1"])):
            assert np.all(c["pushg_ref"] == 0)


def test_pushg_restatement_is_nonzero(golden):
    c = golden[-1]
    got = oracle.push_grad(c["g1"], c["phi"], c["f0"].shape, case_settings(c))
    assert np.abs(got).max() > 0


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build (oracle/_ref) absent")
def test_live_random_sweep_against_reference():
    rng = np.random.default_rng(7)
    for trial in range(60):
        d0 = tuple(int(v) for v in rng.integers(1, 9, size=3))
        d1 = tuple(int(v) for v in rng.integers(1, 7, size=3))
        deg = tuple(int(v) for v in rng.integers(0, 5, size=3))
        bnd = tuple(int(v) for v in rng.integers(0, 3, size=3))
        ext = int(rng.integers(0, 2))
        s = S(deg, bnd, ext)
        f0 = rng.standard_normal(d0).astype(np.float32)
        phi = (rng.uniform(-2, 2 * max(d0), size=(3,) + d1)).astype(np.float32)
        f1 = rng.standard_normal(d1).astype(np.float32)
        for fn in (oracle.pull, oracle.pull_grad, oracle.pull_hess):
            assert _eq(fn(f0, phi, s), fn(f0, phi, s, prec="ref")), (trial, fn.__name__)
        assert _eq(oracle.push(f1, phi, d0, s), oracle.push(f1, phi, d0, s, prec="ref")), trial


@pytest.mark.parametrize("op", ["pull_affine", "push_affine"])
def test_affine_restatement_bit_exact(golden_affine, op):
    assert len(golden_affine) >= 20
    bad = []
    for i, c in enumerate(golden_affine):
        s = case_settings(c)
        if op == "pull_affine":
            got = oracle.affine_pull(c["f0"], c["aff0"], tuple(c["d1"]), s)[0, 0]
        else:
            got = oracle.affine_push(c["f1"], c["aff0"], c["f0"].shape, s)[0, 0]
        if not _eq(got, c[op]):
            bad.append(i)
    assert not bad, bad
