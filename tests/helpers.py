"""Shared helpers for the tests: settings views, synthetic volumes and error metrics."""
from __future__ import annotations

import numpy as np


class S:
    """Minimal settings object (deg, bnd list, ext) for the oracle."""

    def __init__(self, deg, bnd, ext):
        self.deg = tuple(int(d) for d in np.atleast_1d(deg)) if np.ndim(deg) else (int(deg),) * 3
        b = np.asarray(bnd)
        if b.ndim == 0:
            self.bnd = [(int(b),) * 3]
        elif b.ndim == 1:
            self.bnd = [tuple(int(v) for v in b)]
        else:
            self.bnd = [tuple(int(v) for v in r) for r in b]
        self.ext = int(ext)


def case_settings(c):
    return S(c["deg"], c["bnd"], int(c["ext"]))


def rel_err(a, b):
    """||a-b|| / ||b|| over entries finite in b (NaN masks are compared separately)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = np.isfinite(b)
    den = np.linalg.norm(b[m])
    num = np.linalg.norm(a[m] - b[m])
    return num / den if den > 0 else num


def same_nan_mask(a, b):
    return np.array_equal(np.isnan(a), np.isnan(b))


def identity(d1):
    """1-based identity grid (3, Z, Y, X)."""
    Z, Y, X = d1
    phi = np.empty((3, Z, Y, X), np.float32)
    phi[0] = np.arange(1, X + 1)[None, None, :]
    phi[1] = np.arange(1, Y + 1)[None, :, None]
    phi[2] = np.arange(1, Z + 1)[:, None, None]
    return phi


def gauss_smooth(a, sigma, axes=(-3, -2, -1)):
    """Separable Gaussian smoothing, truncated at 4 sigma, reflect boundary (scipy)."""
    from scipy.ndimage import gaussian_filter1d

    out = np.asarray(a, np.float64)
    for ax in axes:
        out = gaussian_filter1d(out, sigma, axis=ax, mode="reflect", truncate=4.0)
    return out


def smooth_field(rng, d, sigma, maxabs):
    """Per-axis Gaussian noise smoothed by sigma and scaled to max |u| = maxabs."""
    u = gauss_smooth(rng.standard_normal((3,) + tuple(d)), sigma)
    u *= maxabs / np.abs(u).max()
    return u.astype(np.float32)


def max_abs(a):
    """Largest finite |a| (0 when none)."""
    a = np.asarray(a)
    f = a[np.isfinite(a)]
    return float(np.abs(f).max()) if f.size else 0.0


def check_close(got, ref, what, tol=1e-5, scale=None):
    """NaN masks equal; ||got - ref|| <= tol*||ref|| over finite entries, plus (when `scale`,
    the largest input magnitude, is given) an absolute floor 1e-6*sqrt(N)*scale at FP32
    rounding level for outputs that are mathematically (near) zero."""
    got = np.asarray(got)
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (what, got.shape, ref.shape)
    assert same_nan_mask(got, ref), f"{what}: NaN mask differs"
    m = np.isfinite(ref)
    num = float(np.linalg.norm(got[m].astype(np.float64) - ref[m]))
    den = float(np.linalg.norm(ref[m].astype(np.float64)))
    floor = 1e-6 * np.sqrt(max(1, int(m.sum()))) * scale if scale is not None else 0.0
    assert num <= tol * den + floor, f"{what}: rel err {num / den if den else num:.3g}"
    return num / den if den else num


def report(name, value):
    """Append `name value` to the file named by PP_PARITY_LOG (when set): the measured
    errors of the parity tests, for the summary of errors per workload."""
    import os

    path = os.environ.get("PP_PARITY_LOG")
    if path:
        with open(path, "a") as fh:
            fh.write(f"{name} {float(value):.6g}\n")
