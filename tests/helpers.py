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
