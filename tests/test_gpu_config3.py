"""Differentiation workload at full size (needs a GPU): L = 0.5*||pull(f, phi) - g||^2 at
256^3 Float32, degree 2, reflect boundary, ext=1; f and g independent smoothed-noise
volumes, phi = identity + smooth displacement of max 6 voxels. dL/df (push of the residual)
and dL/dphi (the residual contracted with the gradient of pull) from torch.autograd through
the rrule, and from the fused objective, against central finite differences with h = 1e-3
on 64 seeded voxels each, evaluated locally in FP64.

Measure (test/utils.jl:3-25): sum((g_ad - g_fd)^2) / sum(g_ad^2), bar 1e-3 as in
test/runtests.jl:44,63-67.
"""
import numpy as np
import pytest

import oracle
from helpers import S, identity, report

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
N = 256
H_FD = 1e-3
NVOX = 64
MARGIN = 20  # FD voxels stay this far from the faces: the local blocks (+-18) fit, no folds


def _smooth32(rng, shape, sigma):
    from scipy.ndimage import gaussian_filter1d

    a = rng.standard_normal(shape, dtype=np.float32)
    for ax in (-3, -2, -1):
        a = gaussian_filter1d(a, sigma, axis=ax, mode="reflect", truncate=4.0, output=np.float32)
    return a


@pytest.fixture(scope="module")
def case():
    rng = np.random.default_rng(3)
    d = (N, N, N)
    f = _smooth32(rng, d, 2.0)
    f /= f.std()
    g = _smooth32(rng, d, 2.0)
    g /= g.std()
    u = np.stack([_smooth32(rng, d, 4.0) for _ in range(3)])
    u *= 6.0 / np.abs(u).max()
    phi = identity(d) + u
    return f, g, phi.astype(np.float32)


def _b2(u):
    """Quadratic B-spline (textbook form)."""
    a = np.abs(u)
    return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))


def _pull64(f, x, y, z):
    """Degree-2 pull of f at one 1-based coordinate (interior: no folds), in FP64."""
    xs, ys, zs = x - 1.0, y - 1.0, z - 1.0
    ox, oy, oz = (int(np.ceil(c - 1.5)) for c in (xs, ys, zs))
    ix, iy, iz = np.arange(ox, ox + 3), np.arange(oy, oy + 3), np.arange(oz, oz + 3)
    wx, wy, wz = _b2(xs - ix), _b2(ys - iy), _b2(zs - iz)
    blk = f[iz[0]:iz[-1] + 1, iy[0]:iy[-1] + 1, ix[0]:ix[-1] + 1].astype(np.float64)
    return float(np.einsum("k,j,i,kji->", wz, wy, wx, blk))


def _voxels(rng):
    return rng.integers(MARGIN, N - MARGIN, size=(NVOX, 3))  # (z, y, x)


def _fd_phi(f, g, phi, vox, axes):
    """Central differences of L along phi[a, v]: only sample v changes."""
    out = np.empty(len(vox))
    for n, ((z, y, x), a) in enumerate(zip(vox, axes)):
        c = phi[:, z, y, x].astype(np.float64)

        def ell(cc):
            return 0.5 * (_pull64(f, cc[0], cc[1], cc[2]) - float(g[z, y, x])) ** 2

        e = np.zeros(3)
        e[a] = H_FD
        out[n] = (ell(c + e) - ell(c - e)) / (2 * H_FD)
    return out


def _fd_f(f, g, phi, vox):
    """Central differences of L along f[u], on the samples within 10 voxels of u (the only
    ones that reach u: displacement <= 6, support 1.5) and a block of f around them (taps
    within 18 voxels), with the FP64 oracle; L is quadratic in f, so the differences are
    exact up to rounding."""
    so = S(2, 1, 1)
    out = np.empty(len(vox))
    for n, (z, y, x) in enumerate(vox):
        w0, w1 = np.array([z, y, x]) - 10, np.array([z, y, x]) + 11
        b0, b1 = np.array([z, y, x]) - 18, np.array([z, y, x]) + 19
        ph = phi[:, w0[0]:w1[0], w0[1]:w1[1], w0[2]:w1[2]].copy()
        ph[0] -= b0[2]  # block-local 1-based coordinates (exact: integer shifts)
        ph[1] -= b0[1]
        ph[2] -= b0[0]
        gw = g[w0[0]:w1[0], w0[1]:w1[1], w0[2]:w1[2]].astype(np.float64)
        blk = f[b0[0]:b1[0], b0[1]:b1[1], b0[2]:b1[2]].astype(np.float64)
        assert ph.min() > 2.5 and ph[0].max() < blk.shape[2] - 1.5 and ph[2].max() < blk.shape[0] - 1.5
        c = (18, 18, 18)

        def loss(bb):
            r = oracle.pull(bb, ph, so, prec="f64")[0, 0] - gw
            return 0.5 * float(np.sum(r * r))

        blk[c] += H_FD
        lp = loss(blk)
        blk[c] -= 2 * H_FD
        lm = loss(blk)
        out[n] = (lp - lm) / (2 * H_FD)
    return out


def _measure(ad, fd):
    return float(np.sum((ad - fd) ** 2) / np.sum(ad ** 2))


def test_config3_gradients_match_central_differences(case):
    f, g, phi = case
    s = pp.Settings(2, 1, 1)
    tf = torch.from_numpy(f).to(DEV).requires_grad_(True)
    tp = torch.from_numpy(phi).to(DEV).requires_grad_(True)
    tg = torch.from_numpy(g).to(DEV)
    r = pp.pull(tf, tp, s) - tg
    (0.5 * (r.double() ** 2).sum()).backward()
    gf, gp = tf.grad.cpu().numpy(), tp.grad.cpu().numpy()
    loss_o, gp_o, gf_o = pp.objective_grad(tf.detach(), tp.detach(), tg, s)
    gf_o, gp_o = gf_o.cpu().numpy(), gp_o.cpu().numpy()

    rng = np.random.default_rng(33)
    vp, axes = _voxels(rng), rng.integers(0, 3, NVOX)
    vf = _voxels(rng)
    fd_p = _fd_phi(f, g, phi, vp, axes)
    fd_f = _fd_f(f, g, phi, vf)
    ad_p = np.array([gp[a, z, y, x] for (z, y, x), a in zip(vp, axes)], np.float64)
    ad_f = np.array([gf[z, y, x] for z, y, x in vf], np.float64)
    ob_p = np.array([gp_o[a, z, y, x] for (z, y, x), a in zip(vp, axes)], np.float64)
    ob_f = np.array([gf_o[z, y, x] for z, y, x in vf], np.float64)
    errs = {"config3 dL/dphi autograd vs FD": _measure(ad_p, fd_p), "config3 dL/df autograd vs FD": _measure(ad_f, fd_f),
            "config3 dL/dphi objective vs FD": _measure(ob_p, fd_p), "config3 dL/df objective vs FD": _measure(ob_f, fd_f)}
    for k, v in errs.items():
        report(k, v)
    bad = {k: v for k, v in errs.items() if not v < 1e-3}
    assert not bad, bad
