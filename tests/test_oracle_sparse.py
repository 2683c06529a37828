"""Pins the sparse-regulariser oracle (oracle/sp_oracle.c) and the package's host-side
operator construction (registration_operator, reduce2fit, sparsify of src/operator.jl).

* the reference's own artifacts/C/sparse_operator.c, compiled from source, gives the same
  bits as the restatement for vel2mom (with and without accumulation at the edges) and
  relax (no H, diagonal H, full H), on every boundary code, including circulant folds
  whose period is not a multiple of the stencil (where the sweep order matters);
* registration_operator reproduces the textbook stencils: the 7-point Laplacian
  (membrane), the 25-point biharmonic (bending: centre 42, faces -12, edges 2, axial
  distance two 1) and the grad-div cross terms, with the voxel-size scalings;
* vel2mom of polynomials gives the derivatives the operators stand for (membrane of x^2
  is -2, bending of x^4 is 24, the divergence cross term of x*y is -1), which pins the
  channel orientation of the lists;
* vel2mom equals a dense FP64 correlation built with numpy.pad (wrap / symmetric /
  symmetric with sign change) for every boundary code;
* reduce2fit on a volume one voxel thick gives the same result as the full stencil folded
  at that axis;
* sparsify's lists rebuild L, and its sum slots hold the non-centre sums;
* many relax sweeps converge to numpy.linalg.solve of the dense (L + H) v = g.
"""
import numpy as np
import pytest

import oracle_sparse as osp
import pushpull as pp
from pushpull.regulariser import DEFAULT_BND

ref_only = pytest.mark.skipif(not osp.ref_available(), reason="reference build (make -C oracle ref) missing")


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def op(lam, vx=(1, 1, 1), vx0=None):
    f = np.float32
    return pp.registration_operator(f(vx), f(lam), None if vx0 is None else f(vx0))


# ------------------------------------------------------------ closed forms ----
def test_membrane_stencil():
    L = op([0.5, 2.0, 0, 0])
    assert L.shape == (3, 3, 3, 3)
    for c in range(3):
        assert L[c, 1, 1, 1] == np.float32(0.5) + np.float32(12.0)
        for idx in [(1, 1, 0), (1, 1, 2), (1, 0, 1), (1, 2, 1), (0, 1, 1), (2, 1, 1)]:
            assert L[c][idx] == -2.0
        assert np.isclose(L[c].sum(), 0.5)  # L applied to a constant is the absolute term


def test_membrane_voxel_sizes():
    L = op([0, 1.0, 0, 0], vx=(1.0, 2.0, 4.0))
    # channel 0 (vx0 = vx): x neighbours vx0_0^2/vx_x^2 = 1, y: 1/4, z: 1/16
    assert L[0, 1, 1, 0] == -1.0 and L[0, 1, 0, 1] == -0.25 and L[0, 0, 1, 1] == -0.0625
    assert L[2, 1, 1, 0] == -16.0  # channel 2: vx0_2^2 / vx_x^2


def test_bending_biharmonic():
    L = op([0, 0, 1.0, 0])
    assert L.shape == (3, 5, 5, 5)
    s = L[0]
    assert s[2, 2, 2] == 42 and s[2, 2, 1] == -12 and s[2, 1, 2] == -12 and s[1, 2, 2] == -12
    assert s[2, 2, 0] == 1 and s[0, 2, 2] == 1 and s[2, 1, 1] == 2 and s[1, 2, 1] == 2 and s[1, 1, 2] == 2
    assert s[1, 1, 1] == 0 and s.sum() == 0
    assert np.count_nonzero(s) == 25


def test_divergence_cross_terms():
    L = op([0, 0, 0, 4.0])
    assert L.shape == (6, 3, 3, 3)
    xy = L[3, 1]  # (y, x) plane at the centre z
    assert xy[0, 0] == -1.0 and xy[2, 2] == -1.0 and xy[0, 2] == 1.0 and xy[2, 0] == 1.0
    assert L[0, 1, 1, 0] == -4.0 and L[0, 1, 0, 1] == 0.0  # channel 0: d2/dx2 only


def test_integer_arguments_like_julia():
    # registration_operator([1,1,1], [0,1,0,1]) of test/test_relax.jl:21
    L = pp.registration_operator([1, 1, 1], [0, 1, 0, 1])
    assert L.shape == (6, 3, 3, 3) and L.dtype == np.float32
    assert L[0, 1, 1, 1] == 8.0 and L[1, 1, 1, 1] == 8.0 and L[3, 1, 0, 0] == -0.25


# ------------------------------------------------------------ sparsify ----
@pytest.mark.parametrize("lam", [[1e-3, 1, 0, 0], [0, 0.1, 0.01, 0.05], [2.0, 0, 0, 0]])
def test_sparsify_rebuilds_L(lam):
    L = op(lam)
    d = (5, 6, 7)
    k = pp.sparsify(L, d)
    nc, (kz, ky, kx) = k.nchan, k.stride
    cross = {(0, 1): 3, (0, 2): 4, (1, 2): 5}
    for j in range(nc):
        for i in range(nc):
            o, n = int(k.offset[j, i]), int(k.length[j, i])
            dense = np.zeros((kz, ky, kx), np.float32)
            for e in range(o, o + n):
                t = int(k.patch_indices[e])
                dense[t // (kx * ky), (t // kx) % ky, t % kx] = k.values[e]
                a, b, c = t % kx - kx // 2, (t // kx) % ky - ky // 2, t // (kx * ky) - kz // 2
                assert k.indices[e] == a + d[2] * (b + d[1] * (c + d[0] * j))
            assert k.patch_indices[o] == (kx * ky * kz) // 2  # the centre comes first
            rest = k.values[o + 1:o + n].astype(np.float64)  # the slot adds them in FP32, in order
            assert abs(k.values[o - 1] - rest.sum()) <= 4 * np.finfo(np.float32).eps * np.abs(rest).sum()
            want = L[i] if i == j else (L[cross[tuple(sorted((i, j)))]] if L.shape[0] == 6 else np.zeros_like(dense))
            assert np.array_equal(dense, want)


# ------------------------------------------------------------ vel2mom semantics ----
def test_vel2mom_polynomials():
    d = (12, 12, 12)
    z, y, x = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in d], indexing="ij")
    inner = (slice(None), slice(2, -2), slice(2, -2), slice(2, -2))
    # membrane: -d2/dx2 of x^2 is -2 (times lambda 3)
    k = pp.sparsify(op([0, 3.0, 0, 0]), d)
    v = np.stack([x * x, np.zeros(d), np.zeros(d)]).astype(np.float32)
    u = osp.vel2mom(v, k, DEFAULT_BND)
    assert np.allclose(u[inner][0], -6.0, atol=1e-3) and np.all(u[inner][1:] == 0)
    # bending: d4/dx4 of x^4 is 24
    k = pp.sparsify(op([0, 0, 1.0, 0]), d)
    v = np.stack([np.zeros(d), (x / 4) ** 4, np.zeros(d)]).astype(np.float32)
    u = osp.vel2mom(v, k, DEFAULT_BND)
    assert np.allclose(u[inner][1], 24.0 / 256, atol=1e-3)
    # divergence: u_0 = -lam d/dx (div v); v = (0, x*y, 0) gives -lam, u_1 = 0
    k = pp.sparsify(op([0, 0, 0, 2.0]), d)
    v = np.stack([np.zeros(d), x * y, np.zeros(d)]).astype(np.float32)
    u = osp.vel2mom(v, k, DEFAULT_BND)
    assert np.allclose(u[inner][0], -2.0, atol=1e-4) and np.allclose(u[inner][1], 0.0, atol=1e-4)


def _dense_vel2mom(v, L, bnd):
    """u_i = sum_j (L_ij correlated with v_j), each v_j padded by numpy.pad."""
    nc = v.shape[0]
    kz, ky, kx = L.shape[1:]
    rz, ry, rx = kz // 2, ky // 2, kx // 2
    cross = {(0, 1): 3, (0, 2): 4, (1, 2): 5}
    u = np.zeros(v.shape, np.float64)
    for j in range(nc):
        p = v[j].astype(np.float64)
        for ax, r in ((2, rx), (1, ry), (0, rz)):
            code = bnd[j][2 - ax]
            w = [(0, 0)] * 3
            w[ax] = (r, r)
            q = np.pad(p, w, mode="wrap" if code == 0 else "symmetric")
            if code == 2:
                sl = [slice(None)] * 3
                sl[ax] = slice(0, r)
                q[tuple(sl)] *= -1
                sl[ax] = slice(q.shape[ax] - r, None)
                q[tuple(sl)] *= -1
            p = q
        for i in range(nc):
            if i == j:
                K = L[i]
            elif L.shape[0] == 6:
                K = L[cross[tuple(sorted((i, j)))]]
            else:
                continue
            for c in range(kz):
                for b in range(ky):
                    for a in range(kx):
                        if K[c, b, a] != 0:
                            u[i] += K[c, b, a] * p[c:c + v.shape[1], b:b + v.shape[2], a:a + v.shape[3]]
    return u


@pytest.mark.parametrize("bnd", [DEFAULT_BND, [(0, 0, 0)] * 3, [(1, 1, 1)] * 3, [(2, 2, 2)] * 3])
def test_vel2mom_equals_dense_correlation(bnd):
    rng = np.random.default_rng(3)
    d = (6, 7, 9)
    L = op([1e-3, 0.1, 0.01, 0.05], vx=(1.0, 1.5, 2.0))
    k = pp.sparsify(L, d)
    v = rng.standard_normal((3,) + d).astype(np.float32)
    u = osp.vel2mom(v, k, bnd)
    want = _dense_vel2mom(v, L, bnd)
    assert np.allclose(u, want, rtol=0, atol=2e-6 * np.abs(want).max())


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_reduce2fit_equals_folded_stencil(axis):
    rng = np.random.default_rng(4)
    d = [5, 6, 7]
    d[axis] = 1
    d = tuple(d)
    L = op([1e-3, 0.3, 0.02, 0.1])
    full = _dense_vel2mom(rng.standard_normal((3,) + d).astype(np.float32), L, DEFAULT_BND)  # warm shape check
    v = rng.standard_normal((3,) + d).astype(np.float32)
    k = pp.sparsify(pp.reduce2fit(L, d, DEFAULT_BND), d)
    assert k.stride[axis] == 1
    u = osp.vel2mom(v, k, DEFAULT_BND)
    # the full stencil, folded at the one-voxel axis by the same boundary codes
    want = _dense_fold_1(v, L, DEFAULT_BND, axis)
    assert full.shape == want.shape
    assert np.allclose(u, want, rtol=0, atol=3e-6 * np.abs(want).max())


def _dense_fold_1(v, L, bnd, axis):
    """Dense correlation where the axis of size 1 is folded by hand: under codes 1 and 2
    every position along it maps to the one voxel, with sign (-1)^distance for code 2."""
    n = L.shape[1 + axis]
    c = n // 2
    Lr = np.zeros_like(L, dtype=np.float64)
    sl = [slice(None)] * 4
    sl[1 + axis] = slice(c, c + 1)
    for ch in range(L.shape[0]):
        code = bnd[ch][2 - axis] if ch < 3 else 1
        w = np.array([(-1.0) ** abs(t - c) if code == 2 else 1.0 for t in range(n)])
        shape = [1, 1, 1]
        shape[axis] = n
        Lr[ch][tuple(sl[1:])] = (L[ch].astype(np.float64) * w.reshape(shape)).sum(axis=axis, keepdims=True)
    keep = [slice(None)] * 4
    keep[1 + axis] = slice(c, c + 1)
    return _dense_vel2mom(v, Lr[tuple(keep)], bnd)


def _dense_matrix(L, d, bnd, nc=3):
    n = int(np.prod(d))
    M = np.zeros((nc * n, nc * n))
    for col in range(nc * n):
        e = np.zeros((nc,) + d, np.float32)
        e.reshape(-1)[col] = 1
        M[:, col] = _dense_vel2mom(e, L, bnd).reshape(-1)
    return M


@pytest.mark.parametrize("hfull", [False, True])
def test_relax_converges_to_dense_solve(hfull):
    rng = np.random.default_rng(5)
    d = (4, 5, 6)
    L = op([1e-2, 0.5, 0, 0.2])
    k = pp.sparsify(L, d)
    n = int(np.prod(d))
    g = rng.standard_normal((3,) + d).astype(np.float32)
    h = np.zeros((6 if hfull else 3,) + d, np.float32)
    h[:3] = 2.0 + rng.random((3,) + d)
    if hfull:
        h[3:] = 0.3 * rng.standard_normal((3,) + d)
    M = _dense_matrix(L, d, DEFAULT_BND)
    H = np.zeros_like(M)
    for c in range(3):
        H[c * n + np.arange(n), c * n + np.arange(n)] = h[c].reshape(-1)
    if hfull:
        for ii, (a, b) in enumerate([(0, 1), (0, 2), (1, 2)]):
            H[a * n + np.arange(n), b * n + np.arange(n)] = h[3 + ii].reshape(-1)
            H[b * n + np.arange(n), a * n + np.arange(n)] = h[3 + ii].reshape(-1)
    want = np.linalg.solve(M + H, g.reshape(-1).astype(np.float64)).reshape(g.shape)
    got = osp.relax(np.zeros_like(g), g, h, k, DEFAULT_BND, 200)
    assert np.allclose(got, want, rtol=0, atol=1e-4 * np.abs(want).max())


# ------------------------------------------------------------ reference build ----
CASES = [
    # (lam, vx, d, bnd)
    ([1e-3, 1, 0, 0], (1, 1, 1), (8, 9, 10), DEFAULT_BND),
    ([0, 0.1, 0.01, 0.05], (1, 1.5, 2), (7, 11, 12), DEFAULT_BND),
    ([0, 0.1, 0.01, 0.05], (1, 1, 1), (7, 11, 12), [(0, 0, 0)] * 3),   # circulant, 12 % 5 != 0
    ([1e-3, 0.2, 0, 0.3], (1, 1, 1), (9, 8, 10), [(0, 1, 2), (2, 0, 1), (1, 2, 0)]),
    ([2.0, 0, 0, 0], (1, 1, 1), (3, 4, 5), DEFAULT_BND),
    ([0, 0.1, 0.01, 0], (1, 1, 1), (2, 3, 6), DEFAULT_BND),             # volume thinner than the stencil
]


@ref_only
@pytest.mark.parametrize("case", range(len(CASES)))
def test_vel2mom_matches_reference_build(case):
    lam, vx, d, bnd = CASES[case]
    rng = np.random.default_rng(10 + case)
    k = pp.sparsify(op(lam, vx), d)
    v = rng.standard_normal((3,) + d).astype(np.float32)
    assert bits_equal(osp.vel2mom(v, k, bnd), osp.ref_vel2mom(v, k, bnd))
    u0 = rng.standard_normal((3,) + d).astype(np.float32)  # the reference adds at the edges
    assert bits_equal(osp.vel2mom(v, k, bnd, u=u0, accum_edges=True), osp.ref_vel2mom(v, k, bnd, u=u0))


@ref_only
@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("hcode", [0, 1, 2])
def test_relax_matches_reference_build(case, hcode):
    lam, vx, d, bnd = CASES[case]
    rng = np.random.default_rng(20 + case)
    k = pp.sparsify(op(lam, vx), d)
    g = rng.standard_normal((3,) + d).astype(np.float32)
    v0 = 0.1 * rng.standard_normal((3,) + d).astype(np.float32)
    h = None
    if hcode:
        h = np.zeros((3 if hcode == 1 else 6,) + d, np.float32)
        h[:3] = 1.0 + rng.random((3,) + d)
        if hcode == 2:
            h[3:] = 0.2 * rng.standard_normal((3,) + d)
    assert bits_equal(osp.relax(v0, g, h, k, bnd, 2), osp.ref_relax(v0, g, h, k, bnd, 2))


@ref_only
def test_reduced_2d_kernel_matches_reference_build():
    rng = np.random.default_rng(7)
    d = (1, 9, 10)
    L = pp.reduce2fit(op([1e-3, 0.1, 0.01, 0.05]), d, DEFAULT_BND)
    k = pp.sparsify(L, d)
    assert k.stride == (1, 5, 5)
    v = rng.standard_normal((3,) + d).astype(np.float32)
    assert bits_equal(osp.vel2mom(v, k, DEFAULT_BND), osp.ref_vel2mom(v, k, DEFAULT_BND))
    h = 1.0 + rng.random((3,) + d).astype(np.float32)
    assert bits_equal(osp.relax(v, v, h, k, DEFAULT_BND, 3), osp.ref_relax(v, v, h, k, DEFAULT_BND, 3))


@ref_only
def test_single_channel_kernel_matches_reference_build():
    rng = np.random.default_rng(8)
    d = (6, 7, 8)
    k = pp.sparsify(op([1e-3, 0.5, 0.05, 0]), d, nd=1)
    assert k.nchan == 1 and k.offset.shape == (1, 1)
    v = rng.standard_normal((1,) + d).astype(np.float32)
    bnd = [(1, 1, 1)]
    assert bits_equal(osp.vel2mom(v, k, bnd), osp.ref_vel2mom(v, k, bnd))
    h = 1.0 + rng.random((1,) + d).astype(np.float32)
    assert bits_equal(osp.relax(v, v, h, k, bnd, 2), osp.ref_relax(v, v, h, k, bnd, 2))


@ref_only
@pytest.mark.parametrize("nd", [2, 4])
def test_other_channel_counts_match_reference_build(nd):
    # sparsify takes the first nd channels of L as the diagonal lists (src/operator.jl:237-253)
    rng = np.random.default_rng(30 + nd)
    d = (7, 8, 9)
    k = pp.sparsify(op([1e-3, 0.2, 0.01, 0.05]), d, nd=nd)
    assert k.nchan == nd
    bnd = [(1, 2, 0), (2, 1, 1), (0, 0, 1), (1, 1, 2)][:nd]
    v = rng.standard_normal((nd,) + d).astype(np.float32)
    assert bits_equal(osp.vel2mom(v, k, bnd), osp.ref_vel2mom(v, k, bnd))
    h = 1.0 + rng.random((nd,) + d).astype(np.float32)
    assert bits_equal(osp.relax(v, v, h, k, bnd, 2), osp.ref_relax(v, v, h, k, bnd, 2))
