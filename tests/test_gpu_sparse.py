"""The sparse regulariser on the GPU (pp_vel2mom, pp_relax and the host-pointer vel2mom /
relax) against the oracle (needs a GPU).

Bar: bit-identical to the FP32 oracle, which is itself bit-identical to the reference's
artifacts/C/sparse_operator.c (tests/test_oracle_sparse.py). Cases: membrane, bending +
divergence with anisotropic voxels, the absolute term alone, per-channel boundary codes,
circulant folds whose period is not a multiple of the stencil (the sweep-order case),
volumes thinner than the stencil, a kernel reduced to 2-D, one channel; H absent, diagonal
and full; accumulation into u at the edges; device tensors and host arrays. At 128^3 the
checks are the size-independent ones: linearity of vel2mom, L applied to a constant, and
relax lowering the residual.
"""
import numpy as np
import pytest

import oracle_sparse as osp

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")
from pushpull.regulariser import DEFAULT_BND  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


def bits_equal(a, b):
    return np.array_equal(np.asarray(a, np.float32).view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def op(lam, vx=(1, 1, 1)):
    return pp.registration_operator(np.float32(vx), np.float32(lam))


CASES = [
    # (lam, vx, d (Z, Y, X), bnd)
    ([1e-3, 1, 0, 0], (1, 1, 1), (9, 10, 67), DEFAULT_BND),
    ([0, 0.1, 0.01, 0.05], (1, 1.5, 2), (13, 11, 70), DEFAULT_BND),
    ([0, 0.1, 0.01, 0.05], (1, 1, 1), (11, 12, 66), [(0, 0, 0)] * 3),        # circulant, 66 % 5 != 0
    ([1e-3, 0.2, 0, 0.3], (1, 1, 1), (10, 8, 65), [(0, 1, 2), (2, 0, 1), (1, 2, 0)]),
    ([2.0, 0, 0, 0], (1, 1, 1), (3, 4, 5), DEFAULT_BND),
    ([0, 0.1, 0.01, 0], (1, 1, 1), (2, 3, 6), DEFAULT_BND),                   # thinner than the stencil
    ([0, 0.1, 0.01, 0.05], (1, 1, 1), (1, 1, 1), DEFAULT_BND),                # one voxel
]


def _kernel(case):
    lam, vx, d, bnd = CASES[case]
    return pp.sparsify(pp.reduce2fit(op(lam, vx), d, DEFAULT_BND), d), d, bnd


def _h(rng, hcode, d, nc=3):
    if hcode == 0:
        return None
    h = np.zeros((nc if hcode == 1 else nc * (nc + 1) // 2,) + d, np.float32)
    h[:nc] = 1.0 + rng.random((nc,) + d)
    if hcode == 2:
        h[nc:] = 0.2 * rng.standard_normal(h[nc:].shape)
    return h


@pytest.mark.parametrize("case", range(len(CASES)))
def test_vel2mom_bit_exact(case):
    k, d, bnd = _kernel(case)
    rng = np.random.default_rng(100 + case)
    v = rng.standard_normal((3,) + d).astype(np.float32)
    print(f"vel2mom case {case} d={d} stride={k.stride}", flush=True)
    u = pp.vel2mom(T(v), k, bnd)
    torch.cuda.synchronize()
    assert bits_equal(H(u), osp.vel2mom(v, k, bnd))
    # vel2mom!: the edges add to the u they are given
    u0 = rng.standard_normal((3,) + d).astype(np.float32)
    ut = T(u0)
    pp.vel2mom_(ut, T(v), k, bnd)
    assert bits_equal(H(ut), osp.vel2mom(v, k, bnd, u=u0, accum_edges=True))


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("hcode", [0, 1, 2])
def test_relax_bit_exact(case, hcode):
    k, d, bnd = _kernel(case)
    rng = np.random.default_rng(200 + case)
    g = rng.standard_normal((3,) + d).astype(np.float32)
    v0 = 0.1 * rng.standard_normal((3,) + d).astype(np.float32)
    h = _h(rng, hcode, d)
    print(f"relax case {case} hcode {hcode} d={d}", flush=True)
    v = T(v0)
    pp.relax_(T(g), None if h is None else T(h), k, bnd, 2, v)
    torch.cuda.synchronize()
    assert bits_equal(H(v), osp.relax(v0, g, h, k, bnd, 2))


@pytest.mark.parametrize("case", [1, 2, 3])
def test_host_arrays_bit_exact(case):
    k, d, bnd = _kernel(case)
    rng = np.random.default_rng(300 + case)
    v = rng.standard_normal((3,) + d).astype(np.float32)
    u0 = rng.standard_normal((3,) + d).astype(np.float32)
    u = u0.copy()
    pp.vel2mom_(u, v, k, bnd)
    assert bits_equal(u, osp.vel2mom(v, k, bnd, u=u0, accum_edges=True))
    g = rng.standard_normal((3,) + d).astype(np.float32)
    h = _h(rng, 2, d)
    vv = v.copy()
    pp.relax_(g, h, k, bnd, 2, vv)
    assert bits_equal(vv, osp.relax(v, g, h, k, bnd, 2))


def test_single_channel_bit_exact():
    d = (9, 10, 33)
    k = pp.sparsify(op([1e-3, 0.5, 0.05, 0]), d, nd=1)
    rng = np.random.default_rng(7)
    v = rng.standard_normal((1,) + d).astype(np.float32)
    bnd = [(1, 2, 0)]
    assert bits_equal(H(pp.vel2mom(T(v), k, bnd)), osp.vel2mom(v, k, bnd))
    h = 1.0 + rng.random((1,) + d).astype(np.float32)
    vt = T(v)
    pp.relax_(T(v), T(h), k, bnd, 3, vt)
    assert bits_equal(H(vt), osp.relax(v, v, h, k, bnd, 3))


def test_full_size_properties():
    d = (128, 128, 128)
    k = pp.sparsify(op([1e-3, 0.1, 0.01, 0.05]), d)
    rng = np.random.default_rng(9)
    a = torch.randn((3,) + d, device=DEV)
    b = torch.randn((3,) + d, device=DEV)
    # linearity: L(2a + b) = 2 L a + L b (to rounding)
    lhs = pp.vel2mom(2 * a + b, k)
    rhs = 2 * pp.vel2mom(a, k) + pp.vel2mom(b, k)
    assert float((lhs - rhs).abs().max()) <= 1e-5 * float(rhs.abs().max())
    # a constant field along every channel: only the absolute term remains away from the
    # antisymmetric faces
    c = torch.full((3,) + d, 2.0, device=DEV)
    u = pp.vel2mom(c, k)
    inner = u[:, 3:-3, 3:-3, 3:-3]
    assert torch.allclose(inner, torch.full_like(inner, 2e-3), atol=1e-6)
    # relax lowers the residual of (L + H) v = g
    g = torch.randn((3,) + d, device=DEV)
    h = torch.from_numpy((1.0 + rng.random((3,) + d)).astype(np.float32)).to(DEV)
    v = torch.zeros_like(g)

    def resid(v):
        return float(((pp.vel2mom(v, k) + h * v) - g).pow(2).sum())

    r0 = resid(v)
    pp.relax_(g, h, k, DEFAULT_BND, 2, v)
    r2 = resid(v)
    pp.relax_(g, h, k, DEFAULT_BND, 4, v)
    r6 = resid(v)
    assert r2 < 0.5 * r0 and r6 < r2
    # and the GPU sweeps equal the oracle's on a slab of the same problem (bit-exact)
    d2 = (16, 24, 128)
    k2 = pp.sparsify(op([1e-3, 0.1, 0.01, 0.05]), d2)
    g2 = rng.standard_normal((3,) + d2).astype(np.float32)
    h2 = (1.0 + rng.random((3,) + d2)).astype(np.float32)
    v2 = T(np.zeros_like(g2))
    pp.relax_(T(g2), T(h2), k2, DEFAULT_BND, 1, v2)
    assert bits_equal(H(v2), osp.relax(np.zeros_like(g2), g2, h2, k2, DEFAULT_BND, 1))


def test_errors():
    d = (6, 7, 8)
    k = pp.sparsify(op([1e-3, 1, 0, 0]), d)
    v = torch.zeros((3,) + d, device=DEV)
    g = torch.zeros((3,) + d, device=DEV)
    with pytest.raises(AssertionError, match="inappropriate boundary"):
        pp.relax_(g, None, k, [(3, 1, 1)] * 3, 1, v)
    with pytest.raises(AssertionError, match="inappropriate boundary"):
        pp.relax_(g, None, k, [(1, 1, 1)] * 2, 1, v)
    with pytest.raises(AssertionError, match="incompatible hessian"):
        pp.relax_(g, torch.zeros((4,) + d, device=DEV), k, DEFAULT_BND, 1, v)
    with pytest.raises(AssertionError, match="incompatible dimensions"):
        pp.vel2mom(torch.zeros((3, 6, 7, 9), device=DEV), k)
    with pytest.raises(pp._lib.PPError, match="different arrays"):
        pp.vel2mom_(v, v, k)
    # an even stencil is refused by the library
    with pytest.raises(pp._lib.PPError, match="even"):
        bad = pp.SparseKernel(stride=(2, 3, 3), d=d, nchan=3, offset=k.offset, length=k.length, values=k.values,
                              indices=k.indices, patch_indices=k.patch_indices)
        pp.vel2mom(v, bad)
    # the host entry point reports through the status and fills the output with NaN
    u = np.zeros((3,) + d, np.float32)
    bad = pp.SparseKernel(stride=(3, 3, 7), d=d, nchan=3, offset=k.offset, length=k.length, values=k.values,
                          indices=k.indices, patch_indices=k.patch_indices)
    with pytest.raises(pp._lib.PPError, match="too big"):
        pp.vel2mom_(u, np.zeros_like(u), bad)
    assert np.isnan(u).all()


def test_graph_replays():
    from pushpull import batch as ppb

    d = (12, 13, 40)
    k = pp.sparsify(op([1e-3, 0.1, 0.01, 0.05]), d)
    rng = np.random.default_rng(11)
    g = T(rng.standard_normal((3,) + d).astype(np.float32))
    h = T((1.0 + rng.random((3,) + d)).astype(np.float32))
    v = torch.zeros_like(g)
    want = osp.relax(np.zeros((3,) + d, np.float32), H(g), H(h), k, DEFAULT_BND, 2)
    k.handle(0)  # upload before the capture: nothing is allocated while capturing
    graph = ppb.Graph(lambda stream: pp.relax_(g, h, k, DEFAULT_BND, 2, v))
    for _ in range(3):
        v.zero_()
        torch.cuda.synchronize()
        graph.launch()
        graph.stream.synchronize()
        assert bits_equal(H(v), want)
    graph.close()


@pytest.mark.parametrize("nd", [2, 4])
def test_other_channel_counts_bit_exact(nd):
    d = (10, 11, 70)
    k = pp.sparsify(op([1e-3, 0.2, 0.01, 0.05]), d, nd=nd)
    rng = np.random.default_rng(40 + nd)
    bnd = [(1, 2, 0), (2, 1, 1), (0, 0, 1), (1, 1, 2)][:nd]
    v = rng.standard_normal((nd,) + d).astype(np.float32)
    print(f"channels {nd}", flush=True)
    assert bits_equal(H(pp.vel2mom(T(v), k, bnd)), osp.vel2mom(v, k, bnd))
    g = rng.standard_normal((nd,) + d).astype(np.float32)
    h = np.zeros((nd * (nd + 1) // 2,) + d, np.float32)
    h[:nd] = 1.0 + rng.random((nd,) + d)
    h[nd:] = 0.1 * rng.standard_normal(h[nd:].shape)
    vt = T(v)
    pp.relax_(T(g), T(h), k, bnd, 2, vt)
    assert bits_equal(H(vt), osp.relax(v, g, h, k, bnd, 2))


def test_overwrite_mode_and_host_hessian_layouts():
    d = (9, 10, 67)
    k = pp.sparsify(op([1e-3, 0.1, 0.01, 0.05]), d)
    rng = np.random.default_rng(50)
    v = rng.standard_normal((3,) + d).astype(np.float32)
    # pp_vel2mom(..., accumulate_edges = 0) overwrites u everywhere
    ut = T(rng.standard_normal((3,) + d).astype(np.float32))
    b = pp.regulariser._bnd(DEFAULT_BND, 3)
    L = pp._lib.lib()
    pp._lib.check(L.pp_vel2mom(k.handle(0), pp.ops._ptr(ut), pp.ops._ptr(T(v)), b, 0, pp.ops._stream(ut)), "pp_vel2mom")
    assert bits_equal(H(ut), osp.vel2mom(v, k, DEFAULT_BND))
    # the host entry point with no H and with a diagonal H
    g = rng.standard_normal((3,) + d).astype(np.float32)
    for hcode in (0, 1):
        h = _h(rng, hcode, d)
        vv = v.copy()
        pp.relax_(g, h, k, DEFAULT_BND, 2, vv)
        assert bits_equal(vv, osp.relax(v, g, h, k, DEFAULT_BND, 2))
