"""compose / Exp on the GPU (pp_compose, pp_exp) against the oracle (needs a GPU).

Bar: NaN masks identical; norm-wise relative error <= 1e-5 against the FP32 oracle (the
reference's own arithmetic) and the FP64 oracle for compose. Exp returns Id + u in Float32,
whose spacing at the coordinate (1.9e-6 near 28) dwarfs the displacement's own error, so
every entry must lie within one Float32 spacing of the oracle's deformation plus 1e-5 of the
largest displacement, against the FP32 chain and against the FP64 chain.
"""
import ctypes as ct

import numpy as np
import pytest

import oracle
from helpers import check_close, max_abs, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(params=["auto", "tiled", "per-sample"])
def tiling(request):
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    pp._lib.set_option(pp._lib.PP_OPT_TILING, {"auto": 1, "tiled": 2, "per-sample": 0}[request.param])
    yield request.param
    pp._lib.set_option(pp._lib.PP_OPT_TILING, old)


def _fields(rng, nb, d0, d1, out_frac=0.05):
    phi = np.stack([smooth_field(rng, d0, 2.0, 3.0) for _ in range(nb)])
    theta = np.stack([oracle.identity(d1) + smooth_field(rng, d1, 3.0, 5.0) for _ in range(nb)])
    # send some samples outside the field of view so the sliding signs matter
    m = rng.random((nb,) + d1) < out_frac
    theta[:, 0][m] = -rng.uniform(0.5, 6, m.sum()).astype(np.float32)
    return phi.astype(np.float32), theta.astype(np.float32)


@pytest.mark.parametrize("d0,d1", [((21, 18, 35), (21, 18, 35)), ((17, 29, 23), (12, 31, 40)), ((1, 9, 8), (4, 5, 6))])
def test_compose_matches_oracle(tiling, d0, d1):
    rng = np.random.default_rng(11)
    phi, theta = _fields(rng, 2, d0, d1)
    theta[1, 2, 0, 0, 0] = np.nan  # invalid sample -> NaN in all three channels
    got = H(pp.compose(T(phi), T(theta)))
    ref = oracle.compose(phi, theta, "f32")
    check_close(got, ref, f"compose {d0}->{d1}", scale=max_abs(phi))
    check_close(got, oracle.compose(phi, theta, "f64"), f"compose f64 {d0}->{d1}", scale=max_abs(phi))
    assert np.isnan(got[1, :, 0, 0, 0]).all()


def test_compose_in_place_and_host_arrays():
    rng = np.random.default_rng(12)
    phi, theta = _fields(rng, 1, (10, 11, 12), (10, 11, 12))
    ref = oracle.compose(phi, theta, "f32")
    psi = torch.empty(theta.shape, dtype=torch.float32, device=DEV)
    assert pp.compose_(psi, T(phi), T(theta)) is psi
    check_close(H(psi), ref, "compose!", scale=max_abs(phi))
    check_close(pp.compose(phi, theta), ref, "compose (host arrays)", scale=max_abs(phi))


def _check_exp(got, ref, what):
    """Every entry within one Float32 spacing of ref plus 1e-5 of the largest displacement."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    assert got.shape == ref.shape and np.array_equal(np.isnan(got), np.isnan(ref)), what
    m = np.isfinite(ref)
    Id = oracle.identity(ref.shape[-3:])
    umax = float(np.abs((ref - Id)[m]).max())
    bound = np.spacing(np.abs(ref[m]).astype(np.float32)).astype(np.float64) + 1e-5 * umax
    err = np.abs(got[m] - ref[m])
    assert (err <= bound).all(), f"{what}: worst excess {float((err - bound).max()):.3g} (max |u| {umax:.3g})"


@pytest.mark.parametrize("d,nb", [((24, 20, 28), 1), ((13, 17, 9), 2)])
def test_exp_matches_oracle(tiling, d, nb):
    rng = np.random.default_rng(13)
    v = np.stack([smooth_field(rng, d, 3.0, 4.0) for _ in range(nb)])
    got = H(pp.Exp(T(v)))
    _check_exp(got, oracle.exp(v, 8, "f32"), "Exp vs FP32 chain")
    _check_exp(got, oracle.exp(v, 8, "f64"), "Exp vs FP64 chain")


def test_exp_squarings_and_host_path():
    rng = np.random.default_rng(14)
    d = (9, 10, 11)
    v = smooth_field(rng, d, 2.0, 3.0)[None]
    for nsq in (1, 2, 5):
        _check_exp(H(pp.Exp(T(v), nsq)), oracle.exp(v, nsq, "f64"), f"Exp nsq={nsq}")
    _check_exp(pp.Exp(v[0])[None], oracle.exp(v, 8, "f64"), "Exp host")


def test_exp_graph_replays():
    rng = np.random.default_rng(15)
    d = (16, 12, 20)
    v = T(smooth_field(rng, d, 3.0, 4.0)[None])
    out, work = torch.empty_like(v), torch.empty_like(v)
    dims = pp.compositions._xyz(d)
    L = pp._lib.lib()

    def run(stream=None):
        s = ct.c_void_p(stream.cuda_stream) if stream is not None else pp.ops._stream(v)
        pp._lib.check(L.pp_exp(pp.ops._ptr(out), pp.ops._ptr(v), pp.ops._ptr(work), dims, 1, 8, s), "pp_exp")

    run()
    torch.cuda.synchronize()
    eager = H(out).copy()
    from pushpull import batch as ppb

    g = ppb.Graph(run)
    for _ in range(3):
        out.fill_(0)
        work.fill_(7)
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        assert np.array_equal(H(out), eager)
    g.close()


def test_composition_errors():
    v = torch.zeros((1, 3, 4, 5, 6), device=DEV)
    with pytest.raises(AssertionError):
        pp.Exp(v, 0)
    with pytest.raises(AssertionError):
        pp.compose(torch.zeros((1, 2, 4, 5, 6), device=DEV), v)
    L = pp._lib.lib()
    dims = pp.compositions._xyz((4, 5, 6))
    s = pp.ops._stream(v)
    w = torch.empty_like(v)
    p = pp.ops._ptr
    assert L.pp_exp(p(v), p(v), p(w), dims, 1, 8, s) == pp._lib.PP_EINVAL  # out aliases v
    assert L.pp_compose(p(v), p(v), p(w), dims, dims, 1, s) == pp._lib.PP_EINVAL  # psi aliases phi
