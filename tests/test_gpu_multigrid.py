"""Push with a count image (pp_push_count, pp_affine_push_count), invert_def's pushes and
restrict / prolong on the GPU against the oracle (needs a GPU).

Bar: NaN masks identical; norm-wise relative error <= 1e-5 against the FP32 oracle and the
FP64 oracle. Cases cover the fused launch (triples that fold alike: equal, or differing
between 1 and 2), the split fallback (a circulant channel with a reflect count; degrees 0
and 4), the identity values of invert_def, batches, and accumulation.
"""
import numpy as np
import pytest

import oracle
from helpers import S, check_close, max_abs, smooth_field

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def T(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def H(t):
    return t.detach().cpu().numpy()


@pytest.fixture(params=["auto", "per-sample"])
def tiling(request):
    old = pp._lib.get_option(pp._lib.PP_OPT_TILING)
    pp._lib.set_option(pp._lib.PP_OPT_TILING, 1 if request.param == "auto" else 0)
    yield request.param
    pp._lib.set_option(pp._lib.PP_OPT_TILING, old)


def _phi(rng, nb, d1, d0, out_frac=0.03):
    phi = np.stack([oracle.identity(d1) * (np.array(d0)[::-1, None, None, None] / np.array(d1)[::-1, None, None, None])
                    + smooth_field(rng, d1, 3.0, 3.0) for _ in range(nb)]).astype(np.float32)
    m = rng.random((nb,) + d1) < out_frac
    phi[:, 1][m] = rng.uniform(-5, 0, m.sum()).astype(np.float32)
    return phi


CASES = [
    # (deg, channel triples, count triple, ext)
    (1, [(1, 1, 1)], None, 1),
    (2, [(2, 1, 1), (1, 2, 1), (1, 1, 2)], (1, 1, 1), 0),
    (3, [(0, 0, 0), (0, 0, 0)], (0, 0, 0), 1),
    (3, [(0, 1, 2), (1, 1, 1)], (1, 1, 1), 1),   # circulant x in channel 0: split launches
    (0, [(1, 1, 1)], (2, 2, 2), 1),
    (4, [(1, 2, 1)], (1, 1, 1), 0),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_push_count_matches_oracle(tiling, case):
    deg, bnd, cb, ext = CASES[case]
    rng = np.random.default_rng(20 + case)
    nc, d1, d0 = len(bnd), (14, 17, 19), (12, 16, 21)
    phi = _phi(rng, 2, d1, d0)
    f1 = rng.standard_normal((2, nc) + d1).astype(np.float32)
    sett = pp.Settings(deg, bnd if len(bnd) > 1 else bnd[0], ext)
    f0, c0 = pp.push_count(T(f1), T(phi), d0, sett, count_bnd=cb)
    osett = S(deg, bnd, ext)
    for prec in ("f32", "f64"):
        rf0, rc0 = oracle.push_count(f1, phi, d0, osett, cb, prec)
        check_close(H(f0), rf0, f"f0 {prec}", scale=max_abs(f1))
        check_close(H(c0), rc0, f"c0 {prec}", scale=1.0)


def test_push_count_identity_and_accumulate(tiling):
    rng = np.random.default_rng(30)
    d = (15, 13, 18)
    phi = _phi(rng, 1, d, d)
    sett = pp.Settings(1, oracle.SLIDING, 0)
    f0, c0 = pp.push_count(None, T(phi), d, sett, count_bnd=(1, 1, 1))
    rf0, rc0 = oracle.push_count(None, phi, d, S(1, oracle.SLIDING, 0), (1, 1, 1), "f64")
    check_close(H(f0), rf0, "identity f0", scale=float(max(d)))
    check_close(H(c0), rc0, "identity c0", scale=1.0)
    # accumulate=1 adds onto what the buffers hold
    L = pp._lib.lib()
    desc = pp.multigrid._count_desc(sett, d, d, 3, 1, (1, 1, 1))
    import ctypes as ct
    p = pp.ops._ptr
    pp._lib.check(L.pp_push_count(ct.byref(desc), p(f0), p(c0), p(T(phi)), None, 1, pp.ops._stream(f0)), "acc")
    check_close(H(f0), 2 * rf0, "identity f0 twice", scale=float(max(d)))
    check_close(H(c0), 2 * rc0, "c0 twice", scale=1.0)


def test_invert_def_rhs_matches_oracle(tiling):
    rng = np.random.default_rng(31)
    d = (16, 14, 20)
    phi = (oracle.identity(d) + smooth_field(rng, d, 3.0, 2.5))[None].astype(np.float32)
    g, h = pp.invert_def_rhs(T(phi))
    for prec in ("f32", "f64"):
        rg, rh = oracle.invert_def_rhs(phi, prec)
        check_close(H(g), rg, f"g {prec}", scale=float(max(d)))
        check_close(H(h), rh, f"h {prec}", scale=1.0)


@pytest.mark.parametrize("deg", [1, 2, 3])
def test_restrict_prolong_match_oracle(tiling, deg):
    rng = np.random.default_rng(40 + deg)
    f1 = rng.standard_normal((1, 2, 21, 18, 25)).astype(np.float32)
    d0 = (11, 9, 13)
    got = H(pp.restrict(T(f1), d0, deg))
    check_close(got, oracle.restrict(f1, d0, deg, "f32"), "restrict f32", scale=max_abs(f1))
    check_close(got, oracle.restrict(f1, d0, deg, "f64"), "restrict f64", scale=max_abs(f1))
    u = rng.standard_normal((1, 3, 7, 6, 8)).astype(np.float32)
    got = H(pp.prolong(T(u), (14, 12, 16), deg))
    check_close(got, oracle.prolong(u, (14, 12, 16), deg, "f32"), "prolong f32", scale=max_abs(u))
    check_close(got, oracle.prolong(u, (14, 12, 16), deg, "f64"), "prolong f64", scale=max_abs(u))


def test_host_arrays_match_device():
    rng = np.random.default_rng(50)
    d = (9, 10, 11)
    phi = _phi(rng, 1, d, d)
    f1 = rng.standard_normal((1, 2) + d).astype(np.float32)
    sett = pp.Settings(2, 1, 1)
    a0, a1 = pp.push_count(f1, phi, d, sett)
    b0, b1 = pp.push_count(T(f1), T(phi), d, sett)
    check_close(a0, H(b0), "host f0", scale=max_abs(f1))
    check_close(a1, H(b1), "host c0", scale=1.0)


def test_push_count_errors():
    import ctypes as ct

    d = (4, 5, 6)
    phi = T(oracle.identity(d)[None])
    with pytest.raises(AssertionError):
        pp.push_count(None, phi, d, pp.Settings(1, [(1, 1, 1), (1, 1, 1)], 1))
    L, p = pp._lib.lib(), pp.ops._ptr
    out = torch.zeros((2,) + d, device=DEV)
    cnt = torch.zeros(d, device=DEV)
    desc = pp.ops._desc(pp.Settings(1, 1, 1), d, d, 2, 1)
    s = pp.ops._stream(phi)
    assert L.pp_push_count(ct.byref(desc), p(out), p(cnt), p(phi), None, 0, s) == pp._lib.PP_EINVAL  # identity: 3 ch
    assert L.pp_push_count(ct.byref(desc), p(out), None, p(phi), p(out), 0, s) == pp._lib.PP_EINVAL  # no count
