"""The Tier B memory contract (needs a GPU): device arrays are hipMalloc'ed memory of the
current device. Host memory (pageable or pinned) and managed memory are refused with
PP_EINVAL before any launch, both by the single-volume entry points and by the batch
ones, and a valid call right after still runs (the refused lookup leaves no error behind).
"""
import ctypes as ct

import numpy as np
import pytest

from helpers import identity

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")
from pushpull import _lib, ops  # noqa: E402
from pushpull import batch as ppb  # noqa: E402

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
D = (8, 9, 10)


def _desc():
    return ops._desc(pp.Settings(3, 1, 1), D, D, 1, 1)


def _pull(out_ptr, phi_ptr, f_ptr):
    d = _desc()
    return _lib.lib().pp_pull(ct.byref(d), ct.c_void_p(out_ptr), ct.c_void_p(phi_ptr), ct.c_void_p(f_ptr), None)


@pytest.fixture
def dev_arrays():
    phi = torch.from_numpy(identity(D)).to(DEV)
    f = torch.ones(D, device=DEV)
    out = torch.empty(D, device=DEV)
    return phi, f, out


def test_host_memory_is_refused(dev_arrays):
    phi, f, out = dev_arrays
    host = np.zeros(int(np.prod(D)), np.float32)
    assert _pull(out.data_ptr(), phi.data_ptr(), host.ctypes.data) == -1
    assert "`f0` is not device memory" in _lib.last_error() or "`f0` is not memory allocated" in _lib.last_error()
    pinned = torch.zeros(3 * int(np.prod(D)), pin_memory=True)
    assert _pull(out.data_ptr(), pinned.data_ptr(), f.data_ptr()) == -1
    assert "`phi` is host memory" in _lib.last_error()
    # a valid call right after
    assert _pull(out.data_ptr(), phi.data_ptr(), f.data_ptr()) == 0
    torch.cuda.synchronize()
    assert torch.allclose(out, torch.ones_like(out))


def test_managed_memory_is_refused(dev_arrays):
    phi, f, out = dev_arrays
    hip = ct.CDLL("libamdhip64.so")
    p = ct.c_void_p()
    assert hip.hipMallocManaged(ct.byref(p), ct.c_size_t(4 * int(np.prod(D))), ct.c_uint(1)) == 0
    try:
        assert _pull(p.value, phi.data_ptr(), f.data_ptr()) == -1
        assert "`out` is managed or unified memory" in _lib.last_error()
    finally:
        assert hip.hipFree(p) == 0
    assert _pull(out.data_ptr(), phi.data_ptr(), f.data_ptr()) == 0


def test_batch_volumes_are_checked(dev_arrays):
    phi, f, _ = dev_arrays
    pinned_out = torch.empty((1,) + D, pin_memory=True)
    vol = ppb.PPVolume()
    vol.f, vol.phi, vol.out = f.data_ptr(), phi.data_ptr(), pinned_out.data_ptr()
    for q in range(3):
        vol.d0[q] = vol.d1[q] = D[2 - q]
        vol.bnd[q] = 1
        vol.deg[q] = 3
    vol.nchan, vol.ext = 1, 1
    L = ppb._sigs()
    h = ct.c_void_p()
    assert L.pp_batch_plan_create(0, ct.byref(vol), 1, ct.byref(h)) == -1
    assert "volume 0: `out` is host memory" in _lib.last_error()
    assert L.pp_batch_run_static(0, ct.byref(vol), 1, None) == -1
    assert "volume 0: `out` is host memory" in _lib.last_error()
