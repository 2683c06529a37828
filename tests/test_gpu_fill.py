"""pp_memset_zero (the zero fill before every push-family launch) on unaligned pointers and
odd counts: exactly [p, p + count) is zeroed, the neighbours keep their values (needs a GPU)."""
import ctypes as ct

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pp = pytest.importorskip("pushpull")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("offset", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("count", [0, 1, 2, 3, 4, 5, 7, 8, 9, 1023, 4096, 100003])
def test_zero_fill_exact_range(offset, count):
    buf = torch.full((count + 16,), 7.0, device="cuda:0")
    L = pp._lib.lib()
    p = ct.c_void_p(buf.data_ptr() + 4 * offset)
    pp._lib.check(L.pp_memset_zero(p, count, ct.c_void_p(torch.cuda.current_stream().cuda_stream)), "fill")
    got = buf.cpu().numpy()
    ref = np.full(count + 16, 7.0, np.float32)
    ref[offset:offset + count] = 0.0
    assert np.array_equal(got, ref)
