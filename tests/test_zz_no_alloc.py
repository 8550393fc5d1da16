"""yamc_rspace_full_noalloc (scratch in a static device array, no
allocation) against yamc_rspace_full through the C ABI: two calls in a row,
the second on a smaller system, each bit-identical to the allocating entry
point on a twin context."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

pytestmark = pytest.mark.gpu


def full(lib, ctx, n, fn):
    from yamc_amd._lib import check, dptr
    comps = np.empty(6)
    uu, fx, fy, fz = (np.empty(n) for _ in range(4))
    ov = C.c_int32()
    check(getattr(lib, fn)(ctx, dptr(comps), dptr(uu), dptr(fx), dptr(fy), dptr(fz), C.byref(ov)))
    return comps, uu, fx, fy, fz, ov.value


def test_noalloc_equals_allocating_entry_point():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from yamc_amd import _lib, config
    from yamc_amd.ewald import EwaldMC
    lib = _lib.load()
    assert hasattr(lib, "yamc_rspace_full_noalloc")
    for name in ("c2", "small"):  # the second call on a smaller system
        s = config(name, seed=9)
        a, b = EwaldMC(s), EwaldMC(s)
        ra = full(lib, a.ctx, s.natom, "yamc_rspace_full_noalloc")
        rb = full(lib, b.ctx, s.natom, "yamc_rspace_full")
        for x, y in zip(ra, rb):
            assert np.array_equal(np.asarray(x), np.asarray(y)), name
        assert np.all(np.isfinite(ra[1])) and abs(ra[0][0]) > 0, name
        a.close()
        b.close()
