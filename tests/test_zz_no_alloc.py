"""cress_gaussel_noalloc, the variant of cress_gaussel that allocates no
device memory (its four arrays are staged in a fixed device array): the
same results as cress_gaussel, bitwise, on two calls in a row, the second
on a smaller box (so stale contents of the larger first call would show),
and sizes beyond the fixed array are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

from cress_amd import host

CRESS_ETOOBIG = -7
SCRATCH_FLOATS = 8 * 1024 * 1024


def _system(rng, ni, nj, nk):
    shape = (ni + 2, nj + 2, nk)
    a = [np.asfortranarray(rng.uniform(-1.0, 1.0, shape).astype(np.float32)) for _ in range(3)]
    ss = np.asfortranarray(rng.uniform(3.0, 5.0, shape).astype(np.float32))
    return [a[0], ss, a[1], a[2]]           # rr, ss, tt, ff


def _ptrs(arrs):
    return [x.ctypes.data_as(C.c_void_p) for x in arrs]


def _call(fn, ni, nj, nk, arrs):
    return fn(2, ni - 2, 2, nj - 2, 3, nk - 2, ni, nj, nk, *_ptrs(arrs))


@pytest.mark.gpu
def test_noalloc_gaussel_equals_gaussel_twice():
    L = host.lib()
    rng = np.random.default_rng(17)
    for ni, nj, nk in ((70, 60, 44), (23, 17, 19)):
        src = _system(rng, ni, nj, nk)
        a = [x.copy(order="F") for x in src]
        b = [x.copy(order="F") for x in src]
        host.check(_call(L.cress_gaussel, ni, nj, nk, a), "cress_gaussel")
        host.check(_call(L.cress_gaussel_noalloc, ni, nj, nk, b), "cress_gaussel_noalloc")
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[3], b[3])
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[2], b[2])
        assert not np.array_equal(b[3], src[3])


def test_noalloc_gaussel_refuses_oversize():
    L = host.lib()
    ni = nj = 160
    nk = SCRATCH_FLOATS // 4 // ((ni + 2) * (nj + 2)) + 1
    assert (ni + 2) * (nj + 2) * nk > SCRATCH_FLOATS // 4
    d = np.zeros(1, np.float32)
    p = d.ctypes.data_as(C.c_void_p)
    assert L.cress_gaussel_noalloc(2, ni - 2, 2, nj - 2, 3, nk - 2, ni, nj, nk, p, p, p, p) == CRESS_ETOOBIG
