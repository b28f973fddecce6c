"""qedh_dot_many_noalloc (the Gram-Schmidt coefficients with no allocation, its column table,
coefficients and partial sums in the library's fixed device array) against qedh_dot_many on the same
device vectors: equal bit for bit, on two calls in a row (the second with fewer rows and fewer
columns, so the array holds the first call's leftovers), both checked against numpy; the size limit
is refused before anything runs. Every case prints its name before it touches the device."""

import ctypes as C
import sys

import numpy as np
import pytest

from qed_spin_amd import _lib, errors
from qed_spin_amd.sector import DeviceArray

pytestmark = pytest.mark.gpu

MAX_BASIS = 4096   # QEDH_NOALLOC_MAX_BASIS


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _coeffs(name, cplx, table, cols, dw, n):
    out = np.zeros(cols * (2 if cplx else 1))
    _lib.call(name, cplx, table, cols, dw.ptr, n, _lib.ptr(out, C.c_double))
    return out


@pytest.mark.parametrize("cplx", [0, 1])
def test_noalloc_equals_dot_many_twice(cplx):
    rng = np.random.default_rng(17 + cplx)
    for n, cols in ((100_003, 300), (7_001, 9)):
        say("dot_many_noalloc", cplx, n, cols)
        vs = [rng.normal(size=n) + (1j * rng.normal(size=n) if cplx else 0) for _ in range(cols)]
        w = rng.normal(size=n) + (1j * rng.normal(size=n) if cplx else 0)
        dvs = [DeviceArray.from_host(np.ascontiguousarray(v)) for v in vs]
        dw = DeviceArray.from_host(np.ascontiguousarray(w))
        table = (C.c_void_p * cols)(*[d.ptr.value for d in dvs])
        plain = _coeffs("qedh_dot_many", cplx, table, cols, dw, n)
        fixed = _coeffs("qedh_dot_many_noalloc", cplx, table, cols, dw, n)
        assert np.array_equal(plain, fixed)
        got = fixed.view(np.complex128) if cplx else fixed
        ref = np.array([np.vdot(v, w) for v in vs])
        scale = np.linalg.norm(w) * max(np.linalg.norm(v) for v in vs)
        assert np.abs(got - ref).max() <= 1e-13 * scale


def test_noalloc_refuses_more_columns_than_its_array_holds():
    say("dot_many_noalloc limit")
    d = DeviceArray.from_host(np.zeros(8))
    cols = MAX_BASIS + 1
    table = (C.c_void_p * cols)(*([d.ptr.value] * cols))
    with pytest.raises(errors.InvalidRequest):
        _coeffs("qedh_dot_many_noalloc", 0, table, cols, d, 8)
