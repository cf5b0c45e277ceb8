"""xp_dot_f64_device, the allocation-free entry point, through the C-ABI:
equal to dot_kernel (the entry point it varies: same kernels, split and
summation order, so the same bits) and within the FP64 reduction bound
n eps sum |a_i b_i| of numpy, on two calls in a row, the second smaller (the
shared scratch array keeps the first call's partials)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import xplus_amd as X

pytestmark = pytest.mark.gpu


def test_no_alloc_dot_matches_dot_kernel(gpu):
    f = X._lib.xp_dot_f64_device
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
    rng = np.random.default_rng(7)
    for n in (3_000_001, 777):
        a = rng.standard_normal(n)
        b = rng.standard_normal(n)
        da, db, dout = X.DeviceArray.from_numpy(a), X.DeviceArray.from_numpy(b), X.DeviceArray(1, np.float64)
        rc = f(da.ptr, db.ptr, dout.ptr, n, None)
        assert rc == 0, X.last_error()
        X.synchronize()
        got = float(dout.numpy()[0])
        ref = np.zeros(1)
        X.dot_kernel(a, b, ref, n)
        exact = float(np.dot(a, b))
        print(f"n={n}: no-alloc {got!r} dot_kernel {ref[0]!r} numpy {exact!r}", flush=True)
        assert got == ref[0]
        assert abs(got - exact) <= n * np.finfo(np.float64).eps * float(np.abs(a * b).sum())
        for d in (da, db, dout):
            d.close()
    # n < 0: the reference's silent no-op; null pointers: an error
    assert f(None, None, None, -5, None) == 0
    assert f(None, None, None, 10, None) != 0
