"""
mars_hip_fit_append_noalloc (the append without allocation) against
mars_hip_fit_append, through the C ABI: the same mse and bitwise the same
basis, on two contexts in a row, the second with fewer rows (the shared
scratch keeps the first call's contents).
"""

import ctypes as C

import numpy as np
import pytest

from mars_amd import datagen
from tests import abi

pytestmark = pytest.mark.gpu


def _create(lib, X, y, max_terms):
    n, p = X.shape
    w = np.ones(n, dtype=np.float32)
    ctx = C.c_void_p()
    abi.check(lib, lib.mars_hip_fit_create(abi.ptr(X), n, abi.ptr(y), abi.ptr(w), n, p, max_terms, 0, C.byref(ctx)))
    return ctx


def _basis(lib, ctx, n, m):
    B = np.zeros((n, m), dtype=np.float32)
    Bo = np.zeros((n, m), dtype=np.float32)
    got = C.c_size_t()
    abi.check(lib, lib.mars_hip_fit_basis(ctx, abi.ptr(B), abi.ptr(Bo), C.byref(got)))
    assert got.value == m
    return B, Bo


# (type, xcol, bcol, knot quantile) appended in order; the third reuses a product of the first
STEPS = [(b"l", 0, 0, None), (b"+", 3, 1, 0.4), (b"-", 2, 0, 0.6), (b"l", 1, 2, None), (b"+", 3, 0, 0.4)]


@pytest.mark.parametrize("n", [30000, 7000])
def test_append_noalloc_matches_append(n):
    lib = abi.load()
    lib.mars_hip_fit_append_noalloc.argtypes = lib.mars_hip_fit_append.argtypes
    X, y = datagen.workload2(n=n, p=6)
    print("no-alloc append vs append: n=%d" % n, flush=True)
    ref = _create(lib, X, y, 9)
    new = _create(lib, X, y, 9)
    sz = C.c_size_t
    lib.mars_hip_device_bytes.argtypes = [C.POINTER(sz), C.POINTER(sz), C.c_int]
    try:
        live0, live1, peak = sz(), sz(), sz()
        abi.check(lib, lib.mars_hip_device_bytes(C.byref(live0), C.byref(peak), 1))
        got = []
        for t, xc, bc, q in STEPS:
            h = 0.0 if q is None else float(np.quantile(X[:, xc], q))
            b = C.c_double()
            abi.check(lib, lib.mars_hip_fit_append_noalloc(new, t, xc, bc, C.c_float(h), C.byref(b)))
            got.append(b.value)
        abi.check(lib, lib.mars_hip_device_bytes(C.byref(live1), C.byref(peak), 0))
        assert live1.value == live0.value and peak.value == live0.value  # no library allocation
        m = 1
        for (t, xc, bc, q), b in zip(STEPS, got):
            h = 0.0 if q is None else float(np.quantile(X[:, xc], q))
            a = C.c_double()
            abi.check(lib, lib.mars_hip_fit_append(ref, t, xc, bc, C.c_float(h), C.byref(a)))
            assert a.value == b, (t, xc, bc, a.value, b)
            m += 1 if a.value >= 0 else 0
        for u, v in zip(_basis(lib, ref, n, m), _basis(lib, new, n, m)):
            np.testing.assert_array_equal(u, v)
        # its checks are the append's
        d = C.c_double()
        assert lib.mars_hip_fit_append_noalloc(new, b"l", 0, 99, C.c_float(0), C.byref(d)) == -1
        assert b"invalid basis column number: 99" in lib.mars_hip_last_error()
    finally:
        lib.mars_hip_context_destroy(ref)
        lib.mars_hip_context_destroy(new)
