"""lbm_surrogate_forward_noalloc against lbm_surrogate_forward, through the C
ABI.

The allocation-free entry point stages its host rows in a fixed device array
(up to 65536 rows) instead of allocating device buffers per call; it runs the
same inference kernel, so the outputs must be bitwise those of
lbm_surrogate_forward, on two calls in a row, the second with fewer rows (the
array keeps the first call's rows beyond the second's).  Rows beyond the
limit are refused with LBM_E_ARG.
"""
import ctypes

import numpy as np
import pytest

from lbm_amd import cases
from lbm_amd.capi import LBM_FP64, LBM_MODE_STRICT, MixtureLattice

pytestmark = pytest.mark.gpu

LIMIT = 65536


def forward_noalloc(lat, X):
    L = lat._L
    L.lbm_surrogate_forward_noalloc.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int64,
                                                ctypes.POINTER(ctypes.c_float)]
    L.lbm_surrogate_forward_noalloc.restype = ctypes.c_int
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros_like(X)
    fp = ctypes.POINTER(ctypes.c_float)
    rc = L.lbm_surrogate_forward_noalloc(lat._h, X.ctypes.data_as(fp), X.shape[0], Y.ctypes.data_as(fp))
    return rc, Y


def test_noalloc_forward_matches_forward_on_two_calls():
    lat = MixtureLattice(4, 4, 4, *cases.UNITS_FLAME, reaction=True, mode=LBM_MODE_STRICT, precision=LBM_FP64)
    try:
        lat.surrogate_init(7)
        rng = np.random.default_rng(3)
        for m in (LIMIT, 1000):
            print(f"noalloc forward: {m} rows", flush=True)
            X = rng.standard_normal((m, 16)).astype(np.float32)
            rc, Y = forward_noalloc(lat, X)
            assert rc == 0, lat._L.lbm_last_error(lat._h)
            ref = lat.surrogate_forward(X)
            assert np.array_equal(Y, ref)
            assert np.isfinite(Y).all() and np.any(Y != 0)
        rc, _ = forward_noalloc(lat, np.zeros((LIMIT + 1, 16), np.float32))
        assert rc == -1  # LBM_E_ARG
    finally:
        lat.close()
