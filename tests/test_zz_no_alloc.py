"""pdot_round_device_noalloc, the allocation-free rounding of a device-resident
plan, against pdot_round_matrix (the same rounding of a host plan, which
allocates its device buffers and workspace): bit for bit, on two calls in a
row (the second smaller), in place and out of place, and its error handling
(sizes above its static workspace, host pointers, allocations smaller than the
stated sizes)."""
import ctypes

import numpy as np
import pytest

import pdot_amd as P
from pdot_amd import _lib, instances

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libpdot_hip.so uses."""

    def __init__(self):
        _lib.load_library()
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.held = []

    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.held.append(p.value)
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64) if arr.ndim == 1 else np.asfortranarray(arr, np.float64)
        d = self.malloc(arr.nbytes)
        assert self.rt.hipMemcpy(d, arr.ctypes.data, arr.nbytes, H2D) == 0
        return d

    def download(self, d, shape):
        out = np.empty(shape, dtype=np.float64, order="F")
        assert self.rt.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0
        return out

    def free_all(self):
        for p in self.held:
            self.rt.hipFree(p)
        self.held = []


def _lib_handle():
    return _lib.load_library()


def _plan(n, m, seed):
    """A nonnegative plan near, but off, the transport polytope of (a, b)."""
    prob = instances.uniform_sqeuclid(n, m, seed=seed)
    rng = np.random.default_rng(seed)
    a, b = prob.source_distribution, prob.target_distribution
    X = np.asfortranarray(np.outer(a, b) * rng.uniform(0.5, 1.5, size=(n, m)))
    X[rng.uniform(size=(n, m)) < 0.05] = 0.0
    return a, b, X


def _round_matrix(L, a, b, X):
    Xr = np.array(X, dtype=np.float64, order="F", copy=True)
    err = ctypes.create_string_buffer(512)
    n, m = X.shape
    rc = L.pdot_round_matrix(0, n, m, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             b.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             Xr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), err, 512)
    assert rc == 0, err.value
    return Xr


def _noalloc(L, n, m, da, db, dX, ldx, dY, ldy):
    err = ctypes.create_string_buffer(512)
    rc = L.pdot_round_device_noalloc(0, None, n, m, da, db, dX, ldx, dY, ldy, err, 512)
    return rc, err.value.decode()


def test_noalloc_rounding_matches_round_matrix_twice():
    L, hip = _lib_handle(), Hip()
    try:
        for (n, m, seed) in [(1500, 1300, 3), (700, 257, 4)]:
            print(f"noalloc rounding {n}x{m} seed {seed}", flush=True)
            a, b, X = _plan(n, m, seed)
            ref = _round_matrix(L, a, b, X)
            da, db, dX = hip.upload(a), hip.upload(b), hip.upload(X)
            dY = hip.malloc(8 * n * m)
            rc, msg = _noalloc(L, n, m, da, db, dX, n, dY, n)
            assert rc == 0, msg
            Y = hip.download(dY, (n, m))
            assert np.array_equal(Y, ref)
            # the input is left alone out of place
            assert np.array_equal(hip.download(dX, (n, m)), X)
            # feasibility of the rounded plan
            assert np.allclose(Y.sum(axis=1), a, rtol=0, atol=1e-15)
            assert np.allclose(Y.sum(axis=0), b, rtol=0, atol=1e-15)
            # in place
            rc, msg = _noalloc(L, n, m, da, db, dX, n, dX, n)
            assert rc == 0, msg
            assert np.array_equal(hip.download(dX, (n, m)), ref)
    finally:
        hip.free_all()


def test_noalloc_rounding_with_leading_dimensions():
    """A plan stored with padding rows (ldx > n) rounded into a padded output."""
    L, hip = _lib_handle(), Hip()
    try:
        n, m, ldx, ldy = 333, 190, 384, 400
        a, b, X = _plan(n, m, 9)
        ref = _round_matrix(L, a, b, X)
        Xp = np.full((ldx, m), 7.0, order="F")
        Xp[:n] = X
        da, db, dX = hip.upload(a), hip.upload(b), hip.upload(Xp)
        dY = hip.upload(np.full((ldy, m), -3.0, order="F"))
        rc, msg = _noalloc(L, n, m, da, db, dX, ldx, dY, ldy)
        assert rc == 0, msg
        Y = hip.download(dY, (ldy, m))
        assert np.array_equal(Y[:n], ref)
        assert np.all(Y[n:] == -3.0)
    finally:
        hip.free_all()


def test_noalloc_rounding_rejects_bad_calls():
    L, hip = _lib_handle(), Hip()
    try:
        n, m = 64, 48
        a, b, X = _plan(n, m, 1)
        da, db, dX = hip.upload(a), hip.upload(b), hip.upload(X)
        dY = hip.malloc(8 * n * m)
        # above the static workspace (checked before any pointer is used)
        rc, msg = _noalloc(L, 60000, 60000, da, db, dX, 60000, dY, 60000)
        assert rc == _lib.PDOT_E_ARG and "maximum" in msg, msg
        # host pointers
        rc, msg = _noalloc(L, n, m, a.ctypes.data, db, dX, n, dY, n)
        assert rc == _lib.PDOT_E_ARG and "device memory" in msg, msg
        # an allocation smaller than the stated plan
        rc, msg = _noalloc(L, n + 8, m, da, db, dX, n + 8, dY, n + 8)
        assert rc == _lib.PDOT_E_ARG, msg
        # overlapping distinct output, bad leading dimension, bad device
        rc, msg = _noalloc(L, n, m // 2, da, db, dX, n, dX + 8 * n, n)
        assert rc == _lib.PDOT_E_ARG and "overlaps" in msg, msg
        rc, msg = _noalloc(L, n, m, da, db, dX, n - 1, dY, n)
        assert rc == _lib.PDOT_E_ARG, msg
        err = ctypes.create_string_buffer(256)
        assert L.pdot_round_device_noalloc(99, None, n, m, da, db, dX, n, dY, n, err, 256) == _lib.PDOT_E_ARG
        # a good call afterwards still works
        rc, msg = _noalloc(L, n, m, da, db, dX, n, dY, n)
        assert rc == 0, msg
        assert np.array_equal(hip.download(dY, (n, m)), _round_matrix(L, a, b, X))
    finally:
        hip.free_all()
