"""The point-cost kernels divide each formed cost by the scale s (the largest
cost) with a multiply by y = RN(1/s) and two Markstein corrections instead of
an IEEE division.  Markstein's theorem (y within half an ulp of 1/s, the
corrected quotient within one ulp of a/s) makes the result the correctly
rounded a/s; this checks it bit for bit on the CPU, on random samples over
many scales (with one-ulp neighbours), on exact multiples and on the costs of
a real point-cloud instance, against C's division.  The kernel's own results
are pinned on the GPU by the point-cloud solves equal to the dense ones."""
import ctypes
import subprocess

import numpy as np

from pdot_amd import instances

import os

HERE = os.path.dirname(os.path.abspath(__file__))


def _lib(tmp_path):
    so = str(tmp_path / "quotient_check.so")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    os.path.join(HERE, "native", "quotient_check.c"), "-lm"], check=True)
    lib = ctypes.CDLL(so)
    lib.quotient_mismatches.restype = ctypes.c_long
    lib.quotient_mismatches.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_long, ctypes.c_double]
    lib.quotient_mismatches_random.restype = ctypes.c_long
    lib.quotient_mismatches_random.argtypes = [ctypes.c_double, ctypes.c_long, ctypes.c_uint64]
    lib.quotient_mismatches_tiny.restype = ctypes.c_long
    lib.quotient_mismatches_tiny.argtypes = [ctypes.c_double, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_uint64]
    return lib


def test_quotient_sequence_is_the_correctly_rounded_division(tmp_path):
    lib = _lib(tmp_path)
    rng = np.random.default_rng(7)
    scales = np.concatenate([[1.0, 3.0, 0.1, 7.0 / 3.0, 2.0 ** 0.5, np.nextafter(2.0, 0), np.nextafter(1.0, 2),
                              1e305, 3e-305, 2.0 ** 1000, 2.0 ** -1000],
                             10.0 ** rng.uniform(-6, 6, 25), rng.uniform(1, 2, 25)])
    for k, s in enumerate(scales):
        assert lib.quotient_mismatches_random(float(s), 200_000, 1000 + k) == 0, s
    # exact multiples and tiny, zero and edge values
    for s in (3.0, 0.7, 1234.5678):
        a = np.concatenate([s * np.arange(0, 1025) / 1024, [0.0, 5e-324, 1e-300, s, np.nextafter(s, 0)]])
        a = np.ascontiguousarray(a)
        assert lib.quotient_mismatches(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a), s) == 0


def test_quotient_on_instance_costs(tmp_path):
    lib = _lib(tmp_path)
    pts = instances.gaussian_shift_points(700, 500, 3)
    x, y = pts.source_points, pts.target_points
    d = np.zeros((x.shape[0], y.shape[0]))
    for k in range(x.shape[1]):  # the generator's operation order: sum over coordinates of squares
        t = y[None, :, k] - x[:, None, k]
        d += t * t
    s = float(d.max())
    a = np.ascontiguousarray(d.ravel())
    assert lib.quotient_mismatches(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(a), s) == 0


def test_quotient_tiny_numerators_at_small_scales(tmp_path):
    """Scales below 2^-122 (where the relative bound a >= 2^-900 s underflows),
    not powers of two, with subnormal and near-subnormal numerators: the
    absolute bound a >= 2^-960 sends the numerators whose residuals could be
    inexact to the division."""
    lib = _lib(tmp_path)
    rng = np.random.default_rng(11)
    scales = [float.fromhex("0x1.bc3d8f9ca9bf6p-167"), 3e-300, 7.0 / 3.0 * 2.0 ** -500, 2.0 ** -123 * 1.5,
              1e-37 * np.pi] + list(10.0 ** rng.uniform(-300, -37, 20)) + list(10.0 ** rng.uniform(-6, 6, 5))
    for k, s in enumerate(scales):
        assert lib.quotient_mismatches_tiny(float(s), 100_000, -1074, -900, 500 + k) == 0, s
        assert lib.quotient_mismatches_tiny(float(s), 50_000, -970, -950, 900 + k) == 0, s
    # the case that showed the underflowing guard
    a = np.array([float.fromhex("0x0.0000000000435p-1022")])
    s = float.fromhex("0x1.bc3d8f9ca9bf6p-167")
    assert lib.quotient_mismatches(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1, s) == 0
