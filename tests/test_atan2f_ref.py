"""The library's atan2f (used on the device for wall angles) against the host libm's atan2f, which
the reference's f32::atan2 calls on Linux: bit for bit on random quotients in the range drum
positions take, on random bit patterns (NaN, infinities, subnormals, extremes) and on the special
cases of the algorithm (zeros of both signs, x = 1, infinite arguments)."""
import ctypes

import numpy as np
import pytest


@pytest.fixture(scope="module")
def libm_atan2f():
    libm = ctypes.CDLL("libm.so.6")
    f = libm.atan2f
    f.restype = ctypes.c_float
    f.argtypes = [ctypes.c_float, ctypes.c_float]
    return lambda y, x: np.array([f(float(a), float(b)) for a, b in zip(y, x)], dtype=np.float32)


def _same(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    both_nan = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) == b.view(np.uint32)) | both_nan


def test_drum_range(libm_atan2f):
    from milldynamics3_amd.metrics import atan2f_ref
    rng = np.random.default_rng(11)
    y = ((rng.random(300000) - 0.5) * 0.2).astype(np.float32)
    x = ((rng.random(300000) - 0.5) * 0.2).astype(np.float32)
    assert _same(atan2f_ref(y, x), libm_atan2f(y, x)).all()


def test_random_bit_patterns(libm_atan2f):
    from milldynamics3_amd.metrics import atan2f_ref
    rng = np.random.default_rng(12)
    y = rng.integers(0, 2 ** 32, 200000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    x = rng.integers(0, 2 ** 32, 200000, dtype=np.uint64).astype(np.uint32).view(np.float32)
    assert _same(atan2f_ref(y, x), libm_atan2f(y, x)).all()


def test_special_cases(libm_atan2f):
    from milldynamics3_amd.metrics import atan2f_ref
    v = np.array([0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4e38, 2.0 ** -30, 0.4375,
                  0.6875, 1.1875, 2.4375, 2.0 ** 25, 2.0 ** 26], dtype=np.float32)
    y, x = np.meshgrid(v, v)
    y, x = y.ravel(), x.ravel()
    assert _same(atan2f_ref(y, x), libm_atan2f(y, x)).all()
