"""Spline resampling of the warm-start path (host code, no device): the
not-a-knot cubic spline of DEALII_SCFT/src/spline_chen.c:13-107, checked
against that routine compiled from the reference sources (oracle/_ref),
against scipy's not-a-knot CubicSpline, and by exact reproduction of cubics."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from tests.test_capi_symbols import _lib

_dp = np.ctypeslib.ndpointer(np.float64, flags="C")


def resample(x, y, xp):
    L = _lib()
    L.scft_resample_spline.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
    x, y, xp = (np.ascontiguousarray(v, dtype=np.float64) for v in (x, y, xp))
    out = np.empty(len(xp))
    rc = L.scft_resample_spline(len(x), x, y, len(xp), xp, out)
    return rc, out


def _cases():
    rng = np.random.default_rng(3)
    yield np.linspace(0.0, 2.0, 7)                        # uniform
    yield np.sort(rng.uniform(0, 5, 40))                  # irregular
    yield 3.0 * (np.arange(128) / 127.0) ** 2              # wall-refined grid
    yield np.array([0.0, 0.1, 0.5, 2.0])                   # n = 4


@pytest.mark.parametrize("x", list(_cases()))
def test_spline_matches_reference_spline_chen(x):
    if not O.ref.available():
        pytest.skip("reference NR sources not compiled here (oracle/_ref)")
    y = np.sin(1.7 * x) + 0.3 * x
    xp = np.concatenate([np.linspace(x[0] - 0.2, x[-1] + 0.2, 57), x])
    ref = np.empty(len(xp))
    O.ref.chen.nrref_spline_chen(np.ascontiguousarray(x), np.ascontiguousarray(y),
                                   np.ascontiguousarray(xp), ref, len(x), len(xp))
    rc, out = resample(x, y, xp)
    assert rc == 0
    scale = np.abs(ref).max()
    assert np.abs(out - ref).max() <= 1e-11 * scale


@pytest.mark.parametrize("x", list(_cases()))
def test_spline_matches_scipy_not_a_knot(x):
    from scipy.interpolate import CubicSpline
    y = np.exp(-x) * np.cos(2.0 * x)
    xp = np.linspace(x[0] - 0.1, x[-1] + 0.1, 101)
    rc, out = resample(x, y, xp)
    assert rc == 0
    err = np.abs(out - CubicSpline(x, y, bc_type="not-a-knot")(xp))
    inside = (xp >= x[0]) & (xp <= x[-1])
    assert err[inside].max() <= 1e-12
    # outside the table the end cubics are extrapolated far beyond a short
    # end interval, which amplifies the rounding of either implementation
    assert err[~inside].max() <= 1e-9


def test_spline_reproduces_cubics_exactly():
    x = np.sort(np.random.default_rng(8).uniform(-1, 2, 25))
    p = lambda t: 0.5 - 1.25 * t + 0.75 * t ** 2 - 0.4 * t ** 3
    xp = np.linspace(-1.5, 2.5, 77)
    rc, out = resample(x, p(x), xp)
    assert rc == 0
    assert np.abs(out - p(xp)).max() <= 1e-11


def test_spline_rejects_short_or_unsorted_tables():
    assert resample([0.0, 1.0, 2.0], [0.0, 1.0, 4.0], [0.5])[0] != 0
    assert resample([0.0, 2.0, 1.0, 3.0], [0.0, 1.0, 4.0, 9.0], [0.5])[0] != 0
