"""CPU tests of the time-series evaluation (series.c through libadhk.so, host code only).

The average of a piecewise-linear series over [t1, t2] is pinned by closed forms (a linear ramp
averages to its midpoint value; a constant to itself) and by scipy's adaptive quadrature of numpy's
own linear interpolant over random intervals that span several table entries.
"""
import numpy as np
import pytest
from scipy.integrate import quad

import adh_kraken_amd as A
from adh_kraken_amd import _lib as L


def test_ramp_and_constant_closed_forms():
    t, v = [0.0, 1000.0], [50.0, 600.0]
    for t1, t2 in ((0, 5), (995, 1000), (123.25, 456.5)):
        assert A.series_average(t, v, t1, t2) == pytest.approx(50 + 0.55 * 0.5 * (t1 + t2), rel=1e-15)
    assert A.series_average([3.0], [7.5], -1e9, 1e9) == 7.5
    assert A.series_average(t, v, 500.0, 500.0) == pytest.approx(325.0, rel=1e-15)


def test_piecewise_average_matches_quadrature():
    rng = np.random.default_rng(3)
    t = np.cumsum(rng.uniform(0.5, 3.0, 12))
    v = rng.uniform(-2.0, 5.0, 12)
    for _ in range(50):
        t1, t2 = np.sort(rng.uniform(t[0], t[-1], 2))
        pts = [p for p in t if t1 < p < t2]
        ref = quad(lambda s: np.interp(s, t, v), t1, t2, points=pts or None, epsabs=1e-13, epsrel=1e-13,
                   limit=200)[0] / (t2 - t1)
        assert A.series_average(t, v, t1, t2) == pytest.approx(ref, rel=1e-11, abs=1e-12)


def test_out_of_range_and_bad_tables_fail():
    with pytest.raises(L.AdhkError):
        A.series_average([0.0, 10.0], [1.0, 2.0], 5.0, 10.5)
    with pytest.raises(L.AdhkError):
        A.series_average([0.0, 10.0], [1.0, 2.0], -0.5, 1.0)
    with pytest.raises(L.AdhkError):
        A.series_average([0.0, 0.0], [1.0, 2.0], 0.0, 0.0)
