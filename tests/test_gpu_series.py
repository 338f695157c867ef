"""GPU: a rising inflow hydrograph driven by a time series, against the oracle stepped by hand.

The library averages the series over each step and sets the inflow segments (the time loop's own
forcing); the oracle gets the same boundary values from scipy quadrature of numpy's linear interpolant,
set before each of its single steps.  Depth after the run 1e-8 relative, same Newton counts.
"""
import numpy as np
import pytest
from scipy.integrate import quad

import adh_kraken_amd as A
import oracle as O
from adh_kraken_amd import _lib as L
from helpers import rel_err

pytestmark = pytest.mark.gpu


def test_hydrograph_inflow_matches_oracle(gpu):
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    width = A.FLUME_C1.ymax - A.FLUME_C1.ymin
    t_tab, q_tab = np.array([0.0, 6.0, 20.0]), np.array([10.0, 25.0, 40.0])
    inflow = np.flatnonzero(mesh.seg_kind == L.BC_FLUX)
    dt, nsteps = 1.0, 15
    sm = A.SModelSuper(mesh, dt=dt)
    sm.set_initial(sol0)
    sm.set_series(0, t_tab, q_tab)
    sm.bind_series(inflow, 0, -1.0 / width)
    ng, sg = A.time_loop(sm, nsteps)
    assert ng == nsteps and sm.time == pytest.approx(nsteps * dt)

    orc = O.Oracle(mesh, dt=dt, threads=8)
    orc.set_initial(sol0)
    val = mesh.seg_val.copy()
    its = 0
    for k in range(nsteps):
        t1, t2 = k * dt, (k + 1) * dt
        pts = [p for p in t_tab if t1 < p < t2]
        qavg = quad(lambda s: np.interp(s, t_tab, q_tab), t1, t2, points=pts or None, epsabs=1e-14)[0] / dt
        val[inflow, 0] = -qavg / width
        orc.set_seg_values(val)
        n, st = orc.time_loop(1)
        assert n == 1
        its += st["newton_its"]
    assert sg["newton_its"] == its
    gs, os_ = sm.sol, orc.state()[0]
    assert rel_err(gs[0::3], os_[0::3]) <= 1e-8
    assert rel_err(gs[1::3], os_[1::3]) <= 1e-6


def test_step_outside_the_series_fails_cleanly(gpu):
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    sm = A.SModelSuper(mesh, dt=1.0)
    sm.set_initial(sol0)
    sm.set_series(0, [0.0, 3.0], [10.0, 12.0])
    sm.bind_series(np.flatnonzero(mesh.seg_kind == L.BC_FLUX), 0, -1.0 / 20.0)
    n, _ = A.time_loop(sm, 5)
    assert n == -3  # three steps fit the series; the fourth would need it beyond t = 3 and is not taken
    assert "outside" in L.last_error()
    assert sm.time == pytest.approx(3.0)
