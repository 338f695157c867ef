"""Host parts of the struct-level surface (include/adhk_ref.h), no GPU: the time loop's clock (tc_init,
update_dt), the series averages of advance_time, the friction laws behind fe_sw2_get_roughness, the
single-rank messg reductions, sivar_position_build_dof_map and smodel_super_prep_sol.

Pins: series averages against scipy quadrature of numpy's linear interpolant; the drag coefficient against
the rough-wall log-law profile integrated numerically; the roughness height against the Manning-Strickler
relation solved for n; the dof map against its closed form 3 node + var; the clock against hand-counted
steps of tc_init.c:4-33 and time_loop.c:108-120."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
from scipy.integrate import quad

from adh_kraken_amd import _lib as L
from test_oracle_pins import log_law_drag

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
dp, ip_t = C.POINTER(C.c_double), C.POINTER(C.c_int)
G = 9.8


@pytest.fixture(scope="module")
def lib():
    return C.CDLL(L.LIB_PATH)


@pytest.fixture(scope="module")
def hs(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("refh") / "libref_harness.so")
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run(["gcc", "-std=gnu99", "-O1", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "ref_harness.c"), "-o", out, "-L" + libdir, "-ladhk",
                    "-Wl,-rpath," + libdir], check=True)
    h = C.CDLL(out)
    h.hs_series_avg.argtypes = [C.c_int, dp, dp, dp, C.c_double, C.c_double, dp]
    h.hs_clock.argtypes = [C.c_double, C.c_double, C.c_double, dp, ip_t, dp]
    h.hs_roughness.argtypes = [C.c_double, C.c_double, C.c_int]
    h.hs_roughness.restype = C.c_double
    h.hs_dof_map.argtypes = [C.c_int, ip_t]
    h.hs_prep_sol.argtypes = [C.c_int, dp, dp, dp]
    h.hs_prep_sol.restype = None
    h.hs_time_loop_refusal.argtypes = [C.c_int]
    return h


def d(a):
    return a.ctypes.data_as(dp)


def quad_average(t, v, t1, t2):
    pts = [p for p in t if t1 < p < t2]
    return quad(lambda s: np.interp(s, t, v), t1, t2, points=pts or None, epsabs=1e-14, epsrel=1e-13)[0] / (t2 - t1)


@pytest.mark.parametrize("t1,t2", [(0.5, 0.9), (0.0, 6.0), (2.5, 17.0), (6.0, 20.0), (19.0, 20.0), (0.0, 20.0)])
def test_series_average_vs_quadrature(hs, t1, t2):
    t = np.array([0.0, 1.0, 6.0, 11.5, 20.0])
    v0 = np.array([10.0, 12.0, 25.0, 31.0, 40.0])
    v1 = np.array([-0.5, -0.25, -1.25, 0.75, 2.0])
    out = np.zeros(2)
    assert hs.hs_series_avg(len(t), d(t), d(v0), d(v1), t1, t2, d(out)) == 0  # the third series kept its value
    for k, v in enumerate((v0, v1)):
        want = quad_average(t, v, t1, t2)
        assert abs(out[k] - want) <= 1e-13 * max(1.0, abs(want)), (k, out[k], want)


def test_series_average_outside_the_range_is_an_error(hs):
    t, v = np.array([0.0, 3.0]), np.array([10.0, 12.0])
    out = np.zeros(2)
    assert hs.hs_series_avg(2, d(t), d(v), d(v), 2.0, 3.5, d(out)) == -1
    assert "not in the range" in L.last_error()


@pytest.mark.parametrize("t0,tf,dt,want_dt,want_steps", [
    (0.0, 10.0, 1.0, 1.0, 10),    # whole steps to t_final
    (0.0, 1.5, 1.0, 0.75, 2),     # two steps would overshoot: half the remaining time (tc_init.c:25-26)
    (0.0, 0.8, 1.0, 0.8, 1),      # one step overshoots: exactly the remaining time (tc_init.c:21-22)
    (5.0, 65.0, 2.0, 2.0, 30),
])
def test_clock(hs, t0, tf, dt, want_dt, want_steps):
    dto, tend, n = np.zeros(1), np.zeros(1), C.c_int()
    assert hs.hs_clock(t0, tf, dt, d(dto), C.byref(n), d(tend)) == 0
    assert dto[0] == want_dt and n.value == want_steps
    assert abs(tend[0] - tf) <= 1e-12 * max(1.0, tf)


@pytest.mark.parametrize("depth,ks", [(1.0, 0.05), (0.3, 0.002), (2.5, 0.4), (0.02, 0.3)])
def test_drag_coefficient_vs_log_law(lib, hs, depth, ks):
    lib.fr_bedshstr_drag_coef.restype = C.c_double
    lib.fr_bedshstr_drag_coef.argtypes = [C.c_double, C.c_double]
    cf = lib.fr_bedshstr_drag_coef(depth, ks)
    assert abs(cf - log_law_drag(depth, ks)) <= 1e-12 * cf
    assert hs.hs_roughness(depth, ks, 0) == cf  # fe_sw2_get_roughness: the string's roughness height
    assert hs.hs_roughness(depth, 0.0, 0) == 0.0  # no roughness height: no drag


def test_roughness_laws_off_the_path_are_refused(hs):
    assert np.isnan(hs.hs_roughness(1.0, 0.05, 1))
    assert "vegetation" in L.last_error()


def test_manning_to_roughness_height(lib):
    lib.fr_manningsn_to_rheight.restype = C.c_double
    lib.fr_manningsn_to_rheight.argtypes = [C.c_double, C.c_double, C.c_double]
    for ks in (1e-3, 0.05, 0.7):
        n = ks ** (1.0 / 6.0) / (8.25 * np.sqrt(G))  # Manning-Strickler: n = ks^(1/6) / (8.25 sqrt(g))
        assert abs(lib.fr_manningsn_to_rheight(n, 1.0, G) - ks) <= 1e-13 * ks
    assert lib.fr_manningsn_to_rheight(5e-9, 1.0, G) == 0.0


def test_messg_single_rank_identity(lib):
    for name, ty, x in [("messg_dmax", C.c_double, -2.5), ("messg_dmin", C.c_double, 3.25),
                        ("messg_dsum", C.c_double, 1e300), ("messg_imin", C.c_int, -7), ("messg_imax", C.c_int, 11)]:
        f = getattr(lib, name)
        f.restype, f.argtypes = ty, [ty]
        assert f(x) == x


def test_dof_map(hs):
    n = 37
    out = np.zeros(3 * n, np.int32)
    assert hs.hs_dof_map(n, out.ctypes.data_as(ip_t)) == 3 * n
    assert np.array_equal(out.reshape(3, n), 3 * np.arange(n)[None, :] + np.arange(3)[:, None])


def test_prep_sol_rotates_levels(hs):
    rng = np.random.default_rng(0)
    s, so, soo = rng.normal(size=(3, 12))
    a, b, c = s.copy(), so.copy(), soo.copy()
    hs.hs_prep_sol(12, d(a), d(b), d(c))
    assert np.array_equal(a, s) and np.array_equal(b, s) and np.array_equal(c, so)


def test_time_loop_refusals(hs):
    """time_loop returns 0 with a message, before any device work, for a dt series that does not cover
    t_prev (the reference exits in sseries_get_interval) and for a time stepper other than FE_NEWTON"""
    assert hs.hs_time_loop_refusal(0) == 0
    assert "t_prev is not in its range" in L.last_error()
    assert hs.hs_time_loop_refusal(1) == 0
    assert "FE_NEWTON" in L.last_error()
