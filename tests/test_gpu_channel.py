"""GPU parity on unstructured channel meshes (Poisson-disk points, Delaunay, RCM order) against the oracle.

Same bars as the flume: pattern bit-exact, residual 1e-12, FD Jacobian 1e-7, depth after N steps 1e-8
(velocity 1e-6), mixed precision depth 1e-6 against FP64.  Unstructured meshes add what the flume lacks:
rows of 4 to 10 blocks (uneven SELL-64 slices), two roughness classes, a meandering bed and a depth
condition that varies along the outflow.
"""
import numpy as np
import pytest

import adh_kraken_amd as A
import oracle as O
from adh_kraken_amd import _lib as L
from helpers import perturbed_state, rel_err

pytestmark = pytest.mark.gpu

SPEC = A.ChannelSpec(length=500.0, width=150.0, spacing=3.0, chan_width=40.0, bank_width=10.0, meander_amp=20.0,
                     meander_wavelength=250.0, nbumps=5, inflow_q=60.0, fp_depth_out=1.0, seed=3)


@pytest.fixture(scope="module")
def chan():
    return A.channel_mesh(SPEC)


def pair(mesh, s, so, dt=5.0):
    sm = A.SModelSuper(mesh, dt=dt)
    orc = O.Oracle(mesh, dt=dt, threads=8)
    sm.sol, sm.sol_old, sm.sol_older = s, so, so
    orc.set_state(s, so, so)
    return sm, orc


def test_channel_pattern_residual_jacobian(gpu, chan):
    mesh, sol0 = chan
    s, so = perturbed_state(sol0, seed=21, amp_h=0.05, amp_v=0.2)
    sm, orc = pair(mesh, s, so)
    ip_g, c_g = A.sparsity(sm)
    ip_o, c_o = orc.pattern()
    assert np.array_equal(ip_g, ip_o) and np.array_equal(c_g, c_o)
    rows = np.diff(ip_g) // 3
    assert rows.min() <= 5 and rows.max() >= 9  # uneven rows
    assert rel_err(A.assemble_residual(sm), orc.residual()) <= 1e-12
    assert rel_err(A.assemble_jacobian(sm), orc.jacobian()) <= 1e-7


def test_channel_time_loop_depth(gpu, chan):
    mesh, sol0 = chan
    sm, orc = pair(mesh, sol0, sol0)
    ng, sg = A.time_loop(sm, 6)
    no, so = orc.time_loop(6)
    assert ng == no == 6
    assert sg["newton_its"] == so["newton_its"]
    gs, os_ = sm.sol, orc.state()[0]
    assert rel_err(gs[0::3], os_[0::3]) <= 1e-8
    assert rel_err(gs[1::3], os_[1::3]) <= 1e-6 and rel_err(gs[2::3], os_[2::3]) <= 1e-6


def test_channel_mixed_precision(gpu, chan):
    mesh, sol0 = chan
    a = A.SModelSuper(mesh, dt=5.0)
    b = A.SModelSuper(mesh, dt=5.0)
    a.set_initial(sol0)
    b.set_initial(sol0)
    na, _ = A.time_loop(a, 6, precision=L.PREC_FP64)
    nb, stb = A.time_loop(b, 6, precision=L.PREC_MIXED)
    assert na == nb == 6
    assert rel_err(b.sol[0::3], a.sol[0::3]) <= 1e-6
