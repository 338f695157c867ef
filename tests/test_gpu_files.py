"""GPU parity on a model read from AdH files: the reference's qa/flume mesh, materials, boundary strings,
series and initial state (tests/golden/qa_flume), with its DW material switched to SW2 (this build runs
the SW2 step only).  The inflow velocity boundary takes the series' average over the first step along
the inward normal; residual 1e-12 and FD Jacobian 1e-7 against the oracle, pattern bit-exact.
"""
import os
import shutil

import numpy as np
import pytest

import adh_kraken_amd as A
import oracle as O
from adh_kraken_amd import adh_files as F
from helpers import rel_err

pytestmark = pytest.mark.gpu

QA = os.path.join(os.path.dirname(__file__), "golden", "qa_flume")


def test_qa_flume_files_residual_and_jacobian(gpu, tmp_path):
    for f in os.listdir(QA):
        shutil.copy(os.path.join(QA, f), tmp_path / f)
    p = tmp_path / "test_sm1.params"
    p.write_text(p.read_text().replace("MODEL 2 DW", "MODEL 2 SW2"))
    m = F.load_sw2_model(str(tmp_path / "test.design"))
    mesh = m.mesh
    t, v = m.series[1]
    avg = A.series_average(t, v[:, 0], 0.0, m.dt)
    for idx, sid, comp, scale in m.bindings:
        mesh.seg_val[idx, comp] = scale * avg
    sm = A.SModelSuper(mesh, dt=m.dt)
    orc = O.Oracle(mesh, dt=m.dt, threads=8)
    rng = np.random.default_rng(17)
    s = m.sol0.copy()
    s[0::3] += rng.uniform(-0.5, 0.5, mesh.nnodes)
    sm.sol, sm.sol_old, sm.sol_older = s, m.sol0, m.sol0
    orc.set_state(s, m.sol0, m.sol0)
    ip_g, c_g = A.sparsity(sm)
    ip_o, c_o = orc.pattern()
    assert np.array_equal(ip_g, ip_o) and np.array_equal(c_g, c_o)
    assert rel_err(A.assemble_residual(sm), orc.residual()) <= 1e-12
    assert rel_err(A.assemble_jacobian(sm), orc.jacobian()) <= 1e-7


def load_qa(tmp_path):
    for f in os.listdir(QA):
        shutil.copy(os.path.join(QA, f), tmp_path / f)
    p = tmp_path / "test_sm1.params"
    p.write_text(p.read_text().replace("MODEL 2 DW", "MODEL 2 SW2"))
    return F.load_sw2_model(str(tmp_path / "test.design"))


def test_qa_flume_time_steps_from_its_files(gpu, tmp_path):
    """The model of the files -- mesh, materials, the inflow velocity string driven by its series, the .init
    state and the Newton tolerance of the .params (1e-6) -- time-stepped by the library's own loop, which
    averages the series over each step; the oracle gets the same averages by hand.  The design's time step
    of 100 s fails Newton in both (next test), so the steps here are 1 s, 10 of them."""
    m = load_qa(tmp_path)
    dt, nsteps = 1.0, 10
    sm = A.SModelSuper(m.mesh, dt=dt)
    sm.set_initial(m.sol0)
    F.apply_bindings(sm, m)
    sm.tol_nonlin = m.solver["tol_nonlin"]
    ng, sg = A.time_loop(sm, nsteps)
    assert ng == nsteps
    mesh = m.mesh
    orc = O.Oracle(mesh, dt=dt, threads=8)
    orc.set_initial(m.sol0)
    sv = O.default_solver(tol_nonlin=m.solver["tol_nonlin"])
    val = mesh.seg_val.copy()
    its = 0
    for k in range(nsteps):
        for idx, sid, comp, scale in m.bindings:
            t, v = m.series[sid]
            val[idx, comp] = scale * A.series_average(t, v[:, 0], k * dt, (k + 1) * dt)
        orc.set_seg_values(val)
        n, st = orc.time_loop(1, sv)
        assert n == 1
        its += st["newton_its"]
    assert sg["newton_its"] == its
    gs, os_ = sm.sol, orc.state()[0]
    assert rel_err(gs[0::3], os_[0::3]) <= 1e-8
    assert rel_err(gs[1::3], os_[1::3]) <= 1e-6


def test_qa_flume_design_time_step_fails_as_in_the_oracle(gpu, tmp_path):
    """At the design's own 100 s step Newton does not converge (block-Jacobi BiCGStab, 10 iterations); the
    step is not taken, the state stays at the old level and dt is cut to 25 s, in the oracle and here"""
    m = load_qa(tmp_path)
    sm = A.SModelSuper(m.mesh, dt=m.dt)
    sm.set_initial(m.sol0)
    F.apply_bindings(sm, m)
    sm.tol_nonlin = m.solver["tol_nonlin"]
    n, st = A.time_loop(sm, m.nsteps)
    assert m.dt == 100.0 and m.nsteps == 1
    assert n <= 0 and st["converged"] == 0 and sm.get_dt() == 25.0
    assert np.array_equal(sm.sol, m.sol0)
    mesh = m.mesh
    for idx, sid, comp, scale in m.bindings:
        t, v = m.series[sid]
        mesh.seg_val[idx, comp] = scale * A.series_average(t, v[:, 0], 0.0, m.dt)
    orc = O.Oracle(mesh, dt=m.dt, threads=8)
    orc.set_initial(m.sol0)
    no, so = orc.time_loop(1, O.default_solver(tol_nonlin=m.solver["tol_nonlin"]))
    assert no <= 0 and so["converged"] == 0
