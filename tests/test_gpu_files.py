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
