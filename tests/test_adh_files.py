"""CPU tests of the AdH file readers on the reference's qa/flume model files (tests/golden/qa_flume,
copied unchanged: mesh, materials, parameters, initial state and design of its coupled SW2/DW run).

Pins: card counts and values read back as written; the triangles tile the flume's rectangle; every SEG
card is a boundary edge; the material and string maps count what the files hold; the SW2 loader refuses
the model because material 2 is DW (outside this build), and loads it once that material is SW2.
"""
import os
import shutil

import numpy as np
import pytest

from adh_kraken_amd import _lib as L
from adh_kraken_amd import adh_files as F
from adh_kraken_amd.channel import boundary_edges

QA = os.path.join(os.path.dirname(__file__), "golden", "qa_flume")


def test_geo_cards():
    x, y, z, tri, tfl, seg, sfl = F.read_geo(os.path.join(QA, "test.geo"))
    assert (len(x), len(tri), len(seg)) == (561, 1000, 120)
    assert (x[0], y[0], z[0]) == (0.0, 20.0, -1.0)
    assert tri.min() == 0 and tri.max() == 560 and np.all(tfl == -1) and np.all(sfl == 5)
    a = 0.5 * ((x[tri[:, 1]] - x[tri[:, 0]]) * (y[tri[:, 2]] - y[tri[:, 0]])
               - (x[tri[:, 2]] - x[tri[:, 0]]) * (y[tri[:, 1]] - y[tri[:, 0]]))
    assert abs(np.abs(a).sum() - (x.max() - x.min()) * (y.max() - y.min())) < 1e-9
    t = tri.copy()
    t[a < 0] = t[a < 0][:, [0, 2, 1]]
    be = {tuple(sorted(e)) for e in boundary_edges(t)}
    assert all(tuple(sorted(s)) in be for s in seg) and len(be) == len(seg)


def test_bad_geo_card_is_reported(tmp_path):
    p = tmp_path / "bad.geo"
    p.write_text("ND 1 0 0 0\nND 2 1 0 0\nND 3 0 1 0\nTRI 1 1 2 7 -1\n")
    with pytest.raises(L.AdhkError, match="line 4"):
        F.read_geo(str(p))
    # a repeated node id (3 ids up to 3, but node 2 never given) is refused at its line
    p.write_text("ND 1 0 0 0\nND 3 1 0 0\nND 3 0 1 0\nTRI 1 1 2 3 -1\n")
    with pytest.raises(L.AdhkError, match="line 3"):
        F.read_geo(str(p))


def test_control_files():
    tm, ss = F.read_physics(os.path.join(QA, "test_sm1.physics"), 1000, 120)
    assert np.bincount(tm).tolist() == [520, 480]
    assert np.bincount(ss).tolist() == [10, 50, 10, 50]
    prm = F.read_params(os.path.join(QA, "test_sm1.params"))
    assert prm.models == {1: "SW2", 2: "DW"}
    assert prm.ip == {"NTL": 1e-6, "ITL": 1e-6, "NIT": 10, "MIT": 100}
    t, v = prm.series[1]
    assert t.tolist() == [0, 20, 40, 60, 80, 100] and v[:, 0].tolist() == [0.1, 0.2, 0.3, 0.4, 0.5, 0.6]
    assert prm.bcs == [(3, "SW2", "NB", "VEL", [1])]
    init = F.read_init(os.path.join(QA, "test_sm1.init"), 561)
    assert np.all(init["DEPTH"] == 10.0) and init["DA_VELOCITY"].shape == (561, 2)
    assert np.all(init["DA_VELOCITY"] == [4.0, 2.0])
    d = F.read_design(os.path.join(QA, "test.design"))
    assert d["grid"] == "test.geo" and (d["t0"], d["tf"]) == (0.0, 100.0)
    assert d["mono"] == [(1, "test_sm1"), (2, "test_sm2")]


def test_sw2_loader_refuses_other_physics_and_loads_sw2(tmp_path):
    with pytest.raises(NotImplementedError, match="DW"):
        F.load_sw2_model(os.path.join(QA, "test.design"))
    for f in os.listdir(QA):
        shutil.copy(os.path.join(QA, f), tmp_path / f)
    prm = (tmp_path / "test_sm1.params").read_text().replace("MODEL 2 DW", "MODEL 2 SW2")
    (tmp_path / "test_sm1.params").write_text(prm)
    m = F.load_sw2_model(str(tmp_path / "test.design"))
    assert m.mesh.nnodes == 561 and m.mesh.nelems2d == 1000 and m.mesh.nelems1d == 10
    assert np.all(m.mesh.seg_kind == L.BC_VEL) and (m.dt, m.nsteps) == (100.0, 1)
    # the velocity segments lie on one straight side; the bound scales are the inward normal
    comp0 = [s for idx, sid, c, s in m.bindings if c == 0][0]
    comp1 = [s for idx, sid, c, s in m.bindings if c == 1][0]
    assert np.allclose(np.hypot(comp0, comp1), 1.0)
    assert m.solver == {"tol_nonlin": 1e-6, "max_nonlin_it": 10, "lin_max_it": 100}
    assert m.sol0[0::3].tolist() == [10.0] * 561
