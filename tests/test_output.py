"""CPU tests of the inspection and restart output: the XDMF file parses, names the right sizes, and its
binary data items read back bit for bit; an .init file written from a state reads back bit for bit
through the reader that loads the reference's own .init files."""
import os
import xml.etree.ElementTree as ET

import numpy as np

import adh_kraken_amd as A
from adh_kraken_amd import adh_files as F
from adh_kraken_amd import xdmf as X


def test_xdmf_round_trip(tmp_path):
    mesh, sol0 = A.flume_mesh(A.FlumeSpec(npx=11, npy=5, xmax=10.0, ymax=4.0, nbumps=2, ic_flat_wse=1))
    rng = np.random.default_rng(1)
    steps = [(0.0, sol0), (2.5, sol0 + rng.uniform(-0.1, 0.1, sol0.size))]
    path = X.write_xdmf(str(tmp_path / "run"), mesh, steps)
    root = ET.parse(path).getroot()
    grids = root.findall("./Domain/Grid/Grid")
    assert [float(g.find("Time").get("Value")) for g in grids] == [0.0, 2.5]
    for (t, sol), g in zip(steps, grids):
        topo = g.find("Topology")
        assert int(topo.get("NumberOfElements")) == mesh.nelems2d
        tri = np.fromfile(tmp_path / topo.find("DataItem").text, "<i4").reshape(-1, 3)
        assert np.array_equal(tri, mesh.tri)
        geo = np.fromfile(tmp_path / g.find("Geometry/DataItem").text, "<f8").reshape(-1, 3)
        assert np.array_equal(geo[:, 2], mesh.z)
        att = {a.get("Name"): np.fromfile(tmp_path / a.find("DataItem").text, "<f8") for a in g.findall("Attribute")}
        assert np.array_equal(att["Depth"], sol[0::3])
        assert np.array_equal(att["Velocity"].reshape(-1, 3)[:, :2], np.column_stack([sol[1::3], sol[2::3]]))
        assert np.array_equal(att["Water surface elevation"], sol[0::3] + mesh.z)


def test_init_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    s, so, soo = (rng.normal(size=3 * 37) for _ in range(3))
    p = tmp_path / "state.init"
    X.write_init(str(p), s, so, soo)
    d = F.read_init(str(p), 37)
    assert np.array_equal(d["DEPTH"], s[0::3]) and np.array_equal(d["OLD_DEPTH"], so[0::3])
    assert np.array_equal(d["OLDER_DEPTH"], soo[0::3])
    assert np.array_equal(d["DA_VELOCITY"], np.column_stack([s[1::3], s[2::3]]))
    assert np.array_equal(d["OLDER_DA_VELOCITY"], np.column_stack([soo[1::3], soo[2::3]]))
