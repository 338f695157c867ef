"""Output for inspection and restart: XDMF (XML + raw binary data items) and AdH .init files.

The reference writes XDMF with HDF5 heavy data (src/xdmf/); HDF5 is not in this build, so the heavy
data here are raw little-endian binaries that XDMF's "Binary" data items point to (ParaView and VisIt
read them).  One temporal collection holds the mesh once and depth, velocity and water surface
elevation per written step.  write_init / read_init (adh_files) round-trip the state in the .init
layout the reference's restart reader takes (smodel_super_read_init.c:27-169): DEPTH, OLD_DEPTH,
OLDER_DEPTH, DA_VELOCITY, OLD_DA_VELOCITY, OLDER_DA_VELOCITY, each closed by ENDDS.
"""
import os
from xml.sax.saxutils import quoteattr

import numpy as np


def _item(fname, shape, ntype, prec):
    dims = " ".join(str(d) for d in shape)
    return (f'<DataItem Dimensions="{dims}" NumberType="{ntype}" Precision="{prec}" Format="Binary" '
            f'Endian="Little">{fname}</DataItem>')


def write_xdmf(prefix, mesh, steps):
    """steps: [(time, sol [3N])].  Writes prefix.xmf and prefix_*.bin next to it; returns the .xmf path."""
    d, base = os.path.split(prefix)
    d = d or "."
    N, E = mesh.nnodes, mesh.nelems2d
    geo = np.column_stack([mesh.x, mesh.y, mesh.z]).astype("<f8")
    geo.tofile(os.path.join(d, base + "_geometry.bin"))
    np.ascontiguousarray(mesh.tri, "<i4").tofile(os.path.join(d, base + "_topology.bin"))
    grids = []
    for k, (t, sol) in enumerate(steps):
        sol = np.asarray(sol, float)
        h, u, v = sol[0::3], sol[1::3], sol[2::3]
        np.asarray(h, "<f8").tofile(os.path.join(d, f"{base}_depth_{k}.bin"))
        np.column_stack([u, v, np.zeros(N)]).astype("<f8").tofile(os.path.join(d, f"{base}_velocity_{k}.bin"))
        np.asarray(h + mesh.z, "<f8").tofile(os.path.join(d, f"{base}_wse_{k}.bin"))
        grids.append(
            f'<Grid Name="step {k}" GridType="Uniform"><Time Value="{float(t)!r}"/>'
            f'<Topology TopologyType="Triangle" NumberOfElements="{E}">'
            f'{_item(base + "_topology.bin", (E, 3), "Int", 4)}</Topology>'
            f'<Geometry GeometryType="XYZ">{_item(base + "_geometry.bin", (N, 3), "Float", 8)}</Geometry>'
            f'<Attribute Name="Depth" AttributeType="Scalar" Center="Node">'
            f'{_item(f"{base}_depth_{k}.bin", (N,), "Float", 8)}</Attribute>'
            f'<Attribute Name="Velocity" AttributeType="Vector" Center="Node">'
            f'{_item(f"{base}_velocity_{k}.bin", (N, 3), "Float", 8)}</Attribute>'
            f'<Attribute Name="Water surface elevation" AttributeType="Scalar" Center="Node">'
            f'{_item(f"{base}_wse_{k}.bin", (N,), "Float", 8)}</Attribute></Grid>')
    xml = ('<?xml version="1.0" ?>\n<Xdmf Version="3.0"><Domain>'
           f'<Grid Name={quoteattr(base)} GridType="Collection" CollectionType="Temporal">'
           + "".join(grids) + "</Grid></Domain></Xdmf>\n")
    path = os.path.join(d, base + ".xmf")
    with open(path, "w") as f:
        f.write(xml)
    return path


def write_init(path, sol, sol_old=None, sol_older=None):
    """AdH .init file of the state (depth and depth-averaged velocity at the three levels)."""
    sol = np.asarray(sol, float)
    levels = [("", sol), ("OLD_", sol if sol_old is None else np.asarray(sol_old, float)),
              ("OLDER_", sol if sol_older is None else np.asarray(sol_older, float))]
    with open(path, "w") as f:
        for pre, s in levels:
            f.write(f"{pre}DEPTH\n")
            f.writelines(f"{float(h)!r}\n" for h in s[0::3])
            f.write("ENDDS\n")
        for pre, s in levels:
            f.write(f"{pre}DA_VELOCITY\n")
            f.writelines(f"{float(u)!r} {float(v)!r}\n" for u, v in zip(s[1::3], s[2::3]))
            f.write("ENDDS\n")
