"""Unstructured channel-with-floodplain meshes: Poisson-disk points, Delaunay triangles, RCM node order
(with nodes sorted by row length inside windows of 256, see degree_windows).

The points, the node order, the bed and the roughness come from libadhk.so (meshgen.c); the
triangulation is scipy's Delaunay (Qhull), the one step done outside the library.  The domain is the
rectangle [0, length] x [0, width]; a main channel (straight, or meandering about the sine centreline
y = width/2 + amp sin(2 pi (x - phase) / wavelength)) is cut chan_depth below the floodplain with linear banks.
Boundary conditions, restating what the flume uses (fe_sw2_1d_explicit_flow,
fe_sw2_boundary_integrals.c:56-94, and fe_sw2_bc_h, sw2_bc_h.c:57-178):
  * inflow: the x = 0 edges whose midpoint lies in the channel carry the flux q = -Q / (their total
    length) per unit width (the floodplain edges at x = 0 are walls);
  * outflow: every x = length edge carries the depth condition per node (ADHK_BC_DEPTH_NODAL): the depth
    of a flat water surface (floodplain depth fp_depth_out over the bump-free floodplain) over the bed at
    each of its nodes, so that a lake at rest stays at rest where the bed varies along the outflow (the
    channel banks);
  * the long sides are walls (no boundary element: natural no-flux).
The initial state is that flat water surface with u = v = 0.
"""
import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib as L
from .sw2 import Mesh


@dataclass
class ChannelSpec:
    length: float = 1000.0
    width: float = 200.0
    spacing: float = 3.0
    chan_width: float = 40.0
    chan_depth: float = 3.0
    bank_width: float = 10.0
    meander_amp: float = 0.0
    meander_wavelength: float = 0.0
    slope: float = 1e-4
    nbumps: int = 0
    bump_amp_max: float = 0.3
    bump_sigma_min: float = 5.0
    bump_sigma_max: float = 30.0
    seed: int = 1
    n_channel: float = 0.025
    n_floodplain: float = 0.06
    inflow_q: float = 100.0
    fp_depth_out: float = 1.5
    meander_phase: float = 0.0

    def c(self):
        s = L.ChannelSpec()
        for f in fields(self):
            setattr(s, f.name, getattr(self, f.name))
        return s


# 10 km x 2 km channel with floodplain at 3 m minimum point spacing (1,371,497 points, 2,734,994 triangles:
# Poisson-disk sampling at 3 m packs about 0.069 points per m^2, so ~4M triangles would need ~2.5 m): a 400 m main
# channel 3 m below the floodplain, 50 seeded bumps (amplitude up to 0.3 m, sigma 5-30 m), Manning n
# 0.025 / 0.06, inflow 800 m^3/s; run at dt 5 s
CHANNEL_C3 = ChannelSpec(length=10_000.0, width=2_000.0, spacing=3.0, chan_width=400.0, chan_depth=3.0,
                         bank_width=20.0, slope=1e-4, nbumps=50, bump_amp_max=0.3, bump_sigma_min=5.0,
                         bump_sigma_max=30.0, seed=1, n_channel=0.025, n_floodplain=0.06, inflow_q=800.0,
                         fp_depth_out=1.5)


# 8 km meandering channel (sine centreline, amplitude 300 m, wavelength 2 km, a quarter wavelength of phase so
# that it meets the inflow and outflow sides at right angles) in a 1.8 km wide floodplain at 3 m minimum
# point spacing (about 1M points, 2M triangles): a 200 m channel 3 m deep, 50 seeded bumps, Manning n
# 0.025 / 0.06; the inflow follows the hydrograph C5_HYDROGRAPH (50 -> 600 m^3/s over 200 steps of 5 s)
CHANNEL_C5 = ChannelSpec(length=8_000.0, width=1_800.0, spacing=3.0, chan_width=200.0, chan_depth=3.0,
                         bank_width=20.0, meander_amp=300.0, meander_wavelength=2_000.0, meander_phase=500.0,
                         slope=1e-4, nbumps=50, bump_amp_max=0.3, bump_sigma_min=5.0, bump_sigma_max=30.0, seed=5,
                         n_channel=0.025, n_floodplain=0.06, inflow_q=50.0, fp_depth_out=1.5)
C5_HYDROGRAPH = (np.array([0.0, 1000.0]), np.array([50.0, 600.0]))


def inflow_segments(mesh):
    """Indices of the inflow (flux) segments and their total length."""
    idx = np.flatnonzero(mesh.seg_kind == L.BC_FLUX)
    s = mesh.seg[idx]
    return idx, float(np.hypot(mesh.x[s[:, 0]] - mesh.x[s[:, 1]], mesh.y[s[:, 0]] - mesh.y[s[:, 1]]).sum())


def bind_hydrograph(sm, t, q, sid=0):
    """Drive the inflow segments of sm by the discharge series (t, q) in m^3/s (flux -Q / inflow width)."""
    idx, w = inflow_segments(sm.mesh)
    sm.set_series(sid, t, q)
    sm.bind_series(idx, sid, -1.0 / w)


def poisson_disk(x0, x1, y0, y1, r, seed):
    lib = L.load()
    nmax = lib.adhk_poisson_disk_bound(x0, x1, y0, y1, r)
    if nmax < 0:
        raise ValueError("bad Poisson-disk domain")
    x, y = np.zeros(nmax), np.zeros(nmax)
    n = lib.adhk_poisson_disk(x0, x1, y0, y1, r, seed, nmax, L.dptr(x), L.dptr(y))
    if n < 0:
        raise L.AdhkError("adhk_poisson_disk failed")
    return x[:n].copy(), y[:n].copy()


def delaunay(x, y):
    """Counter-clockwise Delaunay triangles of the points (scipy / Qhull), degenerate ones dropped."""
    from scipy.spatial import Delaunay
    tri = Delaunay(np.column_stack([x, y])).simplices.astype(np.int32)
    a, b, c = tri[:, 0], tri[:, 1], tri[:, 2]
    area2 = (x[b] - x[a]) * (y[c] - y[a]) - (x[c] - x[a]) * (y[b] - y[a])
    cw = area2 < 0
    tri[cw] = tri[cw][:, [0, 2, 1]]
    scale = (x.max() - x.min()) * (y.max() - y.min())
    keep = np.abs(area2) > 1e-14 * scale
    return np.ascontiguousarray(tri[keep])


def rcm_order(nnodes, tri):
    lib = L.load()
    perm = np.zeros(nnodes, np.int32)
    t = np.ascontiguousarray(tri, np.int32)
    L.check(lib.adhk_rcm_order(nnodes, t.shape[0], L.iptr(t), L.iptr(perm)), "adhk_rcm_order")
    return perm


def degree_windows(nnodes, tri, window=256):
    """Stable sort of the nodes by row length (neighbours + 1) inside consecutive windows of `window` nodes:
    perm[new] = old.  Rows of a 64-node SELL slice are padded to the slice's longest row; RCM order mixes
    long and short rows (about 29% padding on these meshes), the sorted windows keep it near 7% while
    every node stays within its window, so neighbour gathers keep their locality."""
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]]).astype(np.int64)
    u = np.unique(np.minimum(e[:, 0], e[:, 1]) * nnodes + np.maximum(e[:, 0], e[:, 1]))
    deg = np.bincount(u // nnodes, minlength=nnodes) + np.bincount(u % nnodes, minlength=nnodes) + 1
    return np.concatenate([w0 + np.argsort(deg[w0:w0 + window], kind="stable")
                           for w0 in range(0, nnodes, window)]).astype(np.int32)


def boundary_edges(tri):
    """Edges used by one triangle only, oriented as in that (counter-clockwise) triangle: domain on the left."""
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    key = np.minimum(e[:, 0], e[:, 1]).astype(np.int64) * (int(tri.max()) + 1) + np.maximum(e[:, 0], e[:, 1])
    _, inv, cnt = np.unique(key, return_inverse=True, return_counts=True)
    return e[cnt[inv] == 1]


def channel_mesh(spec: ChannelSpec):
    """Mesh and initial state of the channel workload.  Returns (Mesh, sol0)."""
    lib = L.load()
    cs = spec.c()
    x, y = poisson_disk(0.0, spec.length, 0.0, spec.width, spec.spacing, spec.seed)
    tri = delaunay(x, y)
    perm = rcm_order(x.size, tri)
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    perm = perm[degree_windows(perm.size, inv[tri])]  # RCM, then rows of like length in each SELL slice
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    x, y = np.ascontiguousarray(x[perm]), np.ascontiguousarray(y[perm])
    tri = inv[tri]
    tri = np.ascontiguousarray(tri[np.lexsort((tri.max(1), tri.min(1)))], np.int32)  # elements in node order
    N, E = x.size, tri.shape[0]
    z = np.zeros(N)
    L.check(lib.adhk_channel_bed(C.byref(cs), N, L.dptr(x), L.dptr(y), L.dptr(z)), "adhk_channel_bed")
    mann = np.zeros(E)
    L.check(lib.adhk_channel_roughness(C.byref(cs), N, L.dptr(x), L.dptr(y), E, L.iptr(tri), L.dptr(mann)),
            "adhk_channel_roughness")
    be = boundary_edges(tri)
    xm, ym = 0.5 * (x[be[:, 0]] + x[be[:, 1]]), 0.5 * (y[be[:, 0]] + y[be[:, 1]])
    tol = 1e-9 * spec.length
    on_in = (np.abs(x[be[:, 0]]) < tol) & (np.abs(x[be[:, 1]]) < tol)
    on_out = (np.abs(x[be[:, 0]] - spec.length) < tol) & (np.abs(x[be[:, 1]] - spec.length) < tol)
    in_chan = np.array([lib.adhk_channel_section_weight(C.byref(cs), a, b) >= 0.5 for a, b in zip(xm, ym)], bool)
    inflow = be[on_in & in_chan]
    inflow = inflow[np.argsort(ym[on_in & in_chan])]
    outflow = be[on_out]
    outflow = outflow[np.argsort(ym[on_out])]
    if inflow.shape[0] == 0 or outflow.shape[0] == 0:
        raise ValueError("channel mesh has no inflow or no outflow boundary")
    w_in = np.hypot(x[inflow[:, 0]] - x[inflow[:, 1]], y[inflow[:, 0]] - y[inflow[:, 1]]).sum()
    wse = -spec.slope * spec.length + spec.fp_depth_out
    seg = np.ascontiguousarray(np.concatenate([inflow, outflow]), np.int32)
    kind = np.concatenate([np.full(inflow.shape[0], L.BC_FLUX),
                           np.full(outflow.shape[0], L.BC_DEPTH_NODAL)]).astype(np.int32)
    val = np.zeros((seg.shape[0], 2))
    val[:inflow.shape[0], 0] = -spec.inflow_q / w_in
    val[inflow.shape[0]:, 0] = wse - z[outflow[:, 0]]
    val[inflow.shape[0]:, 1] = wse - z[outflow[:, 1]]
    h0 = wse - z
    if h0.min() <= 0.0:
        raise ValueError(f"initial water surface leaves dry nodes (min depth {h0.min():.3g} m)")
    sol = np.zeros(3 * N)
    sol[0::3] = h0
    return Mesh(x, y, z, tri, seg, kind, val, mann), sol
