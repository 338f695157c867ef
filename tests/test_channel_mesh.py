"""CPU tests of the unstructured channel meshes (meshgen.c through libadhk.so, host code only).

Pins: the Poisson-disk set keeps every pair at least r apart (brute force over a k-d tree) and is
maximal up to the sampler's 30 tries (no empty disk of radius 2r inside the domain); the triangles
tile the rectangle (areas sum to its area, every interior edge shared by two triangles, boundary
edges only on its sides); the RCM order is a permutation that cuts the bandwidth of a random order
by far; the boundary segments have the domain on their left (outward normals -x at the inflow, +x at
the outflow) and the inflow segments carry exactly Q; the bed and roughness follow the section.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import adh_kraken_amd as A
from adh_kraken_amd import _lib as L
from adh_kraken_amd import channel as CH

SPEC = A.ChannelSpec(length=400.0, width=120.0, spacing=3.0, chan_width=40.0, bank_width=10.0, meander_amp=15.0,
                     meander_wavelength=200.0, nbumps=4, inflow_q=50.0, fp_depth_out=1.0, seed=11)


@pytest.fixture(scope="module")
def chan():
    return A.channel_mesh(SPEC)


def test_poisson_disk_spacing_and_maximality():
    r = 2.0
    x, y = CH.poisson_disk(0.0, 100.0, 0.0, 50.0, r, 5)
    pts = np.column_stack([x, y])
    d, _ = cKDTree(pts).query(pts, k=2)
    assert d[:, 1].min() >= r * (1 - 1e-12)
    assert x.min() >= 0 and x.max() <= 100 and y.min() >= 0 and y.max() <= 50
    gx, gy = np.meshgrid(np.linspace(0, 100, 401), np.linspace(0, 50, 201))
    dg, _ = cKDTree(pts).query(np.column_stack([gx.ravel(), gy.ravel()]))
    assert dg.max() < 2 * r
    # the same seed gives the same set; another seed another one
    x2, _ = CH.poisson_disk(0.0, 100.0, 0.0, 50.0, r, 5)
    x3, _ = CH.poisson_disk(0.0, 100.0, 0.0, 50.0, r, 6)
    assert np.array_equal(x, x2) and not (x.size == x3.size and np.array_equal(x, x3))


def test_triangles_tile_the_rectangle(chan):
    mesh, _ = chan
    x, y, t = mesh.x, mesh.y, mesh.tri
    area2 = (x[t[:, 1]] - x[t[:, 0]]) * (y[t[:, 2]] - y[t[:, 0]]) - (x[t[:, 2]] - x[t[:, 0]]) * (y[t[:, 1]] - y[t[:, 0]])
    assert area2.min() > 0  # counter-clockwise, none degenerate
    assert abs(0.5 * area2.sum() - SPEC.length * SPEC.width) <= 1e-9 * SPEC.length * SPEC.width
    be = CH.boundary_edges(t)
    on_side = ((np.abs(x[be]) < 1e-9) | (np.abs(x[be] - SPEC.length) < 1e-9)).all(1) | \
              ((np.abs(y[be]) < 1e-9) | (np.abs(y[be] - SPEC.width) < 1e-9)).all(1)
    assert on_side.all()
    perim = np.hypot(x[be[:, 0]] - x[be[:, 1]], y[be[:, 0]] - y[be[:, 1]]).sum()
    assert abs(perim - 2 * (SPEC.length + SPEC.width)) <= 1e-9 * perim


def test_rcm_is_a_permutation_with_small_bandwidth(chan):
    mesh, _ = chan
    rng = np.random.default_rng(0)
    rp = rng.permutation(mesh.nnodes).astype(np.int32)  # rp[old] = shuffled
    t_shuffled = rp[mesh.tri]
    perm = CH.rcm_order(mesh.nnodes, t_shuffled)  # perm[new] = shuffled
    assert np.array_equal(np.sort(perm), np.arange(mesh.nnodes))
    inv = np.empty_like(perm)
    inv[perm] = np.arange(mesh.nnodes, dtype=np.int32)
    bw = lambda tt: int(np.abs(tt[:, :, None] - tt[:, None, :]).max())
    b_rcm = bw(inv[t_shuffled])
    assert b_rcm * 10 < bw(t_shuffled)
    # nodes across the width: a breadth-first level holds about that many
    assert b_rcm <= 3.0 * SPEC.width / SPEC.spacing * 1.5
    # the mesh's own order sorts rows inside windows of 256 after RCM: within one window of it
    assert bw(mesh.tri) <= b_rcm + 2 * 256
    w = CH.degree_windows(mesh.nnodes, mesh.tri)
    assert np.array_equal(np.sort(w), np.arange(mesh.nnodes))
    assert all(set(w[k:k + 256]) == set(range(k, min(k + 256, mesh.nnodes))) for k in range(0, mesh.nnodes, 256))


def test_boundary_segments(chan):
    mesh, sol0 = chan
    x, y, s = mesh.x, mesh.y, mesh.seg
    dx, dy = x[s[:, 1]] - x[s[:, 0]], y[s[:, 1]] - y[s[:, 0]]
    ln = np.hypot(dx, dy)
    nx = dy / ln  # outward normal of a segment with the domain on its left (selem_1d.c:53-54)
    fl, dp = mesh.seg_kind == L.BC_FLUX, mesh.seg_kind == L.BC_DEPTH_NODAL
    assert np.allclose(nx[fl], -1.0) and np.allclose(nx[dp], 1.0)
    assert abs(-(mesh.seg_val[fl, 0] * ln[fl]).sum() - SPEC.inflow_q) <= 1e-9 * SPEC.inflow_q
    # the depth condition continues the initial flat surface at both nodes of every outflow segment
    assert np.allclose(mesh.seg_val[dp, 0], sol0[0::3][s[dp, 0]], rtol=0, atol=1e-12)
    assert np.allclose(mesh.seg_val[dp, 1], sol0[0::3][s[dp, 1]], rtol=0, atol=1e-12)
    # inflow only across the channel section
    cs = SPEC.c()
    lib = L.load()
    import ctypes as C
    w = [lib.adhk_channel_section_weight(C.byref(cs), 0.0, 0.5 * (y[a] + y[b])) for a, b in s[fl]]
    assert min(w) >= 0.5


def test_bed_and_roughness_follow_the_section(chan):
    mesh, sol0 = chan
    cs = SPEC.c()
    import ctypes as C
    lib = L.load()
    w = np.array([lib.adhk_channel_section_weight(C.byref(cs), a, b) for a, b in zip(mesh.x, mesh.y)])
    # bump-free bed: z = -slope x - depth w(x, y); the bumps only raise it, by at most bump_amp_max each
    base = -SPEC.slope * mesh.x - SPEC.chan_depth * w
    assert (mesh.z - base).min() >= -1e-12
    assert (mesh.z - base).max() <= SPEC.nbumps * SPEC.bump_amp_max
    t = mesh.tri
    wc = np.array([lib.adhk_channel_section_weight(C.byref(cs), mesh.x[e].mean(), mesh.y[e].mean()) for e in t])
    assert np.array_equal(mesh.mann, np.where(wc >= 0.5, SPEC.n_channel, SPEC.n_floodplain))
    assert np.allclose(sol0[0::3] + mesh.z, -SPEC.slope * SPEC.length + SPEC.fp_depth_out, atol=1e-12)
