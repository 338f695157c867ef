"""The reference's struct-level surface (include/adhk_ref.h) on the GPU against the oracle.

tests/ref_harness.c builds SMODEL_SUPER / SGRID / SLIN_SYS for a flume as a reference caller does (physics
materials carrying fe_sw2_body_resid, fe_sw2_bc_h and fe_sw2_bc_flux; the dof map; boundary values in the
series) and calls libadhk.so by the reference's function names: slin_sys_init_sparsity_mono,
assemble_residual, assemble_jacobian, the elemental residual pointers, load_elem_mat, check_diag,
apply_Dirichlet_BC, get_residual_norms, increment_function, initialize_system and fe_newton in a time loop,
and the reference's own time loop (time_loop over an SMODEL_DESIGN) under an inflow hydrograph."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import adh_kraken_amd as A
import oracle as O
from adh_kraken_amd import _lib as L
from helpers import perturbed_state, rel_err, small_flume

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
dp, ip_t = C.POINTER(C.c_double), C.POINTER(C.c_int)


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    L.load()  # libadhk.so in this process, the same copy the harness links
    out = str(tmp_path_factory.mktemp("refh") / "libref_harness.so")
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run(["gcc", "-std=gnu99", "-O1", "-shared", "-fPIC", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "ref_harness.c"), "-o", out, "-L" + libdir, "-ladhk",
                    "-Wl,-rpath," + libdir], check=True)
    h = C.CDLL(out)
    h.h_create.restype = C.c_void_p
    h.h_create.argtypes = [C.c_int, dp, dp, dp, C.c_int, ip_t, C.c_int, ip_t, ip_t, C.c_double, C.c_double,
                           C.c_double, C.c_double]
    for name, args in [("h_set_manning", [C.c_void_p, dp]), ("h_sparsity", [C.c_void_p]),
                       ("h_init_system", [C.c_void_p]), ("h_residual", [C.c_void_p]),
                       ("h_jacobian", [C.c_void_p]), ("h_check_diag_dirichlet", [C.c_void_p]),
                       ("h_step", [C.c_void_p, ip_t]),
                       ("h_elem_resid", [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, dp]),
                       ("h_elem_mat", [C.c_void_p, C.c_int, dp]), ("h_norms", [C.c_void_p, dp, dp, dp, ip_t]),
                       ("h_increment", [C.c_void_p, dp]),
                       ("h_design", [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp]),
                       ("h_time_loop", [C.c_void_p, dp, dp, ip_t])]:
        getattr(h, name).restype = C.c_int
        getattr(h, name).argtypes = args
    for name, args in [("h_set_state", [C.c_void_p, dp, dp, dp]), ("h_set_bc", [C.c_void_p, ip_t, dp]),
                       ("h_get", [C.c_void_p, dp, dp, ip_t, dp, ip_t]), ("h_pattern", [C.c_void_p, ip_t, ip_t, dp]),
                       ("h_destroy", [C.c_void_p])]:
        getattr(h, name).restype = None
        getattr(h, name).argtypes = args
    return h


def d(a):
    return a.ctypes.data_as(dp)


def i(a):
    return a.ctypes.data_as(ip_t)


class Model:
    """a flume as reference structures in the harness"""

    def __init__(self, h, mesh, dt, manning=False):
        self.h, self.mesh = h, mesh
        self.keep = [np.ascontiguousarray(a) for a in (mesh.x, mesh.y, mesh.z)]
        tri = np.ascontiguousarray(mesh.tri, np.int32)
        seg = np.ascontiguousarray(mesh.seg, np.int32)
        kind = np.ascontiguousarray(mesh.seg_kind, np.int32)
        depth = mesh.seg_val[kind == O.BC_DEPTH, 0]
        flux = mesh.seg_val[kind == O.BC_FLUX, 0]
        assert np.all(depth == depth[0]) and np.all(flux == flux[0])  # one value per kind, as the series give
        self.keep += [tri, seg, kind]
        self.p = h.h_create(mesh.nnodes, d(self.keep[0]), d(self.keep[1]), d(self.keep[2]), mesh.nelems2d, i(tri),
                            mesh.nelems1d, i(seg), i(kind), float(depth[0]), float(flux[0]), dt, 9.8)
        if manning:
            self.mann = np.ascontiguousarray(mesh.mann)
            assert h.h_set_manning(self.p, d(self.mann)) == 0
        self.n = 3 * mesh.nnodes

    def state(self, s, so, soo=None):
        s, so = np.ascontiguousarray(s), np.ascontiguousarray(so)
        soo = so if soo is None else np.ascontiguousarray(soo)
        self.h.h_set_state(self.p, d(s), d(so), d(soo))

    def get(self):
        sol, res, bc, dt = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n, np.int32), np.zeros(1)
        wd = np.zeros(self.mesh.nelems2d, np.int32)
        self.h.h_get(self.p, d(sol), d(res), i(bc), d(dt), i(wd))
        return sol, res, bc, dt[0], wd

    def pattern(self, nnz):
        ip_, cols, vals = np.zeros(self.n + 1, np.int32), np.zeros(nnz, np.int32), np.zeros(nnz)
        self.h.h_pattern(self.p, i(ip_), i(cols), d(vals))
        return ip_, cols, vals

    def close(self):
        self.h.h_destroy(self.p)


@pytest.mark.parametrize("manning", [False, True])
def test_struct_surface_residual_jacobian_pattern(gpu, harness, manning):
    _, (mesh, sol0) = small_flume(npx=15, npy=7, seed=3)
    s, so = perturbed_state(sol0, seed=5)
    m = Model(harness, mesh, 0.8, manning=manning)
    orc = O.Oracle(mesh, dt=0.8, friction="manning" if manning else "const", cf_const=0.002)
    orc.set_state(s, so, so)
    m.state(s, so)
    nnz = harness.h_sparsity(m.p)
    ip_o, c_o = orc.pattern()
    assert nnz == orc.nnz
    ip_g, c_g, _ = m.pattern(nnz)
    assert np.array_equal(ip_g, ip_o) and np.array_equal(c_g, c_o)
    assert harness.h_residual(m.p) == 0
    _, res, _, _, _ = m.get()
    assert rel_err(res, orc.residual()) <= 1e-12
    assert harness.h_jacobian(m.p) == 0
    _, _, vals = m.pattern(nnz)
    assert rel_err(vals, orc.jacobian()) <= 1e-7
    # the elemental residual pointers, perturbed as perturb_var perturbs (variable codes 0, 4, 5)
    out9, out6 = np.zeros(9), np.zeros(6)
    for e in range(0, mesh.nelems2d, 7):
        for pnode, (code, pv) in [(0, (0, 0)), (2, (4, 1)), (1, (5, 2))]:
            assert harness.h_elem_resid(m.p, 0, e, 1e-4, pnode, code, -1, d(out9)) == 0
            ref = orc.body_resid(e, s, 1e-4, pnode, pv, -1)
            assert np.abs(out9 - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    for e in range(mesh.nelems1d):
        assert harness.h_elem_resid(m.p, 1, e, 1e-4, 1, 0, 1, d(out6)) == 0
        ref = orc.seg_resid(e, s, 1e-4, 1, 0, 1)
        assert np.abs(out6 - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # load_elem_mat: the reference's element loop of perturb_var over the residual pointers
    em = np.zeros(81)
    for e in (0, mesh.nelems2d // 2, mesh.nelems2d - 1):
        assert harness.h_elem_mat(m.p, e, d(em)) == 0
        ref = orc.elem_jac_tri(e).ravel()
        assert np.abs(em - ref).max() <= 1e-7 * np.abs(ref).max()
    m.close()


def test_struct_surface_check_diag_dirichlet_norms_increment(gpu, harness):
    _, (mesh, sol0) = small_flume()
    s, so = perturbed_state(sol0, seed=2)
    m = Model(harness, mesh, 1.0)
    orc = O.Oracle(mesh, dt=1.0, friction="const", cf_const=0.002)
    bc = np.zeros(m.n, np.int32)
    bc[[1, 2, 7, 30, 31]] = 1
    dd = np.zeros(m.n)
    m.state(s, so)
    harness.h_set_bc(m.p, i(bc), d(dd))
    orc.set_state(s, so, so)
    orc.set_bc(bc, dd)
    nnz = harness.h_sparsity(m.p)
    assert harness.h_residual(m.p) == 0 and harness.h_jacobian(m.p) == 0
    assert harness.h_check_diag_dirichlet(m.p) == 0
    orc.residual()
    orc.jacobian()
    orc.check_diag()
    orc.apply_dirichlet()
    vo, ro = orc.system()
    _, res, bcg, _, _ = m.get()
    _, _, vals = m.pattern(nnz)
    assert np.array_equal(bcg, orc.bc_mask())
    assert rel_err(vals, vo) <= 1e-7 and rel_err(res, ro) <= 1e-12
    # get_residual_norms over the dofs outside bc_mask, l2 = sqrt(sum r^2) / ndofs as coded
    rng = np.random.default_rng(1)
    r, ds = rng.normal(size=m.n), rng.normal(size=m.n)
    out, idx = np.zeros(3), np.zeros(2, np.int32)
    assert harness.h_norms(m.p, d(r), d(ds), d(out), i(idx)) == 0
    use = bcg != 1
    assert out[0] == np.abs(r[use]).max() and idx[0] == np.flatnonzero(use)[np.argmax(np.abs(r[use]))]
    assert out[2] == np.abs(ds[use]).max() and idx[1] == np.flatnonzero(use)[np.argmax(np.abs(ds[use]))]
    assert abs(out[1] - np.sqrt(np.sum(r[use] ** 2)) / m.n) <= 1e-14 * out[1]
    sol_before = m.get()[0]
    assert harness.h_increment(m.p, d(ds)) == 0
    assert np.array_equal(m.get()[0], sol_before + ds)
    m.close()


def test_struct_surface_time_loop_matches_oracle(gpu, harness):
    """initialize_system then ten time steps of fe_newton on the 200 m flume, against the oracle's"""
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    m = Model(harness, mesh, 1.0)
    orc = O.Oracle(mesh, dt=1.0, friction="const", cf_const=0.002, threads=8)
    m.state(sol0, sol0)
    orc.set_initial(sol0)
    assert harness.h_init_system(m.p) == 0
    its_g, its_o = [], []
    nit = C.c_int()
    for k in range(10):
        assert harness.h_step(m.p, C.byref(nit)) == 1
        its_g.append(nit.value)
        n, st = orc.time_loop(1)
        assert n == 1
        its_o.append(st["newton_its"])
    sol, _, _, dt, _ = m.get()
    so_ = orc.state()[0]
    assert its_g == its_o
    assert dt == 1.0
    assert rel_err(sol[0::3], so_[0::3]) <= 1e-8
    assert rel_err(sol[1::3], so_[1::3]) <= 1e-6
    m.close()


def test_struct_surface_design_time_loop_under_hydrograph(gpu, harness):
    """time_loop(SMODEL_DESIGN *) on the 200 m flume with the inflow flux a time series (a rising
    hydrograph, averaged over each step by advance_time) against the oracle stepped by hand with the
    averages from scipy quadrature: 15 steps of 1 s, the same Newton iterations, depth 1e-8"""
    from scipy.integrate import quad
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    width = A.FLUME_C1.ymax - A.FLUME_C1.ymin
    t_tab, q_tab = np.array([0.0, 6.0, 20.0]), np.array([10.0, 25.0, 40.0])
    qn = np.ascontiguousarray(-q_tab / width)  # flux per unit width, negative into the domain
    dt, nsteps = 1.0, 15
    m = Model(harness, mesh, dt)
    m.state(sol0, sol0)
    assert harness.h_init_system(m.p) == 0
    assert harness.h_design(m.p, 0.0, nsteps * dt, dt, len(t_tab), d(np.ascontiguousarray(t_tab)), d(qn)) == 0
    t_end, dt_end, its = np.zeros(1), np.zeros(1), C.c_int()
    assert harness.h_time_loop(m.p, d(t_end), d(dt_end), C.byref(its)) == 1
    assert t_end[0] == pytest.approx(nsteps * dt) and dt_end[0] == dt

    orc = O.Oracle(mesh, dt=dt, friction="const", cf_const=0.002, threads=8)
    orc.set_initial(sol0)
    inflow = np.flatnonzero(mesh.seg_kind == O.BC_FLUX)
    val = mesh.seg_val.copy()
    its_o = 0
    for k in range(nsteps):
        t1, t2 = k * dt, (k + 1) * dt
        pts = [p for p in t_tab if t1 < p < t2]
        qavg = quad(lambda s: np.interp(s, t_tab, qn), t1, t2, points=pts or None, epsabs=1e-14)[0] / dt
        val[inflow, 0] = qavg
        orc.set_seg_values(val)
        n, st = orc.time_loop(1)
        assert n == 1
        its_o += st["newton_its"]
    assert its.value == its_o
    sol = m.get()[0]
    so_ = orc.state()[0]
    assert rel_err(sol[0::3], so_[0::3]) <= 1e-8
    assert rel_err(sol[1::3], so_[1::3]) <= 1e-6
    m.close()
