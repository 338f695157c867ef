"""Host-side mirror of the reference's SW2 super-model interface over libadhk.so.

Names follow the reference: an ``SModelSuper`` carries ``sol``, ``sol_old``,
``sol_older``, ``bc_mask``, ``dirichlet_data`` and the Newton knobs of
SMODEL_SUPER (src/structs/smodel_super/smodel_super.h:7-113, defaults from
smodel_super_alloc_init.c:55-100); ``assemble_residual``, ``assemble_jacobian``,
``fe_newton`` and ``time_loop`` take it like the reference functions
(residual.h:5, jacobian.h:5, newton.h:19, time_loop.c:14).  Arrays live on the
GPU inside the context; the properties copy to and from host numpy arrays.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


@dataclass
class Mesh:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    tri: np.ndarray            # (E, 3) int32, counter-clockwise
    seg: np.ndarray            # (S, 2) int32, domain on the left
    seg_kind: np.ndarray       # (S,) int32, BC_DEPTH / BC_FLUX / BC_VEL
    seg_val: np.ndarray        # (S, 2) float64
    mann: np.ndarray = None    # (E,) Manning n

    @property
    def nnodes(self):
        return int(self.x.shape[0])

    @property
    def nelems2d(self):
        return int(self.tri.shape[0])

    @property
    def nelems1d(self):
        return int(self.seg.shape[0])


@dataclass
class FlumeSpec:
    xmin: float = 0.0
    xmax: float = 200.0
    ymin: float = 0.0
    ymax: float = 20.0
    npx: int = 201
    npy: int = 21
    slope: float = 1e-4
    nbumps: int = 0
    bump_amp_max: float = 0.0
    bump_sigma_min: float = 5.0
    bump_sigma_max: float = 30.0
    seed: int = 1
    manning_n: float = 0.03
    inflow_q: float = 10.0
    tailwater: float = 1.0
    ic_flat_wse: int = 0

    def c(self):
        s = L.FlumeSpec()
        for k in self.__dataclass_fields__:
            setattr(s, k, getattr(self, k))
        return s


# the configurations of the workload set (sizes and physics as specified)
FLUME_C1 = FlumeSpec()  # 200 m x 20 m at 1 m, n = 0.03, Q = 10 m^3/s, tailwater 1.0 m, h0 = 1.0 m
FLUME_C2 = FlumeSpec(xmax=2000.0, ymax=200.0, npx=2001, npy=201, nbumps=50, bump_amp_max=0.3,
                     bump_sigma_min=5.0, bump_sigma_max=30.0, seed=1, inflow_q=200.0, tailwater=2.0,
                     ic_flat_wse=1)


def flume_mesh(spec: FlumeSpec):
    """Rectangular flume (create_rectangular_grid.c:4-445 numbering) with boundary segments."""
    lib = L.load()
    cs = spec.c()
    n, e, s = C.c_int(), C.c_int(), C.c_int()
    L.check(lib.adhk_flume_counts(C.byref(cs), C.byref(n), C.byref(e), C.byref(s)), "adhk_flume_counts")
    N, E, S = n.value, e.value, s.value
    x, y, z = (np.zeros(N) for _ in range(3))
    tri = np.zeros((E, 3), np.int32)
    mann = np.zeros(E)
    seg = np.zeros((S, 2), np.int32)
    kind = np.zeros(S, np.int32)
    val = np.zeros((S, 2))
    L.check(lib.adhk_flume_build(C.byref(cs), L.dptr(x), L.dptr(y), L.dptr(z), L.iptr(tri), L.dptr(mann),
                                 L.iptr(seg), L.iptr(kind), L.dptr(val)), "adhk_flume_build")
    sol = np.zeros(3 * N)
    L.check(lib.adhk_flume_initial_state(C.byref(cs), N, L.dptr(z), L.dptr(sol)), "adhk_flume_initial_state")
    return Mesh(x, y, z, tri, seg, kind, val, mann), sol


@dataclass
class TimeInfo:  # SDT (src/structs/sdt/sdt.h:7-16)
    dt: float = 1.0
    tau_temporal: float = 0.0


class SModelSuper:
    """One SW2 super model on one GPU (SMODEL_SUPER)."""

    def __init__(self, mesh: Mesh, dt=1.0, gravity=9.8, friction="manning", cf_const=0.002,
                 tau_temporal=0.0, manning_units=1.0, instances=None):
        """instances: None for one mesh, or (node_off, elem_off, seg_off) of an ensemble whose arrays are
        concatenated in mesh (see SModelSuper.ensemble)."""
        self.lib = L.load()
        self.mesh = mesh
        self.gravity = gravity
        p = L.Params(g=gravity, dt=dt, alpha=tau_temporal, cf_const=cf_const, manning_units=manning_units,
                     fric_mode=L.FRIC_MANNING if friction == "manning" else L.FRIC_CONST)
        m = mesh
        mann = m.mann if m.mann is not None else np.zeros(m.nelems2d)
        self._keep = [np.ascontiguousarray(a) for a in (m.x, m.y, m.z)]
        tri = np.ascontiguousarray(m.tri, np.int32)
        seg = np.ascontiguousarray(m.seg, np.int32)
        kind = np.ascontiguousarray(m.seg_kind, np.int32)
        val = np.ascontiguousarray(m.seg_val, np.float64)
        mann = np.ascontiguousarray(mann, np.float64)
        segargs = (L.iptr(seg) if m.nelems1d else None, L.iptr(kind) if m.nelems1d else None,
                   L.dptr(val) if m.nelems1d else None, C.byref(p))
        if instances is None:
            h = self.lib.adhk_sw2_create(m.nnodes, L.dptr(self._keep[0]), L.dptr(self._keep[1]),
                                         L.dptr(self._keep[2]), m.nelems2d, L.iptr(tri), L.dptr(mann), m.nelems1d,
                                         *segargs)
            self.offsets = (np.array([0, m.nnodes], np.int32), np.array([0, m.nelems2d], np.int32),
                            np.array([0, m.nelems1d], np.int32))
        else:
            self.offsets = tuple(np.ascontiguousarray(o, np.int32) for o in instances)
            no, eo, so = self.offsets
            h = self.lib.adhk_sw2_create_batch(len(no) - 1, L.iptr(no), L.iptr(eo), L.iptr(so), m.nnodes,
                                               L.dptr(self._keep[0]), L.dptr(self._keep[1]), L.dptr(self._keep[2]),
                                               m.nelems2d, L.iptr(tri), L.dptr(mann), m.nelems1d, *segargs)
        if not h:
            raise L.AdhkError("adhk_sw2_create failed: " + L.last_error())
        self.h = C.c_void_p(h)
        nd, nnz, nnzb = C.c_int(), C.c_int(), C.c_int()
        self.lib.adhk_sw2_sizes(self.h, C.byref(nd), C.byref(nnz), C.byref(nnzb))
        self.ndofs, self.nnz, self.nnz_blocks = nd.value, nnz.value, nnzb.value
        nsl, nso = C.c_int(), C.c_int()
        self.lib.adhk_sw2_layout(self.h, C.byref(nsl), C.byref(nso))
        self.nslice, self.nslot = nsl.value, nso.value
        # Newton knobs (smodel_super_alloc_init.c:62-65) and linear solver knobs (bcgstab_solver.c:20-22)
        self.solver = L.solver_defaults()
        self.ts = TimeInfo(dt=dt, tau_temporal=tau_temporal)

    @classmethod
    def ensemble(cls, meshes, **kw):
        """One context for several independent meshes (an ensemble): arrays concatenated, node ids offset."""
        no, eo, so = [0], [0], [0]
        for m in meshes:
            no.append(no[-1] + m.nnodes)
            eo.append(eo[-1] + m.nelems2d)
            so.append(so[-1] + m.nelems1d)
        cat = Mesh(np.concatenate([m.x for m in meshes]), np.concatenate([m.y for m in meshes]),
                   np.concatenate([m.z for m in meshes]),
                   np.concatenate([m.tri + o for m, o in zip(meshes, no)]).astype(np.int32),
                   np.concatenate([m.seg + o for m, o in zip(meshes, no)]).astype(np.int32),
                   np.concatenate([m.seg_kind for m in meshes]).astype(np.int32),
                   np.concatenate([m.seg_val for m in meshes]),
                   np.concatenate([m.mann if m.mann is not None else np.zeros(m.nelems2d) for m in meshes]))
        return cls(cat, instances=(no, eo, so), **kw)

    @property
    def ninst(self):
        return len(self.offsets[0]) - 1

    def instance_slice(self, i):
        """dof slice of instance i in the state vectors"""
        no = self.offsets[0]
        return slice(3 * int(no[i]), 3 * int(no[i + 1]))

    # ---- schedule choices (none changes a result) ----
    @property
    def tuning(self):
        t = L.Tuning()
        L.check(self.lib.adhk_sw2_get_tuning(self.h, C.byref(t)), "adhk_sw2_get_tuning")
        return t.as_dict()

    @tuning.setter
    def tuning(self, d):
        t = L.Tuning(**{**self.tuning, **d})
        L.check(self.lib.adhk_sw2_set_tuning(self.h, C.byref(t)), "adhk_sw2_set_tuning")

    def autotune(self, precision=L.PREC_FP64, ninner=2):
        """Time every schedule on this model's own Newton system at the current state; keep the fastest.
        Returns (choice, {elem_jac_ms, krylov_ms, default_krylov_ms, krylov_its})."""
        self.solver.precision = precision
        t = L.Tuning()
        ms = np.zeros(4)
        L.check(self.lib.adhk_sw2_autotune(self.h, C.byref(self.solver), ninner, C.byref(t), L.dptr(ms)),
                "adhk_sw2_autotune")
        return t.as_dict(), {"elem_jac_ms": ms[0], "krylov_ms": ms[1], "default_krylov_ms": ms[2],
                             "krylov_its": int(ms[3])}

    def colouring(self, layout=False):
        """(colour [nnodes], colour count) of the multicolour SGS preconditioner; with layout, also
        {rows, u_slots, l_slots}: the permuted row count and the SELL-64 slot columns of U and L"""
        col = np.zeros(self.mesh.nnodes, np.int32)
        nc = C.c_int()
        lay = np.zeros(3, np.int32)
        L.check(self.lib.adhk_sw2_colouring(self.h, col.ctypes.data_as(L.ip), C.byref(nc), lay.ctypes.data_as(L.ip)),
                "adhk_sw2_colouring")
        if layout:
            return col, nc.value, {"rows": int(lay[0]), "u_slots": int(lay[1]), "l_slots": int(lay[2])}
        return col, nc.value

    def instance_status(self):
        """(alive [ninst], steps_done [ninst]) of an ensemble's members"""
        a, d = np.zeros(self.ninst, np.int32), np.zeros(self.ninst, np.int32)
        L.check(self.lib.adhk_sw2_instance_status(self.h, L.iptr(a), L.iptr(d)), "adhk_sw2_instance_status")
        return a, d

    def instance_stats(self):
        st = (L.NewtonStats * self.ninst)()
        L.check(self.lib.adhk_sw2_instance_stats(self.h, st), "adhk_sw2_instance_stats")
        return [x.as_dict() for x in st]

    def close(self):
        if getattr(self, "h", None):
            self.lib.adhk_sw2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- Newton knobs with the reference's names ----
    tol_nonlin = property(lambda s: s.solver.tol_nonlin, lambda s, v: setattr(s.solver, "tol_nonlin", v))
    inc_nonlin = property(lambda s: s.solver.inc_nonlin, lambda s, v: setattr(s.solver, "inc_nonlin", v))
    max_nonlin_it = property(lambda s: s.solver.max_nonlin_it, lambda s, v: setattr(s.solver, "max_nonlin_it", v))
    max_nonlin_linesearch_cuts = property(lambda s: s.solver.max_linesearch_cuts,
                                          lambda s, v: setattr(s.solver, "max_linesearch_cuts", v))

    # ---- state ----
    def _get(self, which):
        a = np.zeros(self.ndofs)
        args = [None, None, None]
        args[which] = L.dptr(a)
        L.check(self.lib.adhk_sw2_get_state(self.h, *args), "adhk_sw2_get_state")
        return a

    def _set(self, which, v):
        v = np.ascontiguousarray(v, np.float64)
        assert v.shape == (self.ndofs,)
        args = [None, None, None]
        args[which] = L.dptr(v)
        L.check(self.lib.adhk_sw2_set_state(self.h, *args), "adhk_sw2_set_state")

    sol = property(lambda s: s._get(0), lambda s, v: s._set(0, v))
    sol_old = property(lambda s: s._get(1), lambda s, v: s._set(1, v))
    sol_older = property(lambda s: s._get(2), lambda s, v: s._set(2, v))

    @property
    def bc_mask(self):
        a = np.zeros(self.ndofs, np.int32)
        L.check(self.lib.adhk_sw2_get_bc(self.h, L.iptr(a)), "adhk_sw2_get_bc")
        return a

    def set_bc(self, bc_mask=None, dirichlet_data=None):
        bm = None if bc_mask is None else np.ascontiguousarray(bc_mask, np.int32)
        dd = None if dirichlet_data is None else np.ascontiguousarray(dirichlet_data, np.float64)
        L.check(self.lib.adhk_sw2_set_bc(self.h, L.iptr(bm), L.dptr(dd)), "adhk_sw2_set_bc")

    def set_dt(self, dt):
        L.check(self.lib.adhk_sw2_set_dt(self.h, dt), "adhk_sw2_set_dt")
        self.ts.dt = dt

    def get_dt(self):
        return self.lib.adhk_sw2_get_dt(self.h)

    def grid_mass(self, density=1.0):
        """Water volume (x density) at the current depth (tl_find_grid_mass_elem2d)."""
        out = C.c_double()
        L.check(self.lib.adhk_sw2_grid_mass(self.h, float(density), C.byref(out)), "adhk_sw2_grid_mass")
        return out.value

    # time series (SSERIES, src/structs/sseries/sseries.c): boundary values that follow a hydrograph
    def set_series(self, sid, t, v):
        t = np.ascontiguousarray(t, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        L.check(self.lib.adhk_sw2_set_series(self.h, int(sid), t.size, L.dptr(t), L.dptr(v)), "adhk_sw2_set_series")

    def bind_series(self, segs, sid, scale=1.0, component=0):
        """Segments whose value `component` becomes scale * (series sid averaged over each step); sid -1
        unbinds."""
        segs = np.ascontiguousarray(segs, np.int32)
        L.check(self.lib.adhk_sw2_bind_series(self.h, segs.size, L.iptr(segs), int(component), int(sid),
                                              float(scale)), "adhk_sw2_bind_series")

    @property
    def time(self):
        """Time of the old level (t_prev): 0 at creation, + dt per completed step."""
        return self.lib.adhk_sw2_get_time(self.h)

    @time.setter
    def time(self, t):
        L.check(self.lib.adhk_sw2_set_time(self.h, float(t)), "adhk_sw2_set_time")

    def set_initial(self, sol):
        """sol = sol_old = sol_older = the given state (a cold start)."""
        for w in range(3):
            self._set(w, sol)

    def geometry(self):
        E, S = self.mesh.nelems2d, self.mesh.nelems1d
        djac, grad, sd, sn = np.zeros(E), np.zeros(6 * E), np.zeros(max(S, 1)), np.zeros(2 * max(S, 1))
        L.check(self.lib.adhk_sw2_geometry(self.h, L.dptr(djac), L.dptr(grad), L.dptr(sd), L.dptr(sn)),
                "adhk_sw2_geometry")
        return djac, grad.reshape(E, 3, 2), sd[:S], sn[:2 * S].reshape(S, 2)


# ---- reference-named operations ----
def fe_sw2_init(sm: SModelSuper):
    """fe_sw2_init (fe_sw2_init.c:28): wet/dry flags, dry nodes get u,v Dirichlet.  Returns wd flags."""
    wd = np.zeros(sm.mesh.nelems2d, np.int32)
    L.check(sm.lib.adhk_sw2_init(sm.h, L.iptr(wd)), "adhk_sw2_init")
    return wd


def assemble_residual(sm: SModelSuper):
    """assemble_residual (assemble_residual.c:69): residual = -sum of elemental residuals at sm.sol."""
    r = np.zeros(sm.ndofs)
    L.check(sm.lib.adhk_assemble_residual(sm.h, L.dptr(r)), "adhk_assemble_residual")
    return r


def elem_residuals(sm: SModelSuper):
    E, S = sm.mesh.nelems2d, sm.mesh.nelems1d
    out = np.zeros(9 * E + 6 * S)
    L.check(sm.lib.adhk_elem_residuals(sm.h, L.dptr(out)), "adhk_elem_residuals")
    return out[:9 * E].reshape(E, 9), out[9 * E:].reshape(S, 6)


def sparsity(sm: SModelSuper):
    """slin_sys_init_sparsity_mono (slin_sys_init_sparsity_mono.c:28): CSR pattern built on the GPU."""
    indptr = np.zeros(sm.ndofs + 1, np.int32)
    cols = np.zeros(sm.nnz, np.int32)
    L.check(sm.lib.adhk_sparsity(sm.h, L.iptr(indptr), L.iptr(cols)), "adhk_sparsity")
    return indptr, cols


def assemble_jacobian(sm: SModelSuper):
    """assemble_jacobian (assemble_jacobian.c:20): central-FD Jacobian values in the CSR order of sparsity()."""
    vals = np.zeros(sm.nnz)
    L.check(sm.lib.adhk_assemble_jacobian(sm.h, L.dptr(vals)), "adhk_assemble_jacobian")
    return vals


def elem_jacobians(sm: SModelSuper):
    E, S = sm.mesh.nelems2d, sm.mesh.nelems1d
    out = np.zeros(81 * E + 36 * S)
    L.check(sm.lib.adhk_elem_jacobians(sm.h, L.dptr(out)), "adhk_elem_jacobians")
    return out[:81 * E].reshape(E, 9, 9), out[81 * E:].reshape(S, 6, 6)


def elem_jacobians_mixed(sm: SModelSuper):
    """The elemental matrices of the mixed-precision solve (central differences in FP64, stored in FP32),
    widened to float64; same layout as elem_jacobians."""
    E, S = sm.mesh.nelems2d, sm.mesh.nelems1d
    out = np.zeros(81 * E + 36 * S)
    L.check(sm.lib.adhk_elem_jacobians_mixed(sm.h, L.dptr(out)), "adhk_elem_jacobians_mixed")
    return out[:81 * E].reshape(E, 9, 9), out[81 * E:].reshape(S, 6, 6)


def check_diag(sm: SModelSuper):
    L.check(sm.lib.adhk_check_diag(sm.h), "adhk_check_diag")


def apply_Dirichlet_BC(sm: SModelSuper):
    L.check(sm.lib.adhk_apply_dirichlet_bc(sm.h), "adhk_apply_dirichlet_bc")


def get_system(sm: SModelSuper):
    vals, res = np.zeros(sm.nnz), np.zeros(sm.ndofs)
    L.check(sm.lib.adhk_get_system(sm.h, L.dptr(vals), L.dptr(res)), "adhk_get_system")
    return vals, res


def fe_newton(sm: SModelSuper, precision=L.PREC_FP64):
    """fe_newton (fe_newton.c:44): returns (TRUE/FALSE, stats dict)."""
    sm.solver.precision = precision
    st = L.NewtonStats()
    rc = sm.lib.adhk_fe_newton(sm.h, C.byref(sm.solver), C.byref(st))
    L.check(rc, "adhk_fe_newton")
    sm.ts.dt = sm.get_dt()
    return rc, st.as_dict()


def time_loop(sm: SModelSuper, nsteps, precision=L.PREC_FP64):
    """time_loop (time_loop.c:14): nsteps x (prep_sol + fe_newton).  Returns (steps done, stats dict)."""
    sm.solver.precision = precision
    st = L.NewtonStats()
    rc = sm.lib.adhk_time_loop(sm.h, C.byref(sm.solver), nsteps, C.byref(st))
    if rc <= -1000000000:
        raise L.AdhkError("adhk_time_loop failed: " + L.last_error())
    sm.ts.dt = sm.get_dt()
    return rc, st.as_dict()


def series_average(t, v, t1, t2):
    """Average over [t1, t2] of the piecewise-linear series (t, v), as the time loop evaluates it."""
    t = np.ascontiguousarray(t, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    out = C.c_double()
    L.check(L.load().adhk_series_average(t.size, L.dptr(t), L.dptr(v), float(t1), float(t2), C.byref(out)),
            "adhk_series_average")
    return out.value


def flume_ensemble(n=64, seed=4, tri_min=5_000, tri_max=200_000):
    """Uneven flumes of the batched-ensemble workload, all from one seed.

    Per instance: length uniform [200, 4000] m, width uniform [20, 200] m, triangle count log-uniform
    [tri_min, tri_max] (node spacing sqrt(2 L W / T)), Manning n uniform [0.02, 0.05], inflow uniform
    [1, 50] m^3/s per 10 m of width, tailwater depth uniform [0.5, 3] m, bed slope 1e-4 plus seeded bumps
    (at most 0.3 m, capped so the initial flat water surface leaves at least 0.1 m of depth).
    Returns [(mesh, sol0, spec)]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        Lx = rng.uniform(200.0, 4000.0)
        W = rng.uniform(20.0, 200.0)
        T = float(np.exp(rng.uniform(np.log(tri_min), np.log(tri_max))))
        h = np.sqrt(2.0 * Lx * W / T)
        npx = max(3, int(round(Lx / h)) + 1)
        npy = max(3, int(round(W / h)) + 1)
        mann = rng.uniform(0.02, 0.05)
        q10 = rng.uniform(1.0, 50.0)
        tail = rng.uniform(0.5, 3.0)
        slope = 1e-4
        amp = float(np.clip(tail - slope * Lx - 0.1, 0.0, 0.3))
        spec = FlumeSpec(xmin=0.0, xmax=Lx, ymin=0.0, ymax=W, npx=npx, npy=npy, slope=slope,
                         nbumps=8 if amp > 0 else 0, bump_amp_max=amp, bump_sigma_min=5.0, bump_sigma_max=30.0,
                         seed=int(rng.integers(1, 2**62)), manning_n=mann, inflow_q=q10 * W / 10.0, tailwater=tail,
                         ic_flat_wse=1)
        mesh, sol0 = flume_mesh(spec)
        out.append((mesh, sol0, spec))
    return out
