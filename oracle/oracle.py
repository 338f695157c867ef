"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker and the CPU baseline.  The product
(adh_kraken_amd) never imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC_DEPTH, BC_FLUX, BC_VEL, BC_DEPTH_NODAL = 0, 1, 2, 3
FRIC_CONST, FRIC_MANNING = 0, 1


class OrMesh(C.Structure):
    _fields_ = [("nnodes", C.c_int), ("nelems2d", C.c_int), ("nelems1d", C.c_int),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("tri", C.c_void_p),
                ("seg", C.c_void_p), ("seg_kind", C.c_void_p), ("seg_val", C.c_void_p), ("mann", C.c_void_p)]


class OrParams(C.Structure):
    _fields_ = [("g", C.c_double), ("dt", C.c_double), ("alpha", C.c_double), ("cf_const", C.c_double),
                ("fric_mode", C.c_int), ("manning_units", C.c_double)]


class OrSolver(C.Structure):
    _fields_ = [("tol_nonlin", C.c_double), ("inc_nonlin", C.c_double), ("max_nonlin_it", C.c_int),
                ("max_linesearch_cuts", C.c_int), ("lin_tol", C.c_double), ("lin_max_it", C.c_int),
                ("precond", C.c_int)]

PREC_BJ, PREC_SGS = 0, 1


class OrStats(C.Structure):
    _fields_ = [("newton_its", C.c_int), ("linesearch_cuts_total", C.c_int), ("lin_its_total", C.c_int),
                ("lin_its_max", C.c_int), ("resid_max_norm", C.c_double), ("inc_max_norm", C.c_double),
                ("converged", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None
P = C.c_void_p
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "or_create": (P, [C.POINTER(OrMesh), C.POINTER(OrParams)]),
        "or_destroy": (None, [P]),
        "or_ndofs": (C.c_int, [P]),
        "or_nnz": (C.c_int, [P]),
        "or_set_threads": (None, [C.c_int]),
        "or_get_geometry": (None, [P, dp, dp, dp, dp]),
        "or_set_state": (None, [P, dp, dp, dp]),
        "or_get_state": (None, [P, dp, dp, dp]),
        "or_set_bc": (None, [P, ip, dp]),
        "or_get_bc": (None, [P, ip]),
        "or_set_dt": (None, [P, C.c_double]),
        "or_sw2_body_resid": (None, [P, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]),
        "or_sw2_seg_resid": (None, [P, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]),
        "or_wdflag": (None, [P, ip]),
        "or_grid_mass": (C.c_double, [P, C.c_double]),
        "or_sw2_init": (None, [P]),
        "or_assemble_residual": (None, [P, dp]),
        "or_pattern": (None, [P, ip, ip]),
        "or_elem_jacobian_tri": (None, [P, C.c_int, dp]),
        "or_elem_jacobian_seg": (None, [P, C.c_int, dp]),
        "or_assemble_jacobian": (None, [P, dp]),
        "or_check_diag": (None, [P]),
        "or_apply_dirichlet": (None, [P]),
        "or_get_system": (None, [P, dp, dp]),
        "or_set_system": (None, [P, dp, dp]),
        "or_spmv": (None, [C.c_int, ip, ip, dp, dp, dp]),
        "or_scale_linear_system": (None, [C.c_int, ip, ip, dp, dp, dp, dp]),
        "or_unscale": (None, [C.c_int, dp, dp, dp]),
        "or_block_jacobi_prep": (None, [C.c_int, ip, ip, dp, dp]),
        "or_block_jacobi_apply": (None, [C.c_int, dp, dp, dp]),
        "or_bicgstab": (C.c_int, [C.c_int, ip, ip, dp, dp, dp, dp, dp, C.c_double, C.c_int, dp]),
        "or_colour_dsatur": (C.c_int, [C.c_int, ip, ip, ip]),
        "or_colour_reduce": (C.c_int, [C.c_int, ip, ip, ip, C.c_int, C.c_int]),
        "or_bicgstab_sgs": (C.c_int, [C.c_int, ip, ip, dp, dp, ip, C.c_int, dp, dp, dp, C.c_double, C.c_int, dp]),
        "or_residual_norms": (None, [C.c_int, dp, dp, ip, dp, dp, dp]),
        "or_fe_newton": (C.c_int, [P, C.POINTER(OrSolver), C.POINTER(OrStats)]),
        "or_time_loop": (C.c_int, [P, C.POINTER(OrSolver), C.c_int, C.POINTER(OrStats)]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype = r
        f.argtypes = a
    _lib = lib
    return lib


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def _i(a):
    return None if a is None else a.ctypes.data_as(ip)


def default_solver(**kw):
    s = OrSolver(tol_nonlin=1e-8, inc_nonlin=1e-6, max_nonlin_it=10, max_linesearch_cuts=5, lin_tol=1e-11,
                 lin_max_it=100, precond=PREC_BJ)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


class Oracle:
    """CPU restatement of one SW2 super model."""

    def __init__(self, mesh, dt=1.0, g=9.8, friction="manning", cf_const=0.002, alpha=0.0, manning_units=1.0,
                 threads=None):
        self.lib = load()
        if threads:
            self.lib.or_set_threads(int(threads))
        m = mesh
        self._arr = dict(
            x=np.ascontiguousarray(m.x, np.float64), y=np.ascontiguousarray(m.y, np.float64),
            z=np.ascontiguousarray(m.z, np.float64), tri=np.ascontiguousarray(m.tri, np.int32),
            seg=np.ascontiguousarray(m.seg, np.int32), kind=np.ascontiguousarray(m.seg_kind, np.int32),
            val=np.ascontiguousarray(m.seg_val, np.float64),
            mann=np.ascontiguousarray(m.mann if m.mann is not None else np.zeros(len(m.tri)), np.float64))
        a = self._arr
        om = OrMesh(len(a["x"]), len(a["tri"]), len(a["seg"]), a["x"].ctypes.data, a["y"].ctypes.data,
                    a["z"].ctypes.data, a["tri"].ctypes.data, a["seg"].ctypes.data, a["kind"].ctypes.data,
                    a["val"].ctypes.data, a["mann"].ctypes.data)
        p = OrParams(g, dt, alpha, cf_const, FRIC_MANNING if friction == "manning" else FRIC_CONST, manning_units)
        self.h = C.c_void_p(self.lib.or_create(C.byref(om), C.byref(p)))
        self.ndofs = self.lib.or_ndofs(self.h)
        self.nnz = self.lib.or_nnz(self.h)
        self.nnodes = len(a["x"])
        self.E = len(a["tri"])
        self.S = len(a["seg"])

    def __del__(self):
        try:
            if self.h:
                self.lib.or_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_state(self, sol=None, sol_old=None, sol_older=None):
        f = lambda v: None if v is None else np.ascontiguousarray(v, np.float64)
        s, so, soo = f(sol), f(sol_old), f(sol_older)
        self.lib.or_set_state(self.h, _d(s), _d(so), _d(soo))

    def set_initial(self, sol):
        self.set_state(sol, sol, sol)

    def grid_mass(self, density=1.0):
        """Water volume (x density) of the current depth, tl_find_grid_mass_elem2d."""
        return self.lib.or_grid_mass(self.h, float(density))

    def set_seg_values(self, val):
        """Boundary values (S, 2) for the following calls (the context reads this array in place)."""
        self._arr["val"][...] = np.asarray(val, np.float64).reshape(self._arr["val"].shape)

    def state(self):
        a, b, c = (np.zeros(self.ndofs) for _ in range(3))
        self.lib.or_get_state(self.h, _d(a), _d(b), _d(c))
        return a, b, c

    def set_bc(self, bc_mask=None, dirichlet=None):
        bm = None if bc_mask is None else np.ascontiguousarray(bc_mask, np.int32)
        dd = None if dirichlet is None else np.ascontiguousarray(dirichlet, np.float64)
        self.lib.or_set_bc(self.h, _i(bm), _d(dd))

    def bc_mask(self):
        a = np.zeros(self.ndofs, np.int32)
        self.lib.or_get_bc(self.h, _i(a))
        return a

    def set_dt(self, dt):
        self.lib.or_set_dt(self.h, dt)

    def geometry(self):
        djac, grad = np.zeros(self.E), np.zeros(6 * self.E)
        sd, sn = np.zeros(max(self.S, 1)), np.zeros(2 * max(self.S, 1))
        self.lib.or_get_geometry(self.h, _d(djac), _d(grad), _d(sd), _d(sn))
        return djac, grad.reshape(self.E, 3, 2), sd[:self.S], sn[:2 * self.S].reshape(self.S, 2)

    def wdflag(self):
        a = np.zeros(self.E, np.int32)
        self.lib.or_wdflag(self.h, _i(a))
        return a

    def sw2_init(self):
        self.lib.or_sw2_init(self.h)

    def body_resid(self, ie, sol, eps=0.0, pnode=-1, pvar=-1, psign=0):
        r = np.zeros(9)
        sol = np.ascontiguousarray(sol, np.float64)
        self.lib.or_sw2_body_resid(self.h, ie, eps, pnode, pvar, psign, _d(sol), _d(r))
        return r

    def seg_resid(self, ie, sol, eps=0.0, pnode=-1, pvar=-1, psign=0):
        r = np.zeros(6)
        sol = np.ascontiguousarray(sol, np.float64)
        self.lib.or_sw2_seg_resid(self.h, ie, eps, pnode, pvar, psign, _d(sol), _d(r))
        return r

    def residual(self):
        r = np.zeros(self.ndofs)
        self.lib.or_assemble_residual(self.h, _d(r))
        return r

    def pattern(self):
        ip_, c = np.zeros(self.ndofs + 1, np.int32), np.zeros(self.nnz, np.int32)
        self.lib.or_pattern(self.h, _i(ip_), _i(c))
        return ip_, c

    def elem_jac_tri(self, ie):
        m = np.zeros(81)
        self.lib.or_elem_jacobian_tri(self.h, ie, _d(m))
        return m.reshape(9, 9)

    def elem_jac_seg(self, ie):
        m = np.zeros(36)
        self.lib.or_elem_jacobian_seg(self.h, ie, _d(m))
        return m.reshape(6, 6)

    def jacobian(self):
        v = np.zeros(self.nnz)
        self.lib.or_assemble_jacobian(self.h, _d(v))
        return v

    def check_diag(self):
        self.lib.or_check_diag(self.h)

    def apply_dirichlet(self):
        self.lib.or_apply_dirichlet(self.h)

    def system(self):
        v, r = np.zeros(self.nnz), np.zeros(self.ndofs)
        self.lib.or_get_system(self.h, _d(v), _d(r))
        return v, r

    def fe_newton(self, solver=None):
        st = OrStats()
        rc = self.lib.or_fe_newton(self.h, C.byref(solver or default_solver()), C.byref(st))
        return rc, st.as_dict()

    def time_loop(self, nsteps, solver=None):
        st = OrStats()
        rc = self.lib.or_time_loop(self.h, C.byref(solver or default_solver()), nsteps, C.byref(st))
        return rc, st.as_dict()


# plain-array linear algebra
def spmv(indptr, cols, vals, x):
    lib = load()
    y = np.zeros(len(indptr) - 1)
    lib.or_spmv(len(indptr) - 1, _i(indptr), _i(cols), _d(vals), _d(np.ascontiguousarray(x)), _d(y))
    return y


def scale_linear_system(indptr, cols, vals, b, x):
    lib = load()
    vals, b, x = vals.copy(), b.copy(), x.copy()
    s = np.zeros(len(b))
    lib.or_scale_linear_system(len(b), _i(indptr), _i(cols), _d(vals), _d(b), _d(x), _d(s))
    return vals, b, x, s


def block_jacobi(indptr, cols, vals):
    lib = load()
    n = len(indptr) - 1
    m = np.zeros(3 * n)
    lib.or_block_jacobi_prep(n // 3, _i(indptr), _i(cols), _d(vals), _d(m))
    return m


def bicgstab(indptr, cols, vals, minv, b, scale, x0=None, tol=1e-11, max_it=100):
    lib = load()
    n = len(b)
    x = np.zeros(n) if x0 is None else x0.copy()
    bb = b.copy()
    rn = C.c_double()
    it = lib.or_bicgstab(n, _i(indptr), _i(cols), _d(vals), _d(minv), _d(x), _d(bb), _d(scale), tol, max_it,
                         C.byref(rn))
    return x, it, rn.value, bb


def colour_dsatur(indptr, cols):
    """node colours by DSatur (saturation order) and their count"""
    lib = load()
    nn = (len(indptr) - 1) // 3
    col = np.zeros(nn, np.int32)
    nc = lib.or_colour_dsatur(nn, _i(indptr), _i(cols), _i(col))
    return col, nc


KEMPE_CAP = 4096


def colour_reduce(indptr, cols, colour, nc, cap=KEMPE_CAP):
    """a colouring with fewer colours by Kempe-chain exchanges (the input colouring when none frees the top one)"""
    lib = load()
    nn = (len(indptr) - 1) // 3
    col = np.ascontiguousarray(colour, np.int32).copy()
    nc2 = lib.or_colour_reduce(nn, _i(indptr), _i(cols), _i(col), int(nc), int(cap))
    return col, nc2


def colour_sgs(indptr, cols):
    """the SGS colouring the Newton solve uses: DSatur, then the Kempe-chain reduction"""
    col, nc = colour_dsatur(indptr, cols)
    return colour_reduce(indptr, cols, col, nc)


def bicgstab_sgs(indptr, cols, vals, minv, b, scale, colour, ncolour, x0=None, tol=1e-11, max_it=100):
    """split multicolour-SGS BiCGStab; it = -1 when a diagonal block is singular"""
    lib = load()
    n = len(b)
    x = np.zeros(n) if x0 is None else x0.copy()
    rn = C.c_double()
    it = lib.or_bicgstab_sgs(n, _i(indptr), _i(cols), _d(vals), _d(minv), _i(colour), ncolour, _d(x), _d(b),
                             _d(scale), tol, max_it, C.byref(rn))
    return x, it, rn.value
