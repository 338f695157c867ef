"""ctypes binding of libadhk.so (the gfx950 C-ABI declared in include/adhk.h).

The library is built in-tree by ``adh_kraken_amd.build.build()``.  Loading
fails loudly when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libadhk.so")

BC_DEPTH, BC_FLUX, BC_VEL, BC_DEPTH_NODAL = 0, 1, 2, 3
FRIC_CONST, FRIC_MANNING = 0, 1
PREC_FP64, PREC_MIXED = 0, 1
PC_BJ, PC_SGS = 0, 1
KT_NAMES = ["spmv_p", "spmv_s", "krylov_update", "elem_jac", "jac_gather", "elem_resid", "resid_gather", "prep"]


class Params(C.Structure):
    _fields_ = [("g", C.c_double), ("dt", C.c_double), ("alpha", C.c_double), ("cf_const", C.c_double),
                ("manning_units", C.c_double), ("fric_mode", C.c_int)]


class Solver(C.Structure):
    _fields_ = [("tol_nonlin", C.c_double), ("inc_nonlin", C.c_double), ("max_nonlin_it", C.c_int),
                ("max_linesearch_cuts", C.c_int), ("lin_tol", C.c_double), ("lin_max_it", C.c_int),
                ("precision", C.c_int), ("lin_check_every", C.c_int), ("mixed_lin_tol", C.c_double),
                ("warm_gate", C.c_int), ("substep_depth", C.c_int), ("precond", C.c_int)]


class NewtonStats(C.Structure):
    _fields_ = [("newton_its", C.c_int), ("linesearch_cuts_total", C.c_int), ("lin_its_total", C.c_int),
                ("lin_its_max", C.c_int), ("resid_max_norm", C.c_double), ("inc_max_norm", C.c_double),
                ("converged", C.c_int), ("warm_accepted", C.c_int), ("warm_retried", C.c_int),
                ("substeps", C.c_int), ("lin_capped", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Timing(C.Structure):
    _fields_ = [("residual_ms", C.c_double), ("jacobian_ms", C.c_double), ("prep_ms", C.c_double),
                ("krylov_ms", C.c_double), ("update_ms", C.c_double), ("total_ms", C.c_double),
                ("residual_calls", C.c_int), ("jacobian_calls", C.c_int), ("krylov_calls", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("chunk", C.c_int), ("graphs", C.c_int), ("xcd_map", C.c_int), ("elem_tpb", C.c_int),
                ("split", C.c_int), ("stream_matrix", C.c_int), ("sgs_wide", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FlumeSpec(C.Structure):
    _fields_ = [("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double), ("ymax", C.c_double),
                ("npx", C.c_int), ("npy", C.c_int), ("slope", C.c_double), ("nbumps", C.c_int),
                ("bump_amp_max", C.c_double), ("bump_sigma_min", C.c_double), ("bump_sigma_max", C.c_double),
                ("seed", C.c_ulonglong), ("manning_n", C.c_double), ("inflow_q", C.c_double),
                ("tailwater", C.c_double), ("ic_flat_wse", C.c_int)]


class ChannelSpec(C.Structure):
    _fields_ = [("length", C.c_double), ("width", C.c_double), ("spacing", C.c_double),
                ("chan_width", C.c_double), ("chan_depth", C.c_double), ("bank_width", C.c_double),
                ("meander_amp", C.c_double), ("meander_wavelength", C.c_double), ("slope", C.c_double),
                ("nbumps", C.c_int), ("bump_amp_max", C.c_double), ("bump_sigma_min", C.c_double),
                ("bump_sigma_max", C.c_double), ("seed", C.c_ulonglong), ("n_channel", C.c_double),
                ("n_floodplain", C.c_double), ("inflow_q", C.c_double), ("fp_depth_out", C.c_double),
                ("meander_phase", C.c_double)]


P = C.c_void_p
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
fp = C.POINTER(C.c_float)
MLP_NIN, MLP_HID, MLP_NOUT, MLP_NNORM = 12, 64, 3, 36

# (name, restype, argtypes) of every symbol include/adhk.h declares
SIGNATURES = [
    ("adhk_last_error", C.c_char_p, []),
    ("adhk_device_synchronize", C.c_int, []),
    ("adhk_solver_defaults", None, [C.POINTER(Solver)]),
    ("adhk_sw2_create", P, [C.c_int, dp, dp, dp, C.c_int, ip, dp, C.c_int, ip, ip, dp, C.POINTER(Params)]),
    ("adhk_sw2_create_batch", P, [C.c_int, ip, ip, ip, C.c_int, dp, dp, dp, C.c_int, ip, dp, C.c_int, ip, ip, dp,
                                  C.POINTER(Params)]),
    ("adhk_sw2_instances", C.c_int, [P]),
    ("adhk_sw2_instance_status", C.c_int, [P, ip, ip]),
    ("adhk_sw2_instance_stats", C.c_int, [P, C.POINTER(NewtonStats)]),
    ("adhk_sw2_destroy", None, [P]),
    ("adhk_sw2_sizes", C.c_int, [P, ip, ip, ip]),
    ("adhk_sw2_layout", C.c_int, [P, ip, ip]),
    ("adhk_sw2_set_state", C.c_int, [P, dp, dp, dp]),
    ("adhk_sw2_get_state", C.c_int, [P, dp, dp, dp]),
    ("adhk_sw2_set_bc", C.c_int, [P, ip, dp]),
    ("adhk_sw2_get_bc", C.c_int, [P, ip]),
    ("adhk_sw2_set_dt", C.c_int, [P, C.c_double]),
    ("adhk_sw2_get_dt", C.c_double, [P]),
    ("adhk_sw2_get_timing", C.c_int, [P, C.POINTER(Timing)]),
    ("adhk_sw2_kernel_timing", C.c_int, [P, C.c_int]),
    ("adhk_sw2_kernel_timing_mask", C.c_int, [P, C.c_uint, C.c_int]),
    ("adhk_sw2_kernel_times", C.c_int, [P, dp, C.POINTER(C.c_longlong)]),
    ("adhk_copy_probe", C.c_int, [P, C.c_longlong, C.c_int, dp]),
    ("adhk_sw2_geometry", C.c_int, [P, dp, dp, dp, dp]),
    ("adhk_sw2_get_tuning", C.c_int, [P, C.POINTER(Tuning)]),
    ("adhk_sw2_set_tuning", C.c_int, [P, C.POINTER(Tuning)]),
    ("adhk_sw2_autotune", C.c_int, [P, C.POINTER(Solver), C.c_int, C.POINTER(Tuning), dp]),
    ("adhk_sw2_colouring", C.c_int, [P, ip, ip, ip]),
    ("adhk_sw2_init", C.c_int, [P, ip]),
    ("adhk_assemble_residual", C.c_int, [P, dp]),
    ("adhk_elem_residuals", C.c_int, [P, dp]),
    ("adhk_sparsity", C.c_int, [P, ip, ip]),
    ("adhk_assemble_jacobian", C.c_int, [P, dp]),
    ("adhk_elem_jacobians", C.c_int, [P, dp]),
    ("adhk_elem_jacobians_mixed", C.c_int, [P, dp]),
    ("adhk_check_diag", C.c_int, [P]),
    ("adhk_apply_dirichlet_bc", C.c_int, [P]),
    ("adhk_get_system", C.c_int, [P, dp, dp]),
    ("adhk_fe_newton", C.c_int, [P, C.POINTER(Solver), C.POINTER(NewtonStats)]),
    ("adhk_time_loop", C.c_int, [P, C.POINTER(Solver), C.c_int, C.POINTER(NewtonStats)]),
    ("split_CSR_mat_vec_mult", None, [dp, ip, ip, dp, ip, ip, dp, dp, C.c_int, ip, C.c_int]),
    ("scale_linear_system", None, [ip, ip, dp, ip, ip, dp, dp, dp, dp, C.c_int, C.c_int, C.c_int, ip, C.c_int]),
    ("unscale_linear_system", None, [dp, dp, dp, C.c_int]),
    ("prep_umfpack", C.c_int, [ip, ip, dp, C.c_int]),
    ("solve_umfpack", C.c_int, [dp, ip, ip, dp, dp, C.c_int]),
    ("umfpack_clear", None, []),
    ("free_bcgstab", None, []),
    ("solve_linear_sys_bcgstab", C.c_int,
     [dp, ip, ip, dp, ip, ip, dp, dp, dp, C.c_int, C.c_int, C.c_int, ip, C.c_int]),
    ("Screen_print_CSR", None, [ip, ip, dp, C.c_int]),
    ("adhk_la_set_block_size", None, [C.c_int]),
    ("adhk_la_last_status", C.c_int, []),
    ("adhk_elem_residual_one", C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, dp]),
    ("adhk_sw2_set_seg_values", C.c_int, [P, dp]),
    ("adhk_la_last_iterations", C.c_int, []),
    ("adhk_bcgstab_device_noalloc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int)]),
    ("adhk_flume_counts", C.c_int, [C.POINTER(FlumeSpec), ip, ip, ip]),
    ("adhk_flume_build", C.c_int, [C.POINTER(FlumeSpec), dp, dp, dp, ip, dp, ip, ip, dp]),
    ("adhk_flume_initial_state", C.c_int, [C.POINTER(FlumeSpec), C.c_int, dp, dp]),
    ("adhk_sw2_grid_mass", C.c_int, [P, C.c_double, dp]),
    ("adhk_sw2_set_series", C.c_int, [P, C.c_int, C.c_int, dp, dp]),
    ("adhk_sw2_bind_series", C.c_int, [P, C.c_int, ip, C.c_int, C.c_int, C.c_double]),
    ("adhk_sw2_set_time", C.c_int, [P, C.c_double]),
    ("adhk_sw2_get_time", C.c_double, [P]),
    ("adhk_series_average", C.c_int, [C.c_int, dp, dp, C.c_double, C.c_double, dp]),
    ("adhk_geo_counts", C.c_int, [C.c_char_p, ip, ip, ip, ip]),
    ("adhk_geo_read", C.c_int, [C.c_char_p, C.c_int, dp, dp, dp, C.c_int, ip, ip, C.c_int, ip, ip]),
    ("adhk_poisson_disk_bound", C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]),
    ("adhk_poisson_disk", C.c_int, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_ulonglong,
                                    C.c_int, dp, dp]),
    ("adhk_rcm_order", C.c_int, [C.c_int, C.c_int, ip, ip]),
    ("adhk_channel_bed", C.c_int, [C.POINTER(ChannelSpec), C.c_int, dp, dp, dp]),
    ("adhk_channel_roughness", C.c_int, [C.POINTER(ChannelSpec), C.c_int, dp, dp, C.c_int, ip, dp]),
    ("adhk_channel_section_weight", C.c_double, [C.POINTER(ChannelSpec), C.c_double, C.c_double]),
    ("adhk_mlp_param_count", C.c_int, []),
    ("adhk_mlp_create", P, [C.c_ulonglong]),
    ("adhk_mlp_destroy", None, [P]),
    ("adhk_mlp_get_params", C.c_int, [P, fp, fp]),
    ("adhk_mlp_set_params", C.c_int, [P, fp, fp]),
    ("adhk_mlp_data_reserve", C.c_int, [P, C.c_longlong]),
    ("adhk_mlp_collect", C.c_int, [P, P]),
    ("adhk_mlp_data_size", C.c_longlong, [P]),
    ("adhk_mlp_fit_normalization", C.c_int, [P]),
    ("adhk_mlp_train", C.c_int, [P, C.c_int, C.c_int, C.c_double, C.c_ulonglong, dp]),
    ("adhk_mlp_forward", C.c_int, [P, C.c_int, fp, fp]),
    ("adhk_mlp_loss_grad", C.c_int, [P, C.c_int, fp, fp, fp, dp]),
    ("adhk_sw2_features", C.c_int, [P, fp]),
    ("adhk_sw2_set_warm_start", C.c_int, [P, P]),
]

_lib = None


class AdhkError(RuntimeError):
    pass


def load():
    """Load libadhk.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AdhkError(f"{LIB_PATH} is missing: build it with adh_kraken_amd.build.build() "
                        "(the SW2 path has no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().adhk_last_error().decode(errors="replace")


def check(rc, what):
    if rc is None or rc < 0:
        raise AdhkError(f"{what} failed: {last_error()}")
    return rc


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous array expected"
    return a.ctypes.data_as(dp)


def fptr(a):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "float32 C-contiguous array expected"
    return a.ctypes.data_as(fp)


def iptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"], "int32 C-contiguous array expected"
    return a.ctypes.data_as(ip)


def solver_defaults(**kw):
    s = Solver()
    load().adhk_solver_defaults(C.byref(s))
    for k, v in kw.items():
        setattr(s, k, v)
    return s
