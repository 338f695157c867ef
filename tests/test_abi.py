"""The C-ABI library loads and exports every entry point include/adhk.h declares (no GPU calls)."""
import ctypes
import os
import re

import numpy as np

import adh_kraken_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="adhk.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt))
    skip = {"if", "for", "while", "return", "sizeof", "int", "void", "double", "char"}  # incl. function pointers
    return sorted(n for n in names if n not in skip and not n.isupper())


def test_header_declares_reference_surface():
    names = declared_functions()
    for ref in ["split_CSR_mat_vec_mult", "scale_linear_system", "unscale_linear_system", "prep_umfpack",
                "solve_umfpack", "umfpack_clear", "free_bcgstab", "solve_linear_sys_bcgstab", "Screen_print_CSR",
                "adhk_assemble_residual", "adhk_assemble_jacobian", "adhk_fe_newton", "adhk_time_loop",
                "adhk_sparsity"]:
        assert ref in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_exports_the_struct_level_surface():
    lib = ctypes.CDLL(L.LIB_PATH)
    names = declared_functions("adhk_ref.h")
    for ref in ["assemble_residual", "assemble_jacobian", "fe_newton", "check_diag", "apply_Dirichlet_BC",
                "initialize_system", "increment_function", "get_residual_norms", "fe_sw2_body_resid", "fe_sw2_bc_h",
                "fe_sw2_bc_vel", "fe_sw2_init", "slin_sys_init_sparsity_mono", "load_elem_mat", "perturb_var",
                "get_cell_dofs_ivars", "load_global_resid", "add_replace_elem_rhs", "time_loop", "advance_time",
                "update_dt", "tc_init", "smodel_super_prep_sol", "sseries_setall_ts_valuesAVG", "fe_sw2_get_roughness",
                "fr_bedshstr_drag_coef", "messg_dsum", "messg_dmax", "messg_imin", "sivar_position_build_dof_map"]:
        assert ref in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {n for n, _, _ in L.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_load_sets_signatures():
    lib = L.load()
    assert lib.adhk_sw2_create.restype == ctypes.c_void_p


def test_la_errors_return_instead_of_aborting():
    """argument errors of the la.h surface set a status and a message, make no device call and leave the
    outputs untouched (the reference exits through tl_error)"""
    lib = L.load()
    ip_ = np.array([0, 1], np.int32)
    cols = np.array([0], np.int32)
    vals = np.array([2.0])
    x = np.array([1.0])
    y = np.array([7.0])
    off = np.array([0, 0], np.int32)
    lib.split_CSR_mat_vec_mult(L.dptr(y), L.iptr(ip_), L.iptr(cols), L.dptr(vals), L.iptr(off), None, None,
                               L.dptr(x), 1, None, 0)
    assert lib.adhk_la_last_status() == -1 and y[0] == 7.0
    assert b"multi-rank" in lib.adhk_last_error()
    sc = np.ones(1)
    lib.free_bcgstab()
    rc = lib.solve_linear_sys_bcgstab(L.dptr(x), L.iptr(ip_), L.iptr(cols), L.dptr(vals), None, None, None,
                                      L.dptr(y), L.dptr(sc), 1, 1, 0, None, 0)
    assert rc == -1 and lib.adhk_la_last_status() == -1 and x[0] == 1.0
    assert b"prep_umfpack" in lib.adhk_last_error()
