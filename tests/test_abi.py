"""The C-ABI library loads and exports every entry point include/adhk.h declares (no GPU calls)."""
import ctypes
import os
import re

import adh_kraken_amd._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "adhk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", txt))
    return sorted(n for n in names if n not in {"if", "for", "while", "return", "sizeof"} and not n.isupper())


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


def test_binding_covers_header():
    bound = {n for n, _, _ in L.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_load_sets_signatures():
    lib = L.load()
    assert lib.adhk_sw2_create.restype == ctypes.c_void_p
