"""The la.h surface (src/la/la.h) on the GPU: workspace growth, point Jacobi, per-thread workspaces."""
import threading

import numpy as np
import pytest

import oracle as O
from adh_kraken_amd import _lib as L
from helpers import perturbed_state, rel_err, small_flume

pytestmark = pytest.mark.gpu


def scaled_system(npx=11, npy=5, seed=9):
    """the scaled Newton system of a perturbed small flume, from the oracle"""
    _, (mesh, sol0) = small_flume(npx=npx, npy=npy)
    s, so = perturbed_state(sol0, seed=seed)
    orc = O.Oracle(mesh, dt=1.0, threads=4)
    orc.set_state(s, so, so)
    orc.residual()
    vals = orc.jacobian()
    ip_, cols = orc.pattern()
    b = orc.system()[1]
    vo, bo, _, so_ = O.scale_linear_system(ip_, cols, vals, b, np.zeros(len(b)))
    return ip_, cols, vo, bo, so_


def la_solve(lib, ip_, cols, vals, b, scale):
    n = len(b)
    x, bb = np.zeros(n), b.copy()
    rc = lib.solve_linear_sys_bcgstab(L.dptr(x), L.iptr(ip_), L.iptr(cols), L.dptr(vals), None, None, None,
                                      L.dptr(bb), L.dptr(scale), n, n, 0, None, 0)
    return rc, x, bb


def test_prepared_preconditioner_survives_workspace_growth(gpu):
    """prep, then a call on a larger system grows the workspace, then the solve of the prepared size still
    finds its block inverses (no stale or uninitialised preconditioner)"""
    lib = L.load()
    ip_, cols, v, b, sc = scaled_system()
    n = len(b)
    assert lib.prep_umfpack(L.iptr(ip_), L.iptr(cols), L.dptr(v), n) == 0
    rc, x_ref, b_ref = la_solve(lib, ip_, cols, v, b, sc)
    assert rc == 0
    lib.free_bcgstab()
    assert lib.prep_umfpack(L.iptr(ip_), L.iptr(cols), L.dptr(v), n) == 0
    m = 5 * n  # larger than the workspace: every vector is reallocated
    xx, x0, ss = np.ones(m), np.ones(m), np.full(m, 2.0)
    lib.unscale_linear_system(L.dptr(xx), L.dptr(x0), L.dptr(ss), m)
    assert lib.adhk_la_last_status() == 0 and np.all(xx == 4.0)
    rc, x, bb = la_solve(lib, ip_, cols, v, b, sc)
    assert rc == 0
    assert np.array_equal(x, x_ref) and np.array_equal(bb, b_ref)
    lib.free_bcgstab()


def test_solve_without_prep_is_an_error(gpu):
    lib = L.load()
    ip_, cols, v, b, sc = scaled_system()
    lib.free_bcgstab()
    rc, x, _ = la_solve(lib, ip_, cols, v, b, sc)
    assert rc == -1 and lib.adhk_la_last_status() == -1
    assert b"prep_umfpack" in L.load().adhk_last_error()
    assert np.all(x == 0.0)


def test_point_jacobi_matches_oracle_solution(gpu):
    """block size 1: the solution still solves the system (point Jacobi is a weaker preconditioner)"""
    lib = L.load()
    ip_, cols, v, b, sc = scaled_system()
    n = len(b)
    lib.adhk_la_set_block_size(1)
    try:
        assert lib.prep_umfpack(L.iptr(ip_), L.iptr(cols), L.dptr(v), n) == 0
        rc, x, _ = la_solve(lib, ip_, cols, v, b, sc)
    finally:
        lib.adhk_la_set_block_size(3)
        lib.free_bcgstab()
    assert rc == 0
    xs = x / sc  # back to the scaled unknowns
    r = O.spmv(ip_, cols, v, xs) - b
    assert np.abs(r).max() <= 1e-8 * np.abs(b).max()


def test_threads_own_their_workspaces(gpu):
    """two host threads solving different systems at once get the results each gets alone"""
    lib = L.load()
    systems = [scaled_system(seed=9), scaled_system(npx=13, npy=6, seed=4)]
    alone = []
    for ip_, cols, v, b, sc in systems:
        assert lib.prep_umfpack(L.iptr(ip_), L.iptr(cols), L.dptr(v), len(b)) == 0
        alone.append(la_solve(lib, ip_, cols, v, b, sc))
        lib.free_bcgstab()
    out = [None, None]

    def run(k):
        ip_, cols, v, b, sc = systems[k]
        res = []
        for _ in range(3):
            assert lib.prep_umfpack(L.iptr(ip_), L.iptr(cols), L.dptr(v), len(b)) == 0
            res.append(la_solve(lib, ip_, cols, v, b, sc))
        lib.free_bcgstab()
        out[k] = res

    ts = [threading.Thread(target=run, args=(k,)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=60)
    for k in range(2):
        assert out[k] is not None
        for rc, x, bb in out[k]:
            assert rc == 0
            assert np.array_equal(x, alone[k][1]) and np.array_equal(bb, alone[k][2])
    assert rel_err(alone[0][1], alone[0][1]) == 0.0
