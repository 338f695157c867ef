"""adhk_bcgstab_device_noalloc (BiCGStab on device arrays, no allocation) against the allocating la.h path
(prep_umfpack + solve_linear_sys_bcgstab on host arrays): the same kernels, so the same bits."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from adh_kraken_amd import _lib as L
from helpers import perturbed_state, small_flume

pytestmark = pytest.mark.gpu


def scaled_system(npx, npy, seed):
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


def allocating(lib, ip_, cols, v, b, sc, x0):
    n = len(b)
    assert lib.prep_umfpack(L.iptr(ip_), L.iptr(cols), L.dptr(v), n) == 0
    x, bb = x0.copy(), b.copy()
    assert lib.solve_linear_sys_bcgstab(L.dptr(x), L.iptr(ip_), L.iptr(cols), L.dptr(v), None, None, None,
                                        L.dptr(bb), L.dptr(sc), n, n, 0, None, 0) == 0
    return x, bb, lib.adhk_la_last_iterations()


def noalloc(lib, torch, ip_, cols, v, b, sc, x0):
    dev = torch.device("cuda:0")
    t = {k: torch.from_numpy(np.ascontiguousarray(a)).to(dev) for k, a in
         dict(ip=ip_, ci=cols, v=v, b=b, sc=sc, x=x0).items()}
    torch.cuda.synchronize()
    its = C.c_int(-1)
    p = lambda k: C.c_void_p(t[k].data_ptr())  # noqa: E731
    rc = lib.adhk_bcgstab_device_noalloc(p("x"), p("ip"), p("ci"), p("v"), p("b"), p("sc"), len(b), 3, None,
                                         C.byref(its))
    torch.cuda.synchronize()
    assert rc == 0, L.last_error()
    return t["x"].cpu().numpy(), t["b"].cpu().numpy(), its.value


def test_noalloc_matches_allocating_solve(gpu):
    torch = pytest.importorskip("torch")
    lib = L.load()
    rng = np.random.default_rng(3)
    # two calls in a row on the one scratch array, the second on a smaller system
    for npx, npy, seed in ((41, 11, 9), (15, 7, 4)):
        ip_, cols, v, b, sc = scaled_system(npx, npy, seed)
        x0 = rng.uniform(-1e-3, 1e-3, len(b))
        xa, ba, ita = allocating(lib, ip_, cols, v, b, sc, x0)
        xn, bn, itn = noalloc(lib, torch, ip_, cols, v, b, sc, x0)
        assert ita == itn and ita > 0
        assert np.array_equal(xa, xn) and np.array_equal(ba, bn)
        r = O.spmv(ip_, cols, v, xn / sc - x0) - b  # the solution solves the scaled system
        assert np.abs(r).max() <= 1e-8 * np.abs(b).max()
    lib.free_bcgstab()


def test_noalloc_rejects_oversize(gpu):
    lib = L.load()
    n = 262144 + 3
    rc = lib.adhk_bcgstab_device_noalloc(None, None, None, None, None, None, n, 3, None, None)
    assert rc == -1 and b"262144" in lib.adhk_last_error()
