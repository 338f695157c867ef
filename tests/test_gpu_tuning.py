"""GPU tests of the schedule choices, hipGraph replay and the autotuner.

No schedule changes a computed value, so every run below must reproduce the default schedule bit for
bit (state and iteration counts).  With graphs on, each BiCGStab chunk is one captured graph replayed
many times per solve and across solves, with eager kernels and host synchronisations between the
replays; every replay's results feed the compared state.
"""
import numpy as np
import pytest

import adh_kraken_amd as A
from adh_kraken_amd import _lib as L

pytestmark = pytest.mark.gpu

SCHEDULES = [
    {"graphs": 1, "chunk": 8},
    {"graphs": 1, "chunk": 4, "xcd_map": 0},
    {"graphs": 0, "chunk": 16, "elem_tpb": 64},
    {"graphs": 1, "chunk": 16, "elem_tpb": 128},
    {"split": 1},
    {"split": 1, "graphs": 1, "chunk": 8, "xcd_map": 0},
]


def run(mesh, sol0, steps, precision, tuning=None, ensemble=None):
    sm = A.SModelSuper.ensemble(ensemble, dt=1.0) if ensemble else A.SModelSuper(mesh, dt=1.0)
    if tuning:
        sm.tuning = tuning
    sm.set_initial(sol0)
    n, st = A.time_loop(sm, steps, precision=precision)
    return n, st, sm.sol, sm


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_schedules_reproduce_the_default(gpu, precision):
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    n0, st0, s0, _ = run(mesh, sol0, 6, precision)
    for t in SCHEDULES:
        n, st, s, sm = run(mesh, sol0, 6, precision, t)
        assert all(sm.tuning[k] == v for k, v in t.items())
        assert n == n0 and st["newton_its"] == st0["newton_its"] and st["lin_its_total"] == st0["lin_its_total"], t
        assert np.array_equal(s, s0), t


def test_graph_schedule_on_an_ensemble(gpu):
    meshes = [A.flume_mesh(A.FLUME_C1)[0],
              A.flume_mesh(A.FlumeSpec(xmax=60.0, ymax=12.0, npx=61, npy=13, nbumps=3, bump_amp_max=0.2, seed=5,
                                       manning_n=0.04, inflow_q=3.0, tailwater=0.8, ic_flat_wse=1))[0]]
    sols = [A.flume_mesh(A.FLUME_C1)[1],
            A.flume_mesh(A.FlumeSpec(xmax=60.0, ymax=12.0, npx=61, npy=13, nbumps=3, bump_amp_max=0.2, seed=5,
                                     manning_n=0.04, inflow_q=3.0, tailwater=0.8, ic_flat_wse=1))[1]]
    sol0 = np.concatenate(sols)
    n0, st0, s0, _ = run(None, sol0, 5, L.PREC_FP64, None, meshes)
    n1, st1, s1, _ = run(None, sol0, 5, L.PREC_FP64, {"graphs": 1, "chunk": 8}, meshes)
    assert n0 == n1 == 5 and st0["lin_its_total"] == st1["lin_its_total"]
    assert np.array_equal(s0, s1)


def test_autotune_keeps_results(gpu):
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    n0, st0, s0, _ = run(mesh, sol0, 5, L.PREC_MIXED)
    sm = A.SModelSuper(mesh, dt=1.0)
    sm.set_initial(sol0)
    choice, t = sm.autotune(precision=L.PREC_MIXED, ninner=2)
    assert choice["elem_tpb"] in (64, 128, 256) and choice["chunk"] in (4, 8, 16)
    assert t["krylov_ms"] <= t["default_krylov_ms"] and t["krylov_its"] > 0
    assert np.array_equal(sm.sol, sol0)  # tuning leaves the state alone
    n, st = A.time_loop(sm, 5, precision=L.PREC_MIXED)
    assert n == 5 and st["lin_its_total"] == st0["lin_its_total"]
    assert np.array_equal(sm.sol, s0)
