"""GPU tests of the schedule choices, hipGraph replay and the autotuner.

No schedule changes a computed value, so every run below must reproduce the default schedule bit for
bit (state and iteration counts).  With graphs on, each BiCGStab chunk and each Newton stage (residual
and norms; Jacobian, fold and Krylov start; Krylov finish, increment and residual) is one captured graph
replayed many times per solve and across solves, with eager kernels, work-mask uploads and host
synchronisations between the replays; every replay's results feed the compared state.
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
    {"stream_matrix": 1},
    {"stream_matrix": 1, "split": 1, "graphs": 1, "chunk": 8},
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


def test_stage_graphs_recapture_when_dt_changes(gpu):
    """dt is a by-value kernel argument of the residual and Jacobian stages: after set_dt the captured
    stages must be captured again, or the replays would keep integrating with the old step."""
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    out = []
    for tuning in (None, {"graphs": 1, "chunk": 8}):
        sm = A.SModelSuper(mesh, dt=1.0)
        if tuning:
            sm.tuning = tuning
        sm.set_initial(sol0)
        n1, st1 = A.time_loop(sm, 3, precision=L.PREC_MIXED)
        sm.set_dt(0.5)
        n2, st2 = A.time_loop(sm, 3, precision=L.PREC_MIXED)
        out.append((n1, n2, st1["newton_its"] + st2["newton_its"], st1["lin_its_total"] + st2["lin_its_total"],
                    sm.sol))
        sm.close()
    (a1, a2, an, al, sa), (b1, b2, bn, bl, sb) = out
    assert a1 == a2 == b1 == b2 == 3 and an == bn and al == bl
    assert np.array_equal(sa, sb)


def test_stage_graphs_with_wetting_and_line_search(gpu):
    """A flume whose bed bumps rise above the water (wet/dry queue in every residual and Jacobian stage,
    line-search residuals between the stage replays), FP64 and mixed, graphs against eager launches."""
    spec = A.FlumeSpec(xmin=0.0, xmax=100.0, ymin=0.0, ymax=10.0, npx=101, npy=11, slope=1e-3, nbumps=6,
                       bump_amp_max=1.6, bump_sigma_min=2.0, bump_sigma_max=6.0, seed=3, manning_n=0.03,
                       inflow_q=2.0, tailwater=1.0, ic_flat_wse=1)
    mesh, sol0 = A.flume_mesh(spec)
    for precision in (L.PREC_FP64, L.PREC_MIXED):
        n0, st0, s0, _ = run(mesh, sol0, 8, precision)
        n1, st1, s1, _ = run(mesh, sol0, 8, precision, {"graphs": 1, "chunk": 4})
        assert n0 == n1 == 8, precision
        assert (st0["newton_its"], st0["lin_its_total"], st0["linesearch_cuts_total"]) == \
            (st1["newton_its"], st1["lin_its_total"], st1["linesearch_cuts_total"]), precision
        assert np.array_equal(s0, s1), precision
