"""GPU tests of ensembles: several independent meshes solved in one context (one launch per stage).

Every instance keeps its own BiCGStab scalars, Newton norms, line search and stopping test, and every
reduction sums the instance's own 256-node blocks in order, so an instance inside the ensemble must
reproduce its standalone solve bit for bit -- state and iteration counts.  Standalone solves are
themselves checked against the oracle in test_gpu_parity.py.
"""
import numpy as np
import pytest

import adh_kraken_amd as A
from adh_kraken_amd import _lib as L

pytestmark = pytest.mark.gpu


def uneven_meshes():
    specs = [
        A.FLUME_C1,
        A.FlumeSpec(xmax=60.0, ymax=12.0, npx=61, npy=13, nbumps=3, bump_amp_max=0.2, seed=5, manning_n=0.04,
                    inflow_q=3.0, tailwater=0.8, ic_flat_wse=1),
        # bumps above the still water: partially dry elements
        A.FlumeSpec(xmax=100.0, ymax=10.0, npx=101, npy=11, slope=1e-3, nbumps=6, bump_amp_max=1.6,
                    bump_sigma_min=2.0, bump_sigma_max=6.0, seed=3, manning_n=0.03, inflow_q=2.0, tailwater=1.0,
                    ic_flat_wse=1),
        A.FlumeSpec(xmax=150.0, ymax=6.0, npx=151, npy=7, nbumps=2, bump_amp_max=0.1, seed=9, manning_n=0.025,
                    inflow_q=1.5, tailwater=0.6, ic_flat_wse=1),
    ]
    return [A.flume_mesh(s) for s in specs]


@pytest.mark.parametrize("precision", [L.PREC_FP64, L.PREC_MIXED])
def test_ensemble_reproduces_standalone_solves(gpu, precision):
    cases = uneven_meshes()
    ens = A.SModelSuper.ensemble([m for m, _ in cases], dt=1.0)
    assert ens.ninst == len(cases)
    ens.set_initial(np.concatenate([s for _, s in cases]))
    n, st = A.time_loop(ens, 6, precision=precision)
    assert n == 6
    per = ens.instance_stats()
    sol = ens.sol
    for i, (mesh, sol0) in enumerate(cases):
        sm = A.SModelSuper(mesh, dt=1.0)
        sm.set_initial(sol0)
        ns, sts = A.time_loop(sm, 6, precision=precision)
        assert ns == 6
        assert per[i]["newton_its"] == sts["newton_its"], i
        assert per[i]["lin_its_total"] == sts["lin_its_total"], i
        assert np.array_equal(sol[ens.instance_slice(i)], sm.sol), i
    # the ensemble totals are the instance sums
    assert st["newton_its"] == sum(p["newton_its"] for p in per)
    assert st["lin_its_total"] == sum(p["lin_its_total"] for p in per)


def test_failed_member_drops_out_and_the_others_continue(gpu):
    cases = uneven_meshes()[:3]
    sols = [s.copy() for _, s in cases]
    sols[1][:] = np.nan  # this member cannot converge
    ens = A.SModelSuper.ensemble([m for m, _ in cases], dt=1.0)
    ens.set_initial(np.concatenate(sols))
    n, st = A.time_loop(ens, 4)
    alive, done = ens.instance_status()
    assert list(alive) == [1, 0, 1] and list(done) == [4, 0, 4]
    assert n == 0  # the ensemble as a whole did not complete (its slowest member did no step)
    sol = ens.sol
    for i in (0, 2):
        mesh, sol0 = cases[i]
        sm = A.SModelSuper(mesh, dt=1.0)
        sm.set_initial(sol0)
        assert A.time_loop(sm, 4)[0] == 4
        assert np.array_equal(sol[ens.instance_slice(i)], sm.sol), i
    assert ens.get_dt() == 1.0  # no member's failure cuts the shared time step
    ens.set_initial(np.concatenate([s for _, s in cases]))
    assert list(ens.instance_status()[0]) == [1, 1, 1]


def test_ensemble_of_the_batched_workload_converges(gpu):
    ens_cases = A.flume_ensemble(n=8, seed=4, tri_min=5_000, tri_max=20_000)
    ens = A.SModelSuper.ensemble([m for m, _, _ in ens_cases], dt=1.0)
    ens.set_initial(np.concatenate([s for _, s, _ in ens_cases]))
    n, st = A.time_loop(ens, 3, precision=L.PREC_MIXED)
    assert n == 3 and st["converged"] == 1
    per = ens.instance_stats()
    assert all(p["converged"] == 1 for p in per)
    assert np.all(np.isfinite(ens.sol))
