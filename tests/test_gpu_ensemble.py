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


# members of the batched workload (flume_ensemble(64, seed=4)) with a supercritical inflow against a depth
# condition: at dt = 1 s the first step fails in the oracle as on the GPU; 44 converges at dt / 4 and 49 at
# dt / 16 (oracle runs of the same members); 14 and 55 converge at dt
C4_RECOVERED = (44, 49)
C4_PLAIN = (14, 55)


@pytest.fixture(scope="module")
def c4_cases():
    return A.flume_ensemble(n=64, seed=4)


def test_failed_members_retake_their_step_in_sub_steps(gpu, c4_cases):
    """With substep_depth > 0 a member whose step fails goes back to its previous level and retakes the step
    as quarter steps (a failed quarter again in quarters); it completes the step and stays in the ensemble.
    Every member equals its standalone run with the same setting, bit for bit, and a recovered member's
    first step equals four oracle steps of dt / 4."""
    import oracle as O
    idx = C4_PLAIN + C4_RECOVERED
    cases = [c4_cases[i] for i in idx]
    ens = A.SModelSuper.ensemble([m for m, _, _ in cases], dt=1.0)
    ens.solver.substep_depth = 3
    ens.set_initial(np.concatenate([s for _, s, _ in cases]))
    n, st = A.time_loop(ens, 1)
    assert n == 1 and st["converged"] == 1
    alive, done = ens.instance_status()
    assert list(alive) == [1, 1, 1, 1] and list(done) == [1, 1, 1, 1]
    per = ens.instance_stats()
    assert [p["substeps"] > 0 for p in per] == [False, False, True, True]
    assert per[3]["substeps"] > 4  # member 49 retakes a quarter in sixteenths
    sol = ens.sol
    for k, (mesh, sol0, _) in enumerate(cases):
        sm = A.SModelSuper(mesh, dt=1.0)
        sm.solver.substep_depth = 3
        sm.set_initial(sol0)
        ns, sts = A.time_loop(sm, 1)
        assert ns == 1 and sts["substeps"] == per[k]["substeps"], k
        assert np.array_equal(sol[ens.instance_slice(k)], sm.sol), k
        assert sm.get_dt() == 1.0
    mesh, sol0, _ = cases[2]
    orc = O.Oracle(mesh, dt=0.25, threads=8)
    orc.set_initial(sol0)
    assert orc.time_loop(4)[0] == 4
    ho, hg = orc.state()[0][0::3], sol[ens.instance_slice(2)][0::3]
    assert np.abs(hg - ho).max() <= 1e-8 * np.abs(ho).max()


def test_without_sub_steps_a_failed_member_drops_out(gpu, c4_cases):
    cases = [c4_cases[i] for i in (C4_PLAIN[0], C4_RECOVERED[0])]
    ens = A.SModelSuper.ensemble([m for m, _, _ in cases], dt=1.0)
    ens.set_initial(np.concatenate([s for _, s, _ in cases]))
    A.time_loop(ens, 1)
    alive, done = ens.instance_status()
    assert list(alive) == [1, 0] and list(done) == [1, 0]


def test_full_batched_workload_two_steps(gpu, c4_cases):
    """All 64 members (5k-200k triangles, 3.6M in all), mixed precision with sub-step recovery: 62 members
    complete both steps; sampled members (recovered ones among them) equal their standalone runs bit for
    bit; the two smallest match the FP64 oracle to the mixed-precision tolerance."""
    import oracle as O
    ens = A.SModelSuper.ensemble([m for m, _, _ in c4_cases], dt=1.0)
    ens.solver.substep_depth = 3
    ens.set_initial(np.concatenate([s for _, s, _ in c4_cases]))
    n, st = A.time_loop(ens, 2, precision=L.PREC_MIXED)
    alive, done = ens.instance_status()
    # 49 and 51 (inflow Froude numbers 4.6 and 16.5 against a depth condition) complete the first step in
    # sub-steps and fail shortly after t = 1.1 s at every dt down to 1/64 s, in the oracle as here; they
    # drop out, every other member completes both steps
    stuck = [i for i in range(len(c4_cases)) if not alive[i]]
    assert stuck == [49, 51] and list(done[stuck]) == [1, 1]
    assert all(done[i] == 2 for i in range(len(c4_cases)) if i not in stuck)
    assert n == -1  # the slowest member completed one step
    sol = ens.sol
    sizes = np.array([m.nelems2d for m, _, _ in c4_cases])
    small = list(np.argsort(sizes)[:2])
    for i in sorted(set(small) | {0, 19, 44, int(np.argmax(sizes))}):
        mesh, sol0, _ = c4_cases[i]
        sm = A.SModelSuper(mesh, dt=1.0)
        sm.solver.substep_depth = 3
        sm.set_initial(sol0)
        assert A.time_loop(sm, 2, precision=L.PREC_MIXED)[0] == 2
        assert np.array_equal(sol[ens.instance_slice(i)], sm.sol), i
    for i in small:
        mesh, sol0, _ = c4_cases[i]
        orc = O.Oracle(mesh, dt=1.0, threads=8)
        orc.set_initial(sol0)
        assert orc.time_loop(2)[0] == 2
        ho, hg = orc.state()[0][0::3], sol[ens.instance_slice(i)][0::3]
        assert np.abs(hg - ho).max() <= 1e-6 * np.abs(ho).max(), i
