"""GPU checks of the channel workloads at their configured sizes, and of the learned warm start on a
meandering channel under a rising hydrograph (a scaled copy of the 8 km meander workload).

* 10 km x 2 km channel (1.37M nodes, 4.1M dofs) and the 8 km meander (0.99M nodes, 3.0M dofs): element
  residuals of 2,000 sampled triangles against the oracle at 1e-12, then FP64 and mixed precision over the
  first steps, depth within 1e-6 of each other (the mixed-precision bar).
* scaled meander (2 km, 60 m channel, 5 m spacing) under the hydrograph per unit channel width: FP64 against
  the oracle stepped by hand at depth 1e-8 with the same Newton counts; a warm start trained on three
  straight channels of the same family under the same hydrograph (40 steps each, 4000 Adam steps) is
  accepted on most steps, cuts Newton and BiCGStab iterations, and leaves the depth within 1e-6 of the
  plain FP64 run.
"""
import numpy as np
import pytest
from scipy.integrate import quad

import adh_kraken_amd as A
import oracle as O
from adh_kraken_amd import _lib as L
from adh_kraken_amd import learned as LR
from helpers import perturbed_state, rel_err

pytestmark = pytest.mark.gpu


def sampled_residuals(mesh, sol0, dt, nsample=2000, seed=0):
    s, so = perturbed_state(sol0, seed=seed, amp_h=0.02, amp_v=0.1)
    sm = A.SModelSuper(mesh, dt=dt)
    sm.sol, sm.sol_old, sm.sol_older = s, so, so
    eg, _ = A.elem_residuals(sm)
    orc = O.Oracle(mesh, dt=dt, threads=8)
    orc.set_state(s, so, so)
    idx = np.random.default_rng(seed).choice(mesh.nelems2d, nsample, replace=False)
    eo = np.array([orc.body_resid(int(e), s) for e in idx])
    sm.close()
    return eg[idx], eo


def fp64_vs_mixed(mesh, sol0, dt, nsteps, prepare=None):
    out = []
    for prec in (L.PREC_FP64, L.PREC_MIXED):
        sm = A.SModelSuper(mesh, dt=dt)
        sm.set_initial(sol0)
        if prepare is not None:
            prepare(sm)
        n, st = A.time_loop(sm, nsteps, precision=prec)
        assert n == nsteps, (prec, st)
        out.append(sm.sol)
        sm.close()
    return out


@pytest.mark.timeout(400)
def test_c3_channel_full_size(gpu):
    mesh, sol0 = A.channel_mesh(A.CHANNEL_C3)
    assert mesh.nnodes > 1_300_000
    eg, eo = sampled_residuals(mesh, sol0, 5.0)
    assert np.abs(eg - eo).max() <= 1e-12 * np.abs(eo).max()
    s64, smx = fp64_vs_mixed(mesh, sol0, 5.0, 2)
    assert rel_err(smx[0::3], s64[0::3]) <= 1e-6


@pytest.mark.timeout(400)
def test_c5_meander_full_size(gpu):
    mesh, sol0 = A.channel_mesh(A.CHANNEL_C5)
    assert mesh.nnodes > 900_000
    eg, eo = sampled_residuals(mesh, sol0, 5.0, seed=1)
    assert np.abs(eg - eo).max() <= 1e-12 * np.abs(eo).max()
    s64, smx = fp64_vs_mixed(mesh, sol0, 5.0, 2, prepare=lambda sm: A.bind_hydrograph(sm, *A.C5_HYDROGRAPH))
    assert rel_err(smx[0::3], s64[0::3]) <= 1e-6


# the meander workload scaled by 0.3 (2 km long, 60 m channel) and its hydrograph per unit channel width
SCALED = dict(spacing=5.0, chan_width=60.0, chan_depth=3.0, bank_width=10.0, slope=1e-4, nbumps=8,
              bump_amp_max=0.3, bump_sigma_min=5.0, bump_sigma_max=30.0, n_channel=0.025, n_floodplain=0.06,
              inflow_q=15.0, fp_depth_out=1.5)
MEANDER = A.ChannelSpec(length=2000.0, width=540.0, meander_amp=90.0, meander_wavelength=600.0, meander_phase=150.0,
                        seed=5, **SCALED)
HYDRO = (np.array([0.0, 1000.0]), np.array([15.0, 180.0]))
DT, NSTEPS = 5.0, 40


def test_scaled_meander_hydrograph_fp64_matches_oracle(gpu):
    mesh, sol0 = A.channel_mesh(MEANDER)
    sm = A.SModelSuper(mesh, dt=DT)
    sm.set_initial(sol0)
    A.bind_hydrograph(sm, *HYDRO)
    n, st = A.time_loop(sm, 12)
    assert n == 12
    idx, w = A.inflow_segments(mesh)
    orc = O.Oracle(mesh, dt=DT, threads=8)
    orc.set_initial(sol0)
    val = mesh.seg_val.copy()
    its = 0
    for k in range(12):
        t1, t2 = k * DT, (k + 1) * DT
        qavg = quad(lambda s: np.interp(s, *HYDRO), t1, t2, epsabs=1e-14)[0] / DT
        val[idx, 0] = -qavg / w
        orc.set_seg_values(val)
        no, sto = orc.time_loop(1)
        assert no == 1
        its += sto["newton_its"]
    assert st["newton_its"] == its
    assert rel_err(sm.sol[0::3], orc.state()[0][0::3]) <= 1e-8


def test_scaled_meander_learned_warm_start(gpu):
    mesh, sol0 = A.channel_mesh(MEANDER)
    train = [A.channel_mesh(A.ChannelSpec(length=1000.0, width=300.0, seed=sd, **SCALED)) for sd in (11, 12, 13)]
    prep = lambda sm, i=0: A.bind_hydrograph(sm, *HYDRO)  # noqa: E731
    mlp, hist, _ = LR.train_warm_start(train, nsteps=NSTEPS, steps=4000, batch=8192, lr=2e-3, seed=3,
                                       precision=L.PREC_MIXED, dt=DT, prepare=prep, strict=False)
    assert hist[-50:].mean() < hist[:50].mean()
    runs = {}
    for name, warm, prec in (("fp64", None, L.PREC_FP64), ("mixed", None, L.PREC_MIXED),
                             ("learned", mlp, L.PREC_MIXED)):
        sm = A.SModelSuper(mesh, dt=DT)
        sm.set_initial(sol0)
        prep(sm)
        if warm is not None:
            LR.set_warm_start(sm, warm)
        n, st = A.time_loop(sm, NSTEPS, precision=prec)
        assert n == NSTEPS, name
        runs[name] = (sm.sol, st)
        if warm is not None:
            LR.set_warm_start(sm, None)
        sm.close()
    st_l, st_m = runs["learned"][1], runs["mixed"][1]
    assert st_l["warm_accepted"] >= NSTEPS // 2
    assert st_l["lin_its_total"] < st_m["lin_its_total"] and st_l["newton_its"] < st_m["newton_its"]
    assert st_l["resid_max_norm"] <= 1e-8 or st_l["inc_max_norm"] <= 1e-6
    assert rel_err(runs["learned"][0][0::3], runs["fp64"][0][0::3]) <= 1e-6
