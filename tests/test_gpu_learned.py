"""GPU tests of the learned Newton warm start (libadhk.so's MLP kernels) against the oracle.

Bars: features to float32 rounding; forward outputs 1e-5 and loss / gradient 1e-4
relative to the oracle's max (FP32 arithmetic with BF16 weights, the hidden layer's
products on the matrix cores with BF16 activations, against float64 that rounds the
same operands to BF16).  An activation whose FP32 and float64 values fall on two sides
of a BF16 rounding boundary moves its sample's outputs by up to one BF16 step of that
activation times its weights: the forward bar holds for 99.9% of the samples, and every
sample is within 2e-3.  The matrix-core products are also checked against the oracle
with those roundings switched off, to the size of a BF16 rounding; training lowers the loss; the trained model cuts
Newton iterations on a held-out flume while the solution still meets the same
1e-8 Newton tolerance; a useless model is rejected every step and reproduces
plain Newton bit for bit.
"""
import numpy as np
import pytest

import adh_kraken_amd as A
import mlp_oracle as M
from adh_kraken_amd import learned as LR
from helpers import perturbed_state, rel_err

pytestmark = pytest.mark.gpu


def flume(seed, q=10.0, npx=201, npy=21):
    spec = A.FlumeSpec(xmin=0.0, xmax=float(npx - 1), ymin=0.0, ymax=float(npy - 1), npx=npx, npy=npy, slope=1e-4,
                       nbumps=5, bump_amp_max=0.3, bump_sigma_min=3.0, bump_sigma_max=10.0, seed=seed,
                       manning_n=0.03, inflow_q=q, tailwater=1.0, ic_flat_wse=1)
    return A.flume_mesh(spec)


def random_norm(rng):
    norm = np.zeros(36, np.float32)
    norm[0:12] = rng.uniform(-0.5, 0.5, 12)
    norm[12:24] = rng.uniform(0.5, 2.0, 12)
    norm[24:27] = rng.uniform(-0.1, 0.1, 3)
    norm[27:30] = rng.uniform(0.5, 2.0, 3)
    norm[30:33] = 1.0 / norm[27:30]
    return norm


def test_features_match_oracle(gpu):
    mesh, sol0 = A.flume_mesh(A.FLUME_C1)
    s, so = perturbed_state(sol0, seed=40)
    sm = A.SModelSuper(mesh)
    sm.sol_old, sm.sol_older = s, so
    fg = LR.features(sm).astype(np.float64)
    fo = M.features(mesh, s, so)
    for k in range(12):
        scale = max(np.abs(fo[:, k]).max(), 1e-30)
        assert np.abs(fg[:, k] - fo[:, k]).max() <= 2e-7 * scale, k


def test_forward_and_gradient_match_oracle(gpu):
    rng = np.random.default_rng(5)
    mlp = LR.WarmStartMLP(seed=3)
    norm = random_norm(rng)
    mlp.set_params(norm=norm)
    p, n2 = mlp.get_params()
    assert np.array_equal(n2, norm)
    x = rng.normal(size=(1000, 12)).astype(np.float32)  # 15 full tiles and a partial one
    t = rng.normal(size=(1000, 3)).astype(np.float32)
    y = mlp.forward(x)
    yo = M.forward(p, norm.astype(np.float64), x.astype(np.float64))
    per = np.abs(y - yo).max(axis=1) / np.abs(yo).max()
    assert np.quantile(per, 0.999) <= 1e-5 and per.max() <= 2e-3
    # the matrix-core rounding is real and of BF16 size: against the unrounded hidden layer the outputs
    # move by more than FP32 rounding and less than a few BF16 steps
    yu = M.forward(p, norm.astype(np.float64), x.astype(np.float64), mfma=False)
    assert 1e-6 < rel_err(y, yu) <= 3e-2
    loss, g = mlp.loss_grad(x, t)
    lo, go = M.loss_grad(p, norm.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - lo) <= 1e-5 * lo
    assert rel_err(g, go) <= 1e-4
    # every block of the parameter vector carries gradient
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.NPARAM)):
        assert np.abs(go[a:b]).max() > 0 and rel_err(g[a:b], go[a:b]) <= 1e-3


def test_training_lowers_the_loss(gpu):
    mesh, sol0 = flume(31)
    mlp = LR.WarmStartMLP(seed=4)
    mlp.reserve(12 * mesh.nnodes)
    sm = A.SModelSuper(mesh, dt=1.0)
    sm.set_initial(sol0)
    LR.collect_trajectory(sm, mlp, 12)
    assert mlp.size == 12 * mesh.nnodes
    mlp.fit_normalization()
    hist = mlp.train(400, batch=4096, lr=2e-3, seed=9)
    assert np.all(np.isfinite(hist))
    assert hist[-20:].mean() < 0.3 * hist[:5].mean()


@pytest.fixture(scope="module")
def trained():
    cases = [flume(s, q=q) for s, q in ((21, 9.0), (22, 10.0), (23, 11.0))]
    mlp, hist, _ = LR.train_warm_start(cases, nsteps=30, steps=1500, batch=8192, lr=2e-3, seed=5)
    return mlp, hist


def test_warm_start_cuts_newton_iterations(gpu, trained):
    mlp, hist = trained
    mesh, sol0 = flume(24, q=10.5)  # held out
    plain = A.SModelSuper(mesh, dt=1.0)
    plain.set_initial(sol0)
    n0, st0 = A.time_loop(plain, 30)
    warm = A.SModelSuper(mesh, dt=1.0)
    warm.set_initial(sol0)
    LR.set_warm_start(warm, mlp)
    n1, st1 = A.time_loop(warm, 30)
    assert n0 == n1 == 30
    assert st1["warm_accepted"] >= 20
    assert st1["newton_its"] < st0["newton_its"]
    assert st1["resid_max_norm"] <= 1e-8 or st1["inc_max_norm"] <= 1e-6
    # both solutions meet the same Newton tolerance: they agree to within it
    assert rel_err(warm.sol[0::3], plain.sol[0::3]) <= 1e-6


def test_useless_model_falls_back_to_plain_newton(gpu):
    mesh, sol0 = flume(25)
    mlp = LR.WarmStartMLP(seed=6)
    norm = np.zeros(36, np.float32)
    norm[12:24] = 1.0
    norm[24:27] = 5.0  # every prediction moves the state by 5 (m, m/s): never better than the old level
    norm[27:30] = 1.0
    norm[30:33] = 1.0
    mlp.set_params(norm=norm)
    a = A.SModelSuper(mesh, dt=1.0)
    a.set_initial(sol0)
    na, sta = A.time_loop(a, 5)
    b = A.SModelSuper(mesh, dt=1.0)
    b.set_initial(sol0)
    LR.set_warm_start(b, mlp)
    nb, stb = A.time_loop(b, 5)
    assert na == nb == 5 and stb["warm_accepted"] == 0
    assert sta["newton_its"] == stb["newton_its"] and sta["lin_its_total"] == stb["lin_its_total"]
    assert np.array_equal(a.sol, b.sol)


def test_failed_learned_start_is_solved_again_from_the_previous_level(gpu, trained):
    """Plain Newton stays the fallback: a solve that does not converge from the accepted learned start
    (here: one Newton iteration allowed) runs again from the previous level before dt is cut, so the
    step ends exactly as it would without the model (state restored, dt cut by 0.25).  The flume is
    advanced step by step until a step whose learned start is accepted."""
    mlp, _ = trained
    mesh, sol0 = flume(24, q=10.5)  # held out
    warm = A.SModelSuper(mesh, dt=1.0)
    warm.set_initial(sol0)
    LR.set_warm_start(warm, mlp)
    plain = A.SModelSuper(mesh, dt=1.0)
    for _ in range(20):
        snap = (warm.sol.copy(), warm.sol_old.copy(), warm.sol_older.copy())
        warm.max_nonlin_it = 1
        n1, st1 = A.time_loop(warm, 1)
        if st1["warm_retried"]:
            break
        # the start was rejected: back to the snapshot, then one ordinary step
        warm.sol, warm.sol_old, warm.sol_older = snap
        warm.set_dt(1.0)
        warm.max_nonlin_it = 10
        assert A.time_loop(warm, 1)[0] == 1
    else:
        pytest.fail("no learned start accepted in 20 steps")
    plain.sol, plain.sol_old, plain.sol_older = snap
    plain.max_nonlin_it = 1
    n0, st0 = A.time_loop(plain, 1)
    assert n0 == n1 == -1
    assert st1["warm_retried"] == 1 and st1["warm_accepted"] == 0
    assert st1["newton_its"] == st0["newton_its"] + 1
    assert warm.get_dt() == plain.get_dt() == 0.25
    assert np.array_equal(warm.sol, plain.sol)
