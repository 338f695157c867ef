"""Learned warm start of the Newton iterate (host mirror of the adhk_mlp_* entry points).

A per-node MLP (12 -> 64 -> 64 -> 3, GELU, BF16 weights, FP32 arithmetic) predicts
the change of (h, u, v) over one time step.  Data collection, training (forward,
backward, Adam) and inference all run in libadhk.so's gfx950 kernels; this module
only moves handles and host arrays.  With a model attached to a super model,
fe_newton starts from the predicted iterate when its residual is smaller than the
previous level's, and otherwise from the previous level, as the reference does
(smodel_super_prep_sol.c:2-8); the Newton iteration and its 1e-8 test are unchanged.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from .sw2 import SModelSuper, time_loop

NPARAM = 5187
NNORM = 36


class WarmStartMLP:
    def __init__(self, seed=1):
        self.lib = L.load()
        h = self.lib.adhk_mlp_create(seed)
        if not h:
            raise L.AdhkError("adhk_mlp_create failed: " + L.last_error())
        self.h = C.c_void_p(h)
        assert self.lib.adhk_mlp_param_count() == NPARAM

    def close(self):
        if getattr(self, "h", None):
            self.lib.adhk_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- parameters ----
    def get_params(self):
        p, n = np.zeros(NPARAM, np.float32), np.zeros(NNORM, np.float32)
        L.check(self.lib.adhk_mlp_get_params(self.h, L.fptr(p), L.fptr(n)), "adhk_mlp_get_params")
        return p, n

    def set_params(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, np.float32)
        L.check(self.lib.adhk_mlp_set_params(self.h, L.fptr(p), L.fptr(n)), "adhk_mlp_set_params")

    # ---- data and training ----
    def reserve(self, capacity):
        L.check(self.lib.adhk_mlp_data_reserve(self.h, int(capacity)), "adhk_mlp_data_reserve")

    def collect(self, sm: SModelSuper):
        L.check(self.lib.adhk_mlp_collect(self.h, sm.h), "adhk_mlp_collect")

    @property
    def size(self):
        return int(self.lib.adhk_mlp_data_size(self.h))

    def fit_normalization(self):
        L.check(self.lib.adhk_mlp_fit_normalization(self.h), "adhk_mlp_fit_normalization")

    def train(self, steps, batch=8192, lr=1e-3, seed=7):
        hist = np.zeros(max(steps, 1))
        L.check(self.lib.adhk_mlp_train(self.h, int(steps), int(batch), float(lr), int(seed), L.dptr(hist)),
                "adhk_mlp_train")
        return hist[:steps]

    # ---- checking surface ----
    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((len(x), 3), np.float32)
        L.check(self.lib.adhk_mlp_forward(self.h, len(x), L.fptr(x), L.fptr(y)), "adhk_mlp_forward")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(NPARAM, np.float32)
        loss = C.c_double()
        L.check(self.lib.adhk_mlp_loss_grad(self.h, len(x), L.fptr(x), L.fptr(t), L.fptr(g), C.byref(loss)),
                "adhk_mlp_loss_grad")
        return loss.value, g


def features(sm: SModelSuper):
    """(nnodes, 12) float32 features of the super model's (sol_old, sol_older)."""
    f = np.zeros((sm.mesh.nnodes, 12), np.float32)
    L.check(sm.lib.adhk_sw2_features(sm.h, L.fptr(f)), "adhk_sw2_features")
    return f


def set_warm_start(sm: SModelSuper, mlp):
    """Attach a trained model as the Newton warm start of sm (None detaches)."""
    L.check(sm.lib.adhk_sw2_set_warm_start(sm.h, mlp.h if mlp is not None else None), "adhk_sw2_set_warm_start")
    sm._warm = mlp  # keep the model alive while attached


def collect_trajectory(sm: SModelSuper, mlp: WarmStartMLP, nsteps, precision=L.PREC_FP64, strict=True):
    """Advance sm by up to nsteps plain time steps, adding each converged step's (features, change) of
    every node.  A step that fails raises (strict).  Otherwise the failed step, which the time loop
    leaves at its old level with dt cut to a quarter, is taken as four quarter steps that are not
    collected, and the trajectory goes on at its own dt; a quarter step that fails ends the
    trajectory, keeping the steps already collected.  The step after the quarter steps is not
    collected either: its features would hold a quarter step's change as the previous change.
    A step that is not taken because a boundary series does not cover it (the time loop leaves dt as it
    was and runs no Newton iteration) ends the trajectory.
    Returns (Newton iterations, steps collected)."""
    its, k, dt0, resumed = 0, 0, sm.get_dt(), False
    for _ in range(nsteps):
        n, st = time_loop(sm, 1, precision=precision)
        its += st["newton_its"]
        if n == 1:
            if not resumed:
                mlp.collect(sm)
                k += 1
            resumed = False
            continue
        if sm.get_dt() == dt0:  # the forcing stopped the loop before any Newton iteration
            break
        if strict:
            raise L.AdhkError(f"time step failed while collecting: {st}")
        sm.set_dt(dt0 / 4)
        n, st = time_loop(sm, 4, precision=precision)
        its += st["newton_its"]
        if n != 4:
            break
        sm.set_dt(dt0)
        resumed = True
    return its, k


def train_warm_start(cases, nsteps, steps=2000, batch=8192, lr=2e-3, seed=1, precision=L.PREC_FP64, dt=1.0,
                     prepare=None, strict=True, validate=None, checkpoints=1, precond=L.PC_BJ):
    """Train a model on solver trajectories: cases = [(mesh, sol0), ...], nsteps steps each;
    prepare(sm, i), when given, sets up case i's model (e.g. its boundary time series) before it runs.
    strict=False steps over a failed step instead of raising, as collect_trajectory describes (the
    steps collected from each trajectory are in model.collected_steps).  With validate (model -> {"cost": ...}, e.g.
    validation_cost on a case held out of the training set), the steps run in `checkpoints` chunks
    and the chunk end with the lowest cost is kept (model.validation records the choice).

    precond is the BiCGStab preconditioner of the trajectories (ADHK_PC_*).

    Returns (model, loss history, data-generation Newton iterations)."""
    mlp = WarmStartMLP(seed=seed)
    mlp.reserve(sum(m.nnodes for m, _ in cases) * nsteps)
    mlp.collected_steps = []
    gen_its = 0
    for i, (mesh, sol0) in enumerate(cases):
        sm = SModelSuper(mesh, dt=dt)
        sm.solver.precond = precond
        sm.set_initial(sol0)
        if prepare is not None:
            prepare(sm, i)
        its, k = collect_trajectory(sm, mlp, nsteps, precision, strict=strict)
        gen_its += its
        mlp.collected_steps.append(k)
        sm.close()
    if mlp.size == 0:
        raise L.AdhkError("no time step converged while collecting training trajectories")
    mlp.fit_normalization()
    if validate is None:
        hist = mlp.train(steps, batch=batch, lr=lr, seed=seed + 100)
        return mlp, hist, gen_its
    # training in `checkpoints` chunks, each chunk's end a candidate; the candidate whose warm start costs
    # the held-out case the least solver work is kept (ties: the earlier one)
    per = max(1, steps // checkpoints)
    hist, snaps, costs = [], [], []
    for c in range(checkpoints):
        hist.append(mlp.train(per, batch=batch, lr=lr, seed=seed + 100 + c))
        snaps.append(mlp.get_params())
        costs.append(validate(mlp))
    best = min(range(checkpoints), key=lambda c: (costs[c]["cost"], c))
    mlp.set_params(*snaps[best])
    mlp.validation = {"checkpoints": [per * (c + 1) for c in range(checkpoints)], "costs": costs, "chosen": best}
    return mlp, np.concatenate(hist), gen_its


def validation_cost(mesh, sol0, nsteps, precision=L.PREC_FP64, dt=1.0, prepare=None, newton_weight=None,
                    precond=L.PC_BJ):
    """A validate() for train_warm_start: the solver work of nsteps time steps of a held-out case with the
    model as warm start, as BiCGStab iterations + newton_weight x Newton iterations (a Newton iteration
    also builds and folds the Jacobian, about as much work as 15 block-Jacobi or 5 multicolour-SGS
    BiCGStab iterations on the flume and channel workloads: the default for precond).  model=None gives
    the same case without a warm start."""
    if newton_weight is None:
        newton_weight = 5.0 if precond == L.PC_SGS else 15.0

    def cost(mlp):
        sm = SModelSuper(mesh, dt=dt)
        sm.solver.precond = precond
        sm.set_initial(sol0)
        if prepare is not None:
            prepare(sm)
        if mlp is not None:
            set_warm_start(sm, mlp)
        n, st = time_loop(sm, nsteps, precision=precision)
        if mlp is not None:
            set_warm_start(sm, None)
        sm.close()
        ok = n == nsteps
        return {"steps": n, "newton_its": st["newton_its"], "bicgstab_its": st["lin_its_total"],
                "warm_accepted": st["warm_accepted"],
                "cost": (st["lin_its_total"] + newton_weight * st["newton_its"]) if ok else float("inf")}
    return cost
