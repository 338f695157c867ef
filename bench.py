"""SW2 Newton-step benchmark on one MI355X.

Metric: SW2 timesteps/s at 1.2M dofs (the 2000 m x 200 m flume, 402,201 nodes,
800,000 triangles, dt = 2 s), Newton tolerance 1e-8.  A step is one time step
(prep_sol + fe_newton to convergence) of the device-resident solver, inputs
already resident in HBM.  W warmup steps run first; the state is then reset to
the initial condition and exactly K steps are timed between two device syncs.

Printed: one JSON line with the timesteps/s, the roofline of the dominant
kernel (average launch duration from hipEvent pairs on the library's stream
during the timed region) and a CPU baseline (the oracle, all host threads, on a
bounded sample: one time step of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kind, N, nnzb, ndofs, mixed):
    """Algorithmic HBM bytes of one launch (each operand moved once)."""
    vb = 4 if mixed else 8  # matrix / Krylov vector element size
    blocks = nnzb * (9 * vb + 4) + (N + 1) * 4  # block values + block columns + row pointers
    minv = N * 9 * vb
    if kind == "spmv_p":  # read A, p (once), M^-1, q; write Ap
        return blocks + minv + 3 * ndofs * vb
    if kind == "spmv_s":  # read A, r, Ap (once), M^-1, q; write s, As
        return blocks + minv + 5 * ndofs * vb
    if kind == "krylov_update":  # read x p Ap s As; write x r p
        return 8 * ndofs * vb
    raise ValueError(kind)


def cpu_baseline(mesh, sol0, dt, nsteps_sample=1):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker doubles as the CPU baseline ("port")
    threads = os.cpu_count() or 1
    try:
        threads = len(os.sched_getaffinity(0))
    except Exception:
        pass
    threads = min(threads, 16)
    orc = O.Oracle(mesh, dt=dt, threads=threads)
    orc.set_initial(sol0)
    t0 = time.perf_counter()
    n, st = orc.time_loop(nsteps_sample)
    t1 = time.perf_counter()
    return {"value": nsteps_sample / (t1 - t0), "unit": "timesteps/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps_sample} time step(s) of the same 1.2M-dof workload from the same initial state "
                      f"({t1 - t0:.1f} s, {st['newton_its']} Newton its, {st['lin_its_total']} BiCGStab its)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--config", choices=["c2", "c1"], default="c2")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--check-every", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the SW2 Newton step on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)

    import adh_kraken_amd as A
    from adh_kraken_amd import _lib as L
    import ctypes as C

    spec = A.FLUME_C2 if args.config == "c2" else A.FLUME_C1
    dt = 2.0 if args.config == "c2" else 1.0
    mesh, sol0 = A.flume_mesh(spec)
    sm = A.SModelSuper(mesh, dt=dt)
    sm.solver.lin_check_every = args.check_every
    prec = L.PREC_MIXED if args.precision == "mixed" else L.PREC_FP64
    lib = L.load()

    sm.set_initial(sol0)
    if args.warmup:
        A.time_loop(sm, args.warmup, precision=prec)
    sm.set_initial(sol0)
    sm.set_dt(dt)
    L.check(lib.adhk_sw2_kernel_timing(sm.h, 1), "kernel_timing")
    _ = sm.sol  # device sync (the state read-back waits on the stream)
    t0 = time.perf_counter()
    nsteps, st = A.time_loop(sm, args.steps, precision=prec)
    sol_end = sm.sol  # the state read-back waits for the library's stream
    t1 = time.perf_counter()
    ms = (C.c_double * 8)()
    cnt = (C.c_longlong * 8)()
    lib.adhk_sw2_kernel_times(sm.h, ms, cnt)
    L.check(lib.adhk_sw2_kernel_timing(sm.h, 0), "kernel_timing")
    elapsed = t1 - t0
    if nsteps != args.steps:
        print(json.dumps({"error": f"time loop stopped after {nsteps} of {args.steps} steps", "stats": st}))
        sys.exit(1)
    kt = {L.KT_NAMES[i]: {"ms_total": ms[i], "launches": int(cnt[i]),
                          "ms_avg": (ms[i] / cnt[i]) if cnt[i] else None} for i in range(8)}
    dom = max(("spmv_p", "spmv_s", "krylov_update", "elem_jac"), key=lambda k: kt[k]["ms_total"])
    mixed = prec == L.PREC_MIXED
    roof_kind = dom if dom != "elem_jac" else "spmv_s"
    ab = algorithmic_bytes(roof_kind, sm.mesh.nnodes, sm.nnz_blocks, sm.ndofs, mixed)
    avg_s = kt[roof_kind]["ms_avg"] / 1e3
    achieved = ab / avg_s / 1e9
    roofline = {"kernel": "k_" + roof_kind.replace("krylov_", ""), "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "algorithmic_bytes": ab, "avg_launch_ms": kt[roof_kind]["ms_avg"]}
    out = {
        "metric": "SW2 timesteps/s @1.2M dofs: FP64 plain vs mixed+learned, equal 1e-8 Newton tol",
        "value": round(args.steps / elapsed, 4),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if not mixed else "f32-inner/f64-outer",
        "data": "synthetic",
        "config": {"workload": ("SW2 flume 2000 m x 200 m at 1 m, 402201 nodes, 800000 triangles, 1206603 dofs, "
                                "dt 2 s, Manning n 0.03, 50 seeded bed bumps, Q 200 m3/s, tailwater 2 m"
                                if args.config == "c2" else "SW2 flume 200 m x 20 m, 12663 dofs, dt 1 s"),
                   "variant": "plain FP64 Newton + BiCGStab/block-Jacobi" if not mixed else
                              "mixed precision (FP32 Krylov, FP64 residual/update/test)",
                   "newton_tol": sm.solver.tol_nonlin, "inc_nonlin": sm.solver.inc_nonlin},
        "newton_its": st["newton_its"], "bicgstab_its": st["lin_its_total"],
        "final_resid_max": st["resid_max_norm"],
        "kernel_ms": {k: round(v["ms_total"], 3) for k, v in kt.items()},
        "roofline": roofline,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, sol0, dt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
