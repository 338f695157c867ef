"""SW2 Newton-step benchmark on one MI355X.

Metric: SW2 timesteps/s at 1.2M dofs (the 2000 m x 200 m flume, 402,201 nodes,
800,000 triangles, dt = 2 s), Newton tolerance 1e-8, plain FP64 against the
mixed-precision variant.  A step is one time step (prep_sol + fe_newton to
convergence) of the device-resident solver, inputs already resident in HBM.
For each variant W warmup steps run first; the state is then reset to the
initial condition and exactly K steps are timed between two device syncs.
``value`` is the plain FP64 rate; the mixed-precision rate, its speed-up and its
depth/velocity error against the FP64 result of the same K steps are reported
beside it.

A third variant adds the learned warm start to the mixed-precision solver: a per-node MLP trained
(untimed) on the GPU on 100-step trajectories of four other flumes of the same family (bump seeds
2-5), the training checkpoint kept being the one that costs a fifth held-out flume (seed 6) the
least solver work.

A fourth variant runs plain FP64 at the mixed variant's inner (BiCGStab) tolerance of 1e-6, so that the
mixed speed-up splits into what FP32 storage gives and what the looser inner solve gives.

The roofline is for the dominant kernel of the plain run: algorithmic bytes of
one launch divided by its average launch time, measured with hipEvent pairs on
the library's stream during the timed region.  The CPU baseline is the oracle
on the host threads this job may use (the affinity set, or OMP_NUM_THREADS when that names fewer) over
a bounded sample (the first five time steps of the same workload, median of three runs), with a
one-thread sample of the first step beside it; the line states the thread choice, the host's CPU model
and counts, and the sample's step count against the GPU's.

The ensemble leg runs the 64-member batched workload for its 50 steps, failed steps retaken in
sub-steps, against the same members solved one after another.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
KT_STRIDE = 16
T_START = time.perf_counter()


def log(msg):
    """Progress on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench {time.perf_counter() - T_START:7.1f}s] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(kind, N, nslot, ndofs, mixed, split=False):
    """HBM bytes of one launch with every operand moved once.

    matrix: the left-preconditioned M^-1 A in SELL-64 without its diagonal blocks (the identity, slot 0 of
    every slice, not read): (nslot - nslice)*64 slots x (9 values + 1 column index), slice pointers, and a
    byte per node that marks the identity diagonals; vectors: ndofs values each.
    """
    vb = 4 if mixed else 8
    nslice = (N + 63) // 64
    mat = (nslot - nslice) * 64 * (9 * vb + 4) + (nslice + 1) * 4 + N
    if kind == "spmv_p" and split:  # k_spmv_p: read p q, write Ap
        return mat + 3 * ndofs * vb
    if kind == "spmv_p":  # k_iter_a: read x s As p Ap q, write x r p' Ap'
        return mat + 10 * ndofs * vb
    if kind == "spmv_s":  # k_iter_s: read r Ap q, write s As
        return mat + 5 * ndofs * vb
    raise ValueError(kind)


def sgs_product_bytes(layout, counts, mixed):
    """HBM bytes of one multicolour SGS product (the S step: 2 nc - 2 colour-phase launches) with every operand
    moved once, and the matrix's share of them: the U and L parts of the folded matrix in SELL-64 over the
    permuted rows (slot columns x 64 x (9 values + 1 column index), slice pointers); r, Ap and q read and s and
    As written for every node; t written for the middle colours (1 .. nc - 2) and u for every colour but the
    top one (colour 0's t and the top colour's t and u stay in registers).  counts: nodes per colour."""
    vb = 4 if mixed else 8
    rows = layout["rows"]
    mat = (layout["u_slots"] + layout["l_slots"]) * 64 * (9 * vb + 4) + 2 * (rows // 64 + 1) * 4
    nodes, mid, nontop = int(sum(counts)), int(sum(counts[1:-1])), int(sum(counts[:-1]))
    return mat + 3 * vb * (5 * nodes + mid + nontop), mat


def host_cpu():
    """The host's processor model, logical CPU count, the affinity set and the OpenMP thread setting."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except Exception:
        affinity = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS")
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity": affinity,
            "omp_num_threads_env": int(omp) if omp and omp.isdigit() else None}


def cpu_baseline(mesh, sol0, dt, gpu_steps, nsteps_sample=5, seg_val=None, serial_steps=0, repeats=3):
    """The oracle over nsteps_sample steps on every host thread this process may use, the median of `repeats`
    runs; with serial_steps, also on one thread (the reference itself is single-threaded) over that many
    steps.  The threads are the affinity set, unless OMP_NUM_THREADS names fewer: that is the CPU share a
    one-GPU job gets on a shared host, and more threads than that would time other jobs' cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker doubles as the CPU baseline ("port")
    host = host_cpu()
    threads = host["affinity"]
    why = "every CPU of the affinity set"
    if host["omp_num_threads_env"] and host["omp_num_threads_env"] < threads:
        threads = host["omp_num_threads_env"]
        why = f"OMP_NUM_THREADS={threads}, this job's CPU share of the {host['affinity']}-CPU affinity set"

    def timed(nthreads, nsteps):
        orc = O.Oracle(mesh, dt=dt, threads=nthreads)
        orc.set_initial(sol0)
        if seg_val is not None:  # boundary values of the first step of a time-series forcing
            orc.set_seg_values(seg_val)
        t0 = time.perf_counter()
        n, st = orc.time_loop(nsteps)
        return time.perf_counter() - t0, st

    runs = [timed(threads, nsteps_sample) for _ in range(repeats)]
    dt_all, st = sorted(runs, key=lambda r: r[0])[len(runs) // 2]  # the median run
    out = {"value": round(nsteps_sample / dt_all, 5), "unit": "timesteps/s", "cores": threads, "kind": "port",
           "threads_chosen_as": why, "host": host,
           "sample": f"the first {nsteps_sample} of the {gpu_steps} time steps the GPU leg times, same "
                     f"{3 * mesh.nnodes}-dof workload and initial state, OpenMP C oracle, median of {repeats} runs "
                     f"({dt_all:.1f} s; runs {', '.join('%.1f' % r[0] for r in runs)} s; {st['newton_its']} Newton "
                     f"its, {st['lin_its_total']} BiCGStab its)",
           "sample_steps": nsteps_sample, "gpu_steps": gpu_steps}
    if serial_steps:
        dt_one, st1 = timed(1, serial_steps)
        out["serial"] = {"value": round(serial_steps / dt_one, 5), "unit": "timesteps/s", "cores": 1,
                         "sample": f"the first {serial_steps} time step(s) on one thread ({dt_one:.1f} s)"}
    return out


def run_variant(A, L, mesh, sol0, dt, precision, steps, warmup, check_every, warm=None, schedule=None, prepare=None,
                solver=None):
    lib = L.load()
    sm = A.SModelSuper(mesh, dt=dt)
    sm.solver.lin_check_every = check_every
    for k, v in (solver or {}).items():  # e.g. the inner (BiCGStab) tolerance
        setattr(sm.solver, k, v)
    if prepare is not None:  # e.g. the inflow hydrograph
        prepare(sm)
    if warm is not None:
        from adh_kraken_amd import learned as LR
        LR.set_warm_start(sm, warm)
    sm.set_initial(sol0)
    if schedule is None:  # schedule choices timed on this model's own system (they never change a result)
        log(f"autotune (precision {precision})")
        tuned, _ = sm.autotune(precision=precision, ninner=2)
    else:
        tuned = {**sm.tuning, **schedule}
    tuned["graphs"] = 0  # the timed region samples BiCGStab launches with event pairs, which need eager launches
    sm.tuning = tuned
    if warmup:
        A.time_loop(sm, warmup, precision=precision)
    sm.set_initial(sol0)
    sm.set_dt(dt)
    sm.time = 0.0
    # event pairs around the first and then every KT_STRIDE-th launch of each kernel class: timing every
    # launch would add two stream records per kernel (about 16% of the step on this workload)
    L.check(lib.adhk_sw2_kernel_timing_mask(sm.h, 0xFF, KT_STRIDE), "kernel_timing")
    log(f"timed run: {steps} steps, precision {precision}{', learned warm start' if warm is not None else ''}")
    L.check(lib.adhk_device_synchronize(), "device_synchronize")  # the device is idle when the clock starts
    t0 = time.perf_counter()
    nsteps, st = A.time_loop(sm, steps, precision=precision)
    L.check(lib.adhk_device_synchronize(), "device_synchronize")  # and when it stops
    t1 = time.perf_counter()
    sol_end = sm.sol
    ms = (C.c_double * 8)()
    cnt = (C.c_longlong * 8)()
    L.check(lib.adhk_sw2_kernel_times(sm.h, ms, cnt), "kernel_times")
    L.check(lib.adhk_sw2_kernel_timing(sm.h, 0), "kernel_timing")
    if nsteps != steps:
        raise RuntimeError(f"time loop stopped after {nsteps} of {steps} steps: {st}")
    kt = {L.KT_NAMES[i]: {"timed_launches": int(cnt[i]), "ms_avg": (ms[i] / cnt[i]) if cnt[i] else None}
          for i in range(8)}
    return {"sm": sm, "elapsed": t1 - t0, "stats": st, "kt": kt, "sol": sol_end, "tuning": tuned}


def pcie_leg(A, L, mesh, sol0, dt, steps=10):
    """The rate a caller sees through the reference-style per-call boundary: every time step copies the
    state (sol, sol_old, sol_older) in from pageable host arrays, runs one step and copies sol back out,
    as the fe_newton shim over the C-ABI does.  The same steps from the same state are also timed with
    the state resident, so the two rates differ by the copies alone; FP64, default schedule."""
    lib = L.load()
    sm = A.SModelSuper(mesh, dt=dt)
    sm.set_initial(sol0)
    A.time_loop(sm, 1)
    sm.set_initial(sol0)
    sm.set_dt(dt)
    L.check(lib.adhk_device_synchronize(), "device_synchronize")
    t0 = time.perf_counter()
    n, _ = A.time_loop(sm, steps)
    L.check(lib.adhk_device_synchronize(), "device_synchronize")
    t_res = time.perf_counter() - t0
    if n != steps:
        raise RuntimeError("time step failed in the resident half of the PCIe leg")
    s, so, soo = sol0.copy(), sol0.copy(), sol0.copy()
    sm.set_dt(dt)
    L.check(lib.adhk_device_synchronize(), "device_synchronize")
    t0 = time.perf_counter()
    for _ in range(steps):
        L.check(lib.adhk_sw2_set_state(sm.h, L.dptr(s), L.dptr(so), L.dptr(soo)), "set_state")
        n, _ = A.time_loop(sm, 1)
        if n != 1:
            raise RuntimeError("time step failed in the PCIe leg")
        soo, so = so, s
        s = np.zeros_like(s)
        L.check(lib.adhk_sw2_get_state(sm.h, L.dptr(s), None, None), "get_state")
    t_pcie = time.perf_counter() - t0
    sm.close()
    mb = 4 * sol0.size * 8 / 1e6
    return {"value": round(steps / t_pcie, 4), "unit": "timesteps/s", "steps": steps,
            "resident_same_steps": round(steps / t_res, 4),
            "copy_ms_per_step": round((t_pcie - t_res) / steps * 1e3, 3),
            "copied_mb_per_step": round(mb, 2),
            "note": "state copied host->device (3 x ndofs doubles, pageable) before and device->host "
                    "(ndofs doubles) after every step, FP64, default schedule; resident_same_steps times "
                    "the same steps from the same state without the copies"}


def ensemble_run(A, L, steps, substep_depth=3):
    """64 independent uneven flumes (length 200-4000 m, width 20-200 m, 5k-200k triangles, seeded Manning n,
    inflow, tailwater, bumps), mixed precision, dt 1 s, a failed member's step retaken in sub-steps down to
    dt / 4^substep_depth: one context with every member in each launch, against the same members solved one
    after another with the same settings."""
    cases = A.flume_ensemble(n=64, seed=4)
    e0 = np.concatenate([s for _, s, _ in cases])
    ens = A.SModelSuper.ensemble([m for m, _, _ in cases], dt=1.0)
    ens.solver.substep_depth = substep_depth
    ens.set_initial(e0)
    tuned, _ = ens.autotune(precision=L.PREC_MIXED, ninner=2)  # graphs are a choice here: no kernel timers
    ens.set_initial(e0)
    A.time_loop(ens, 1, precision=L.PREC_MIXED)  # warm-up
    ens.set_initial(e0)
    _ = ens.sol
    L.check(L.load().adhk_device_synchronize(), "device_synchronize")
    t0 = time.perf_counter()
    _, st = A.time_loop(ens, steps, precision=L.PREC_MIXED)
    L.check(L.load().adhk_device_synchronize(), "device_synchronize")
    t_batch = time.perf_counter() - t0
    sol_b = ens.sol
    alive, done = ens.instance_status()
    per = ens.instance_stats()
    log("ensemble: the same members one after another")
    t_seq, same, differ, seq_done = 0.0, 0, [], 0
    for i, (m, s0, _) in enumerate(cases):
        sm = A.SModelSuper(m, dt=1.0)
        sm.solver.substep_depth = substep_depth
        sm.tuning = dict(tuned)
        sm.set_initial(s0)
        A.time_loop(sm, 1, precision=L.PREC_MIXED)
        sm.set_initial(s0)
        sm.set_dt(1.0)
        _ = sm.sol
        L.check(L.load().adhk_device_synchronize(), "device_synchronize")
        t0 = time.perf_counter()
        n_alone, _ = A.time_loop(sm, steps, precision=L.PREC_MIXED)
        L.check(L.load().adhk_device_synchronize(), "device_synchronize")
        t_seq += time.perf_counter() - t0
        seq_done += max(n_alone, -n_alone)
        s1 = sm.sol
        if done[i] == steps:  # bit-identity is claimed for the members that completed every step
            if np.array_equal(s1, sol_b[ens.instance_slice(i)]):
                same += 1
            else:
                differ.append({"member": i, "max_abs_diff": float(np.nanmax(np.abs(s1 - sol_b[ens.instance_slice(i)])))})
        sm.close()
    complete = [i for i in range(len(cases)) if done[i] == steps]
    member_steps = int(done.sum())
    return {"members": len(cases), "triangles": int(sum(m.nelems2d for m, _, _ in cases)),
            "dofs": int(ens.ndofs), "steps": steps, "precision": "mixed", "substep_depth": substep_depth,
            "batched_ensemble_steps_per_s": round(steps / t_batch, 3),
            "batched_member_steps_per_s": round(member_steps / t_batch, 2),
            "sequential_ensemble_steps_per_s": round(steps / t_seq, 3),
            "sequential_member_steps_per_s": round(seq_done / t_seq, 2),
            "batched_speedup": round(t_seq / t_batch, 2),
            "members_completed_all_steps": len(complete),
            "members_recovered_by_substeps": [i for i in complete if per[i]["substeps"] > 0],
            "members_dropped": [{"member": i, "steps_done": int(done[i])} for i in range(len(cases)) if done[i] < steps],
            "bit_identical_to_standalone_of_completed": f"{same} of {len(complete)}",
            "members_differing": differ, "schedule": tuned, "newton_its": st["newton_its"],
            "bicgstab_its": st["lin_its_total"], "substeps": st["substeps"],
            "note": "member-steps/s counts the time steps every member completed; dropped members (a supercritical "
                    "inflow against a depth condition that fails at every dt down to 1/64 s, in the oracle too) "
                    "stop contributing after their last completed step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 50; c1 100, c3 20, c5 200)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", choices=["c2", "c1", "c3", "c5"], default="c2",
                    help="c2: the 1.2M-dof flume (the metric); c1: the 200 m flume; c3: the 10 km x 2 km "
                         "unstructured channel with floodplain (FP64 against mixed precision; no learned leg); "
                         "c5: the 8 km meandering channel under a rising hydrograph, learned warm start trained "
                         "on channels of the c3 family")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--train-steps", type=int, default=None,
                    help="time steps per training trajectory (default 100; c5 200)")
    ap.add_argument("--adam-steps", type=int, default=6000, help="optimizer steps of the warm-start training")
    ap.add_argument("--checkpoints", type=int, default=12,
                    help="training chunks whose ends are validated on a held-out case; the best is kept")
    ap.add_argument("--val-steps", type=int, default=30, help="time steps of the held-out validation case")
    ap.add_argument("--train-cases", type=int, default=4,
                    help="training meshes: flumes with bed-bump seeds 2, 3, ... (c1, c2) or channels of the "
                         "c3 family with seeds 11, 12, ... (c5)")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=50, help="time steps of the 64-member ensemble (the "
                    "workload's 50 by default)")
    ap.add_argument("--check-every", type=int, default=8)
    ap.add_argument("--precond", choices=["sgs", "bj"], default="sgs",
                    help="BiCGStab preconditioner of the timed variants: multicolour SGS (default; block-Jacobi "
                         "is then timed beside the FP64 variant) or block-Jacobi alone")
    ap.add_argument("--schedule", default=None,
                    help='fixed solver schedule as JSON, e.g. \'{"split": 1, "chunk": 8}\' (default: autotune)')
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the SW2 Newton step on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not implemented)"}))
        sys.exit(2)

    import adh_kraken_amd as A
    from adh_kraken_amd import _lib as L

    if args.steps is None:
        args.steps = {"c1": 100, "c2": 50, "c3": 20, "c5": 200}[args.config]
    if args.train_steps is None:
        args.train_steps = 200 if args.config == "c5" else 100
    prepare = None
    if args.config == "c5":
        t0 = time.perf_counter()
        spec = A.CHANNEL_C5
        mesh, sol0 = A.channel_mesh(spec)
        mesh_s = time.perf_counter() - t0
        dt = 5.0
        args.no_ensemble = True
        hydro = A.C5_HYDROGRAPH
        prepare = lambda sm: A.bind_hydrograph(sm, *hydro)  # noqa: E731
        workload = (f"SW2 meandering channel 8 km x 1.8 km (sine centreline, amplitude 300 m, wavelength 2 km), "
                    f"Poisson-disk points at 3 m, Delaunay, RCM order: {mesh.nnodes} nodes, {mesh.nelems2d} "
                    f"triangles, {3 * mesh.nnodes} dofs; 200 m channel 3 m deep, 50 seeded bumps, Manning n "
                    f"0.025 / 0.06; inflow hydrograph 50 -> 600 m3/s over 1000 s, dt 5 s "
                    f"(mesh built in {mesh_s:.1f} s, not timed)")
    elif args.config == "c3":
        t0 = time.perf_counter()
        spec = A.CHANNEL_C3
        mesh, sol0 = A.channel_mesh(spec)
        mesh_s = time.perf_counter() - t0
        dt = 5.0
        args.no_learned = args.no_ensemble = True
        workload = (f"SW2 channel with floodplain 10 km x 2 km, Poisson-disk points at 3 m minimum spacing, Delaunay, "
                    f"RCM order: {mesh.nnodes} nodes, {mesh.nelems2d} triangles, {3 * mesh.nnodes} dofs; 400 m "
                    f"channel 3 m deep, 50 seeded bumps, Manning n 0.025 / 0.06, Q 800 m3/s, dt 5 s "
                    f"(mesh built in {mesh_s:.1f} s, not timed)")
    else:
        spec = A.FLUME_C2 if args.config == "c2" else A.FLUME_C1
        dt = 2.0 if args.config == "c2" else 1.0
        mesh, sol0 = A.flume_mesh(spec)
        workload = ("SW2 flume 2000 m x 200 m at 1 m, 402201 nodes, 800000 triangles, 1206603 dofs, "
                    "dt 2 s, Manning n 0.03, 50 seeded bed bumps, Q 200 m3/s, tailwater 2 m"
                    if args.config == "c2" else "SW2 flume 200 m x 20 m, 12663 dofs, dt 1 s")

    log(f"workload ready: {mesh.nnodes} nodes, {mesh.nelems2d} triangles")
    schedule = json.loads(args.schedule) if args.schedule else None
    pc = L.PC_SGS if args.precond == "sgs" else L.PC_BJ
    traffic_tab = {}
    tpath = os.path.join(ROOT, "profiles", "sw2_newton_c2_traffic.json")
    if args.config == "c2" and os.path.exists(tpath):  # HBM bytes per launch from the rocprofv3 counter runs
        traffic_tab = json.load(open(tpath))["bytes_per_launch"]
    plain = run_variant(A, L, mesh, sol0, dt, L.PREC_FP64, args.steps, args.warmup, args.check_every,
                        schedule=schedule, prepare=prepare, solver={"precond": pc})
    sm = plain["sm"]
    probe_ms = C.c_double()
    nb = 1 << 30
    L.check(L.load().adhk_copy_probe(sm.h, nb, 10, C.byref(probe_ms)), "copy_probe")
    copy_gbs = 2 * nb / (probe_ms.value / 1e3) / 1e9

    def mall_note(mat_bytes):
        return (f"the {mat_bytes / 1e6:.0f} MB matrix fits the 256 MB MALL, which serves part of each product's "
                f"reads: the rate can exceed the 1 GiB copy kernel's, and is not HBM bandwidth alone"
                if mat_bytes < 256e6 else
                f"the {mat_bytes / 1e6:.0f} MB matrix exceeds the 256 MB MALL and streams from HBM")

    def roofline_bj(run):
        """the block-Jacobi solve's dominant kernel: the BiCGStab SpMV with the larger time per iteration"""
        kt = run["kt"]
        kind = max(("spmv_p", "spmv_s"), key=lambda k: kt[k]["ms_avg"] or 0.0)
        split = bool(run["tuning"] and run["tuning"].get("split"))
        ab = algorithmic_bytes(kind, sm.mesh.nnodes, sm.nslot, sm.ndofs, False, split)
        avg_ms = kt[kind]["ms_avg"]
        achieved = ab / (avg_ms / 1e3) / 1e9
        kname = {"spmv_p": "k_spmv_p" if split else "k_iter_a", "spmv_s": "k_iter_s"}[kind]
        return {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_tab.get(kname),
                "algorithmic_bytes": ab, "avg_launch_ms": round(avg_ms, 5), "copy_kernel_gbs": round(copy_gbs, 1),
                "frac_of_copy_kernel": round(achieved / copy_gbs, 3),
                "note": mall_note(algorithmic_bytes(kind, sm.mesh.nnodes, sm.nslot, 0, False, split))}

    def roofline_sgs(run):
        """the SGS solve's dominant unit: the S product, one pass over the folded matrix in 2 nc - 2 colour-phase
        launches of k_sgs_phase, timed by one event pair around the sequence"""
        col, nc, lay = run["sm"].colouring(layout=True)
        ab, mat = sgs_product_bytes(lay, np.bincount(col, minlength=nc), False)
        avg_ms = run["kt"]["spmv_s"]["ms_avg"]
        achieved = ab / (avg_ms / 1e3) / 1e9
        return {"kernel": f"k_sgs_phase<double, S> x {2 * nc - 2} (one SGS product)", "bound": "hbm",
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_tab.get("sgs_s_product"),
                "algorithmic_bytes": ab, "avg_launch_ms": round(avg_ms, 5), "copy_kernel_gbs": round(copy_gbs, 1),
                "frac_of_copy_kernel": round(achieved / copy_gbs, 3), "colours": nc, "launches": 2 * nc - 2,
                "note": mall_note(mat) + "; the event pair spans the launch gaps between the phases"}

    def pc_desc(p):
        return ("BiCGStab + multicolour symmetric Gauss-Seidel on the block-Jacobi-folded matrix, split form "
                "(Eisenstat products)" if p == L.PC_SGS else "BiCGStab + 3x3 block-Jacobi folded into the matrix")

    roofline = roofline_sgs(plain) if pc == L.PC_SGS else roofline_bj(plain)
    kt = plain["kt"]
    value = args.steps / plain["elapsed"]
    out = {
        "metric": ("SW2 timesteps/s @1.2M dofs: FP64 plain vs mixed+learned, equal 1e-8 Newton tol"
                   if args.config == "c2" else f"SW2 timesteps/s, {args.config} workload: FP64 plain vs mixed, "
                                               f"equal 1e-8 Newton tol"),
        "value": round(value, 4),
        "unit": "timesteps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * plain["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "variant": f"plain FP64: FD Jacobian, {pc_desc(pc)} (tol 1e-11, max 100)",
                   "newton_tol": sm.solver.tol_nonlin, "inc_nonlin": sm.solver.inc_nonlin},
        "newton_its": plain["stats"]["newton_its"], "bicgstab_its": plain["stats"]["lin_its_total"],
        "bicgstab_its_per_solve": round(plain["stats"]["lin_its_total"] / max(plain["stats"]["newton_its"], 1), 1),
        "bicgstab_its_max": plain["stats"]["lin_its_max"],
        "solves_at_iteration_cap": f"{plain['stats']['lin_capped']} of {plain['stats']['newton_its']} (cap 100)",
        "final_resid_max": plain["stats"]["resid_max_norm"],
        "kernel_avg_us": {k: round(1e3 * v["ms_avg"], 2) for k, v in kt.items() if v["ms_avg"]},
        "kernel_timing": f"hipEvent pairs on the first and every {KT_STRIDE}th launch of each kernel class",
        "schedule": plain["tuning"],
        "roofline": roofline,
    }
    if pc == L.PC_SGS:
        # the same FP64 variant with the block-Jacobi preconditioner
        bj = run_variant(A, L, mesh, sol0, dt, L.PREC_FP64, args.steps, args.warmup, args.check_every,
                         schedule=schedule, prepare=prepare, solver={"precond": L.PC_BJ})
        vbj = args.steps / bj["elapsed"]
        out["fp64_block_jacobi"] = {
            "value": round(vbj, 4), "unit": "timesteps/s", "ms_per_step": round(1e3 * bj["elapsed"] / args.steps, 3),
            "sgs_speedup_over_this": round(value / vbj, 3),
            "variant": f"plain FP64: FD Jacobian, {pc_desc(L.PC_BJ)} (tol 1e-11, max 100)",
            "depth_rel_err_vs_sgs": float(np.abs(bj["sol"][0::3] - plain["sol"][0::3]).max() /
                                          np.abs(plain["sol"][0::3]).max()),
            "newton_its": bj["stats"]["newton_its"], "bicgstab_its": bj["stats"]["lin_its_total"],
            "bicgstab_its_per_solve": round(bj["stats"]["lin_its_total"] / max(bj["stats"]["newton_its"], 1), 1),
            "solves_at_iteration_cap": bj["stats"]["lin_capped"],
            "kernel_avg_us": {k: round(1e3 * v["ms_avg"], 2) for k, v in bj["kt"].items() if v["ms_avg"]},
            "schedule": bj["tuning"], "roofline": roofline_bj(bj)}
    if not args.no_mixed:
        mx = run_variant(A, L, mesh, sol0, dt, L.PREC_MIXED, args.steps, args.warmup, args.check_every,
                         schedule=schedule, prepare=prepare, solver={"precond": pc})
        h64, hmx = plain["sol"][0::3], mx["sol"][0::3]
        u64 = np.hypot(plain["sol"][1::3], plain["sol"][2::3])
        umx = np.hypot(mx["sol"][1::3], mx["sol"][2::3])
        vmx = args.steps / mx["elapsed"]
        out["mixed"] = {"value": round(vmx, 4), "unit": "timesteps/s", "speedup_vs_fp64": round(vmx / value, 3),
                        "variant": "FP32 Jacobian storage, SpMV and BiCGStab (inner tol 1e-6); FP64 residual, "
                                   "Newton update and convergence test; " + pc_desc(pc),
                        "depth_rel_err_vs_fp64": float(np.abs(hmx - h64).max() / np.abs(h64).max()),
                        "speed_rel_err_vs_fp64": float(np.abs(umx - u64).max() / max(np.abs(u64).max(), 1e-300)),
                        "newton_its": mx["stats"]["newton_its"], "bicgstab_its": mx["stats"]["lin_its_total"],
                        "bicgstab_its_max": mx["stats"]["lin_its_max"], "solves_at_iteration_cap": mx["stats"]["lin_capped"],
                        "kernel_avg_us": {k: round(1e3 * v["ms_avg"], 2) for k, v in mx["kt"].items() if v["ms_avg"]}}
        # FP64 throughout at the mixed variant's inner tolerance: splits the mixed speed-up into what the FP32
        # storage gives and what the looser inner solve gives
        f6 = run_variant(A, L, mesh, sol0, dt, L.PREC_FP64, args.steps, args.warmup, args.check_every,
                         schedule=schedule, prepare=prepare, solver={"lin_tol": 1e-6, "precond": pc})
        v6 = args.steps / f6["elapsed"]
        out["fp64_inner_tol_1e-6"] = {
            "value": round(v6, 4), "unit": "timesteps/s", "speedup_vs_fp64": round(v6 / value, 3),
            "mixed_speedup_over_this": round(vmx / v6, 3),
            "variant": "plain FP64 with the inner BiCGStab tolerance of the mixed variant (1e-6); Newton tolerance 1e-8",
            "depth_rel_err_vs_fp64": float(np.abs(f6["sol"][0::3] - h64).max() / np.abs(h64).max()),
            "newton_its": f6["stats"]["newton_its"], "bicgstab_its": f6["stats"]["lin_its_total"]}
    if not args.no_learned:
        from adh_kraken_amd import learned as LR
        t0 = time.perf_counter()
        if args.config == "c5":
            # trajectories of straight channels of the c3 family (2 km x 0.5 km, 100 m channel, seeds 11,
            # 12, ...) under the same hydrograph per unit channel width (25 -> 300 m3/s), mixed precision
            specs = [A.ChannelSpec(length=2000.0, width=500.0, spacing=3.0, chan_width=100.0, chan_depth=3.0,
                                   bank_width=20.0, slope=1e-4, nbumps=10, seed=sd, n_channel=0.025,
                                   n_floodplain=0.06, inflow_q=25.0, fp_depth_out=1.5)
                     for sd in range(11, 12 + args.train_cases)]  # the last one is held out for validation
            cases = [A.channel_mesh(ts) for ts in specs]
            tq = (hydro[0], hydro[1] * 100.0 / 200.0)
            tr_prepare = lambda sm_, i: A.bind_hydrograph(sm_, *tq)  # noqa: E731
            val_prepare = lambda sm_: A.bind_hydrograph(sm_, *tq)  # noqa: E731
            tr_desc = "%d channels of the c3 family (2 km x 0.5 km, seeds %s) x %d steps" % (
                args.train_cases, ", ".join(str(11 + i) for i in range(args.train_cases)), args.train_steps)
            val_desc = "channel seed %d, %d steps" % (11 + args.train_cases, args.val_steps)
        else:
            # trajectories of other flumes of the same family (bed-bump seeds 2, 3, ...), mixed precision
            seeds = tuple(range(2, 3 + args.train_cases))  # the last one is held out for validation
            cases = [A.flume_mesh(A.FlumeSpec(**{**spec.__dict__, "seed": sd})) for sd in seeds]
            tr_prepare = val_prepare = None
            tr_desc = "%d flumes x %d steps (bump seeds %s)" % (args.train_cases, args.train_steps,
                                                                ", ".join(str(sd) for sd in seeds[:-1]))
            val_desc = "flume bump seed %d, %d steps" % (seeds[-1], args.val_steps)
        log("learned warm start: trajectories, training and checkpoint validation")
        val = LR.validation_cost(*cases[-1], args.val_steps, precision=L.PREC_MIXED, dt=dt, prepare=val_prepare,
                                 precond=pc)
        mlp, hist, gen_its = LR.train_warm_start(cases[:-1], nsteps=args.train_steps, steps=args.adam_steps,
                                                 batch=16384, lr=2e-3, seed=5, precision=L.PREC_MIXED, dt=dt,
                                                 prepare=tr_prepare, strict=False, validate=val,
                                                 checkpoints=args.checkpoints, precond=pc)
        val_plain = val(None)
        t_train = time.perf_counter() - t0
        ml = run_variant(A, L, mesh, sol0, dt, L.PREC_MIXED, args.steps, args.warmup, args.check_every, warm=mlp,
                         schedule=schedule, prepare=prepare, solver={"precond": pc})
        h64, hml = plain["sol"][0::3], ml["sol"][0::3]
        vml = args.steps / ml["elapsed"]
        out["mixed_learned"] = {
            "value": round(vml, 4), "unit": "timesteps/s", "speedup_vs_fp64": round(vml / value, 3),
            "variant": "mixed precision + learned per-node MLP warm start (12-64-64-3 GELU, BF16 weights); the "
                       "start with the smaller l2-norm residual of {learned, previous level} is kept; " + pc_desc(pc),
            "warm_accepted_steps": ml["stats"]["warm_accepted"], "warm_retried_steps": ml["stats"]["warm_retried"],
            "newton_its": ml["stats"]["newton_its"],
            "bicgstab_its": ml["stats"]["lin_its_total"],
            "depth_rel_err_vs_fp64": float(np.abs(hml - h64).max() / np.abs(h64).max()),
            "plain_mixed_newton_its": out["mixed"]["newton_its"] if "mixed" in out else None,
            "speedup_vs_plain_mixed": round(vml / out["mixed"]["value"], 3) if "mixed" in out else None,
            # steps after which the untimed data generation, training and validation are paid back by the
            # time each step saves against plain mixed precision
            "break_even_steps": (round(t_train / (1.0 / out["mixed"]["value"] - 1.0 / vml))
                                 if "mixed" in out and vml > out["mixed"]["value"] else None),
            "training": {"trajectories": tr_desc, "steps_collected": mlp.collected_steps,
                         "samples": mlp.size, "adam_steps": args.adam_steps, "batch": 16384,
                         "loss_first": float(hist[:10].mean()), "loss_last": float(hist[-10:].mean()),
                         "seconds_incl_data_generation_and_validation": round(t_train, 2)},
            "validation": {"case": val_desc,
                           "weight": f"BiCGStab its + {5 if pc == L.PC_SGS else 15} x Newton its",
                           "checkpoints": mlp.validation["checkpoints"], "chosen": mlp.validation["chosen"],
                           "costs": [c["cost"] for c in mlp.validation["costs"]], "cost_without_model": val_plain["cost"],
                           "chosen_its": {k: mlp.validation["costs"][mlp.validation["chosen"]][k]
                                          for k in ("newton_its", "bicgstab_its", "warm_accepted")},
                           "its_without_model": {k: val_plain[k] for k in ("newton_its", "bicgstab_its")}}}
    if args.config == "c2":
        log("PCIe-inclusive leg")
        out["pcie_inclusive"] = pcie_leg(A, L, mesh, sol0, dt)
    if not args.no_ensemble:
        log("ensemble")
        out["ensemble"] = ensemble_run(A, L, args.ensemble_steps)
    if not args.no_cpu:
        seg_val = None
        if args.config == "c5":
            idx, w_in = A.inflow_segments(mesh)
            seg_val = mesh.seg_val.copy()
            seg_val[idx, 0] = -A.series_average(hydro[0], hydro[1], 0.0, dt) / w_in
        log("CPU baseline (oracle)")
        big = args.config in ("c3", "c5")
        cb = cpu_baseline(mesh, sol0, dt, args.steps, nsteps_sample=1 if big else 5, seg_val=seg_val,
                          serial_steps=0 if big else 1, repeats=1 if big else 3)
        cb["gpu_fp64_speedup"] = round(value / cb["value"], 1)
        if "serial" in cb:
            cb["serial"]["gpu_fp64_speedup"] = round(value / cb["serial"]["value"], 1)
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
