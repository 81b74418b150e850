"""Benchmark of the Nyles time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: Taylor-Green vortex on a triply periodic 256^3 grid (the grid the
headline metric is quoted on), LES model with RK3_SSP stepping, WENO5
vortex force and tracer advection, mgfor-equivalent multigrid projection,
fixed dt = 0.5 / max|U| at t = 0.  One step = one full RK3 step (3
right-hand sides + 3 projections); one cell-step = one interior cell advanced
by one step.  The state is resident in HBM before the timed region starts.

The headline value is the mixed-precision path (FP32 stencils and inner
V-cycles, FP64 outer residual); the FP64 path (bit-identical to the oracle)
is timed in the same run, and both are advanced --err-steps steps from the
same initial state to state the error of the mixed path against FP64 (KE,
velocity, buoyancy).

Beside the headline, the same run measures an ensemble of 64 members of
uneven sizes advanced by batched steps against the same members stepped one
after another, a serving session (Poisson job arrivals, continuous batching)
reported as p50/p99 time-to-completion and throughput, and a Rayleigh-Taylor
256 x 256 x 512 run with the projection converged to |r|/|b| <= 1e-10 in
both precisions.

Prints ONE JSON line.  ``roofline`` prices the dominant kernel of the
headline path, the two-sweep damped-Jacobi smoothing pass on the finest
multigrid level (k_mg_sweep2t: both pre-smoothing sweeps with the residual
norm of the convergence test, or both post-smoothing sweeps; the kernel with
the largest time per step): achieved = its algorithmic HBM bytes per launch
(each input field read once, each output written once: 12 bytes per cell in
FP32 for two sweeps) divided by its mean launch time, both measured with HIP
events on the library's stream over the timed steps.  ``cpu_baseline`` times the
FP64 CPU oracle (C restatement of the reference kernels + numpy
orchestration, OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
METRIC = "Mcell-steps/s at 256³ mixed vs FP64 (same err); % HBM peak; p99 job latency"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256, help="grid cells per axis")
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="path of the headline value (the other one is timed beside it)")
    ap.add_argument("--roofline-kernel", default="mg_sweep2_fine",
                    help="kernel class priced against the roofline (default: the two-sweep "
                         "smoothing pass on the finest level, the kernel with the most time "
                         "per step)")
    ap.add_argument("--err-steps", type=int, default=200,
                    help="steps of the mixed-vs-FP64 error check (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="RK3 steps of the CPU baseline on all host cores")
    ap.add_argument("--cpu-steps-1thread", type=int, default=1,
                    help="RK3 steps of the CPU baseline on one thread (0: skip)")
    ap.add_argument("--breakdown", action="store_true",
                    help="report the per-class kernel times of both paths")
    ap.add_argument("--ensemble-members", type=int, default=64,
                    help="members of the ensemble run (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=300,
                    help="batched steps of the ensemble run")
    ap.add_argument("--ensemble-seq-steps", type=int, default=10,
                    help="steps of the one-member-after-another reference point")
    ap.add_argument("--serving-jobs", type=int, default=48,
                    help="jobs of the serving run (the first arrivals of the stream; 0: skip)")
    ap.add_argument("--serving-rate", type=float, default=1.0, help="job arrivals per second")
    ap.add_argument("--rt-steps", type=int, default=200,
                    help="steps of the Rayleigh-Taylor 256x256x512 run at |r|/|b| <= 1e-10 "
                         "(0: skip)")
    ap.add_argument("--single", action="store_true",
                    help="time only the headline precision (short runs under a profiler)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="only copy this many MB device to device 5 times (counter calibration)")
    return ap.parse_args()


def copy_calibration(mb):
    """A plain device-to-device copy of known size, to scale HBM byte counters."""
    import torch
    n = mb * (1 << 20)
    a = torch.ones(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n, "copy_calibration_bytes_written": n,
                      "launches": 5}), flush=True)


def c4_sizes(members, seed):
    """N log-uniform in [32, 96], rounded to a multiple of 8"""
    import numpy as np
    rng = np.random.default_rng(seed)
    return [int(8 * round(v / 8)) for v in np.exp(rng.uniform(np.log(32), np.log(96), members))]


def ensemble_run(members, steps, seq_steps, precision="mixed", seed=1000):
    """Ensemble of stratified shear layers (perio_xy) with N^3 cells, N
    log-uniform in [32, 96] rounded to a multiple of 8: the members advanced
    by batched steps (one set of launches per RK stage, hipGraph replays,
    autotuned stage tile), and the same members stepped one after another
    through the single-model API (over seq_steps steps) as the reference
    point."""
    import torch
    from nyles_amd import serving
    from nyles_amd.device import DeviceBatch, tile_name
    sizes = c4_sizes(members, seed)
    jobs = [serving.Job("shear", n, steps, seed=seed + q) for q, n in enumerate(sizes)]
    cells = sum(n ** 3 for n in sizes)
    out = {"members": members,
           "sizes": {str(n): sizes.count(n) for n in sorted(set(sizes))},
           "cells": cells, "precision": precision, "steps": steps,
           "sequential_steps": seq_steps,
           "workload": "stratified shear layer perio_xy (tanh jet, N^2 = 0.1, noise 1e-3), "
                       "N log-uniform in [32, 96] rounded to 8, fixed dt at CFL 0.5"}
    for mode in ("sequential", "batched"):
        ms = [serving.make_model(j, precision) for j in jobs]
        models = [m for m, _ in ms]
        dts = [dt for _, dt in ms]
        nsteps = steps if mode == "batched" else seq_steps
        nite = 0
        if mode == "batched":
            b = DeviceBatch(models)
            out["tile"], out["tune_ms"] = tile_name(b.tile), b.tune_ms
            b.step_rk3(dts)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(nsteps):
                nite += sum(s["nite"] for s in b.step_rk3(dts))
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            b.close()
            out["vcycles_per_member_step"] = round(nite / (nsteps * members), 3)
        else:
            for m, dt in zip(models, dts):
                m.step_rk3(dt)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(nsteps):
                for m, dt in zip(models, dts):
                    m.step_rk3(dt)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
        for m in models:
            m.close()
        out[mode + "_Mcell_steps_per_s"] = round(cells * nsteps / el / 1e6, 3)
        out[mode + "_s"] = round(el, 3)
    out["batching_speedup"] = round(out["batched_Mcell_steps_per_s"] /
                                    out["sequential_Mcell_steps_per_s"], 3)
    return out


def state_errors(m, a):
    """errors of the mixed model m against the FP64 model a: KE (evaluated in
    FP64 by both), velocity and buoyancy (relative L2)"""
    import numpy as np
    from nyles_amd import _lib as L
    fm = {k: m.download(f) for k, f in (("b", L.NYL_F_B), ("u", L.NYL_F_U), ("v", L.NYL_F_V),
                                        ("w", L.NYL_F_W))}
    fa = {k: a.download(f) for k, f in (("b", L.NYL_F_B), ("u", L.NYL_F_U), ("v", L.NYL_F_V),
                                        ("w", L.NYL_F_W))}
    kem, kea = m.reduce(L.NYL_RED_KE), a.reduce(L.NYL_RED_KE)
    du2 = sum(float(np.sum((fm[k] - fa[k]) ** 2)) for k in ("u", "v", "w"))
    u2 = sum(float(np.sum(fa[k] ** 2)) for k in ("u", "v", "w"))
    bn = float(np.sqrt(np.sum(fa["b"] ** 2)))
    return {
        "ke_fp64": kea, "ke_mixed": kem,
        "ke_rel_err": abs(kem - kea) / abs(kea) if kea != 0 else None,
        "u_l2_rel_err": (du2 / u2) ** 0.5 if u2 > 0 else None,
        "b_max_abs_err": float(np.max(np.abs(fm["b"] - fa["b"]))),
        "b_l2_rel_err": (float(np.sqrt(np.sum((fm["b"] - fa["b"]) ** 2))) / bn) if bn > 0
        else None,
        "nonfinite_mixed": m.reduce(L.NYL_RED_NONFINITE),
    }


def rt_run(steps, tol2=1e-20, maxite=50):
    """Rayleigh-Taylor 256 x 256 x 512 (perio_xy, heavy fluid on top), the
    projection converged to |r|/|b| <= 1e-10 (unsquared; tol 1e-20 on mgfor's
    squared ratio): mixed (FP32 stencils, FP64 outer defect correction) and
    FP64 side by side from the same state, each timed over `steps` steps (one
    untimed step first), then the mixed state compared with FP64 after the
    same steps."""
    import numpy as np
    import torch
    from nyles_amd import _lib as L
    from nyles_amd import workloads
    from nyles_amd.device import DeviceModel
    st, meta = workloads.rayleigh_taylor()
    (Lx, Ly, Lz), (nz, ny, nx) = meta["L"], st["b"].shape
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    dt = 0.5 * dz / np.sqrt(1.0 * Lz)          # buoyancy jump 1, depth Lz
    out = {"grid": [nx, ny, nz], "steps": steps + 1, "timed_steps": steps, "dt": dt,
           "tolerance": "|r|/|b| <= 1e-10 (unsquared), at most %d V-cycles" % maxite}
    models = {}
    for prec, code in (("fp64", L.NYL_PREC_FP64), ("mixed", L.NYL_PREC_MIXED)):
        m = DeviceModel(nx, ny, nz, (1, 1, 0), dx, dy, dz, idx2=1 / dx ** 2, idy2=1 / dy ** 2,
                        idz2=1 / dz ** 2, dx2=dx ** 2, precision=code, mg_tol=tol2,
                        mg_maxite=maxite)
        for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                       ("u_k", L.NYL_F_W)):
            m.upload(fid, st[k])
        m.diagnose()
        m.step_rk3(dt)
        torch.cuda.synchronize()
        nite, res = 0, 0.0
        t0 = time.perf_counter()
        for _ in range(steps):
            s = m.step_rk3(dt)
            nite += s["nite"]
            res = max(res, s["res"])
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[prec] = {"Mcell_steps_per_s": round(nx * ny * nz * steps / el / 1e6, 3),
                     "vcycles_per_step": round(nite / steps, 2),
                     "max_final_residual_unsquared": float(np.sqrt(res))}
        models[prec] = m
    out["speedup_mixed_vs_fp64"] = round(out["mixed"]["Mcell_steps_per_s"] /
                                         out["fp64"]["Mcell_steps_per_s"], 3)
    out["error_vs_fp64"] = state_errors(models["mixed"], models["fp64"])
    for m in models.values():
        m.close()
    return out


def lock_exchange_error(steps):
    """Lock exchange 64 x 32 x 32 (perio_x, walls in y and z): mixed and FP64
    advanced `steps` RK3 steps at dt = 0.5 dx / sqrt(db Lz) from the same
    state; errors of the mixed state against FP64."""
    import numpy as np
    from nyles_amd import _lib as L
    from nyles_amd import workloads
    from nyles_amd.device import DeviceModel
    st, meta = workloads.lock_exchange()
    (Lx, Ly, Lz), (nz, ny, nx) = meta["L"], st["b"].shape
    dx, dy, dz = Lx / nx, Ly / ny, Lz / nz
    dt = 0.5 * dx / np.sqrt(1.0 * Lz)
    models = {}
    for prec, code in (("fp64", L.NYL_PREC_FP64), ("mixed", L.NYL_PREC_MIXED)):
        m = DeviceModel(nx, ny, nz, (1, 0, 0), dx, dy, dz, idx2=1 / dx ** 2, idy2=1 / dy ** 2,
                        idz2=1 / dz ** 2, dx2=dx ** 2, precision=code)
        for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                       ("u_k", L.NYL_F_W)):
            m.upload(fid, st[k])
        m.diagnose()
        for _ in range(steps):
            m.step_rk3(dt)
        models[prec] = m
    out = {"grid": [nx, ny, nz], "steps": steps, "dt": dt}
    out.update(state_errors(models["mixed"], models["fp64"]))
    for m in models.values():
        m.close()
    return out


def serving_run(count, rate, precision="mixed"):
    """Poisson arrivals of lock-exchange, Taylor-Green and shear-layer jobs
    (kind uniform), n log-uniform in [32, 256] rounded to a multiple of 8,
    steps uniform in [50, 500], served with continuous batching: the first
    `count` arrivals of the serving stream."""
    from nyles_amd import serving
    jobs = serving.poisson_jobs(count, rate, size_range=(32, 256), steps=(50, 500))
    rep = serving.Server(precision=precision).run(jobs)
    rep["arrival_rate_per_s"] = rate
    rep["sizes"] = sorted(j.n for j in jobs)
    rep["cell_steps_requested"] = int(sum(j.steps * j.cells for j in jobs if j.cells))
    rep["workload"] = ("the first %d arrivals of a Poisson stream at %.2f/s; kind uniform over "
                       "lock exchange (n x n/2 x n/2), Taylor-Green (n^3), shear layer (n^3); "
                       "n log-uniform in [32, 256] rounded to 8; steps uniform in [50, 500]"
                       % (count, rate))
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in rep.items()}


def measured_traffic(kclass, precision, n):
    """HBM bytes per launch of the roofline kernel, from the committed
    rocprofv3 --pmc summary of the same kernel and workload (FETCH_SIZE and
    WRITE_SIZE passes scaled by a copy of known size), when that summary was
    measured on the current kernel sources (its recorded source digest
    matches); None otherwise."""
    from nyles_amd import profiling
    if (precision, n) != ("mixed", 256):
        return None
    path = os.path.join(ROOT, "profiles", "traffic_%s_taylor_green_256_mixed.md" % kclass)
    found, digest = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    found = int(line.split("=")[1])
                elif line.startswith("source_digest"):
                    digest = line.split("=")[1].strip()
    except OSError:
        return None
    if digest != profiling.source_digest():
        return None
    return found


def cpu_baseline(n, steps, steps1):
    """FP64 oracle on the host cores: initial projection untimed, then `steps`
    RK3 steps timed on all the cores of the process, then `steps1` steps on
    one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    from nyles_amd import workloads
    O.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    st, meta = workloads.taylor_green(n, seed=2)
    L = meta["L"]
    orc = O.OracleLES(n, n, n, L[0], L[1], L[2], "perio_xyz")
    orc.set_interior(orc.b, st["b"])
    orc.set_interior(orc.u["i"], st["u_i"])
    orc.set_interior(orc.u["j"], st["u_j"])
    orc.set_interior(orc.u["k"], st["u_k"])
    O.set_threads(cores)
    orc.diagnose_var()
    U = orc.interior(orc.U["i"]), orc.interior(orc.U["j"]), orc.interior(orc.U["k"])
    dt = 0.5 / float(np.sqrt(np.max(U[0] ** 2 + U[1] ** 2 + U[2] ** 2)))
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.rk3_ssp(dt)
    el = time.perf_counter() - t0
    threads = O.num_threads()
    out = {"value": round(n ** 3 * steps / el / 1e6, 4), "unit": "Mcell-steps/s",
           "cores": threads, "kind": "port",
           "sample": "taylor-green %d^3 fp64, %d RK3 step(s) after the initial projection, "
                     "%.1f s wall on %d OpenMP threads" % (n, steps, el, threads)}
    if steps1 > 0:
        O.set_threads(1)
        t0 = time.perf_counter()
        for _ in range(steps1):
            orc.rk3_ssp(dt)
        el1 = time.perf_counter() - t0
        O.set_threads(cores)
        out["single_thread"] = {
            "value": round(n ** 3 * steps1 / el1 / 1e6, 4), "unit": "Mcell-steps/s", "cores": 1,
            "sample": "the same run, %d further RK3 step(s) on 1 thread, %.1f s wall"
                      % (steps1, el1)}
    return out


def make_model(n, precision):
    import numpy as np
    from nyles_amd import _lib as L
    from nyles_amd import workloads
    from nyles_amd.device import DeviceModel
    st, meta = workloads.taylor_green(n, seed=2)
    Lx, Ly, Lz = meta["L"]
    dx, dy, dz = Lx / n, Ly / n, Lz / n
    prec = {"fp64": L.NYL_PREC_FP64, "mixed": L.NYL_PREC_MIXED}[precision]
    dev = DeviceModel(n, n, n, (1, 1, 1), dx, dy, dz, idx2=1 / dx ** 2, idy2=1 / dy ** 2,
                      idz2=1 / dz ** 2, dx2=dx ** 2, precision=prec)
    for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                   ("u_k", L.NYL_F_W)):
        dev.upload(fid, st[k])
    dev.diagnose()
    return dev


def class_times(dev, dt):
    """kernel ms, launches and algorithmic GB/s per class over one step"""
    from nyles_amd import _lib as L
    br = {}
    for kc in L.KERNEL_CLASSES:
        dev.prof_start(kc)
        dev.step_rk3(dt)
        ms, nl, by = dev.prof_stop()
        br[kc] = {"ms_per_step": round(ms, 4), "launches": nl,
                  "GB/s": round(by / (ms / 1e3) / 1e9, 1) if ms > 0 else None}
    return br


def timed(dev, dt, steps, warmup, kclass):
    import torch
    for _ in range(warmup):
        dev.step_rk3(dt)
    torch.cuda.synchronize()
    dev.prof_start(kclass)
    nite = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        nite += dev.step_rk3(dt)["nite"]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kms, klaunch, kbytes = dev.prof_stop()
    return el, nite, (kms, klaunch, kbytes)


def measured_valu(precision, n):
    """VALU issue of the stage kernel as a fraction of the SIMD's ceiling
    (0.25 wave64 instructions per cycle: FP32 unpacked and FP64 alike), mean
    over RK modes 1-3, from the committed rocprofv3 --pmc summary of the same
    workload when it was measured on the current kernel sources; None
    otherwise."""
    from nyles_amd import profiling
    if n != 256:
        return None
    path = os.path.join(ROOT, "profiles", "valu_stage_taylor_green_256.md")
    found, digest = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("valu_issue_frac_%s " % precision):
                    found = float(line.split("=")[1])
                elif line.startswith("source_digest"):
                    digest = line.split("=")[1].strip()
    except OSError:
        return None
    return found if digest == profiling.source_digest() else None


def stage_report(dev, dt, precision, n, steps=2):
    """the stage kernel over `steps` steps: each launch's own duration (HIP
    events in the dispatch), its algorithmic bytes against HBM peak, and its
    VALU issue fraction (the kernel's other bound)"""
    dev.prof_start("stage")
    for _ in range(steps):
        dev.step_rk3(dt)
    ms, nl, by = dev.prof_stop()
    if not nl:
        return None
    us = ms / nl * 1e3
    gbs = by / nl / (us * 1e-6) / 1e9
    return {"kernel": "k_stage (RK3 modes 1-3)", "launches": nl, "mean_launch_us": round(us, 2),
            "alg_bytes_per_launch": by / nl, "achieved": round(gbs, 2), "unit": "GB/s",
            "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
            "valu_issue_frac": measured_valu(precision, n)}


def error_check(n, dt, steps):
    """Advance both paths `steps` steps from the same state; errors of the
    mixed path relative to FP64 (bit-identical to the oracle)."""
    models = {p: make_model(n, p) for p in ("fp64", "mixed")}
    for m in models.values():
        for _ in range(steps):
            m.step_rk3(dt)
    out = {"steps": steps}
    out.update(state_errors(models["mixed"], models["fp64"]))
    for m in models.values():
        m.close()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: the Nyles time step here runs one model per GPU "
              "and has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from nyles_amd import _lib as L

    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb)
        return
    n = args.n
    other = "fp64" if args.precision == "mixed" else "mixed"
    runs = {}
    dt = None
    for prec in (args.precision,) if args.single else (args.precision, other):
        print("timing %s %d^3" % (prec, n), file=sys.stderr, flush=True)
        dev = make_model(n, prec)
        if dt is None:
            dt = 0.5 / float(np.sqrt(dev.reduce(L.NYL_RED_MAXU2)))
        for _ in range(args.warmup):
            dev.step_rk3(dt)
        br = class_times(dev, dt) if args.breakdown else {}
        kclass = args.roofline_kernel
        el, nite, prof = timed(dev, dt, args.steps, 0, kclass)
        runs[prec] = {"el": el, "nite": nite, "prof": prof, "kclass": kclass, "breakdown": br,
                      "stage": stage_report(dev, dt, prec, n)}
        dev.close()
        torch.cuda.synchronize()

    head = runs[args.precision]
    value = n ** 3 * args.steps / head["el"] / 1e6
    kms, klaunch, kbytes = head["prof"]
    achieved = (kbytes / klaunch) / (kms / klaunch / 1e3) / 1e9 if klaunch else None
    oth = runs.get(other)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(head["el"] / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64", "data": "synthetic",
        "config": {
            "workload": "taylor-green vortex %d^3 perio_xyz: LES, RK3_SSP, WENO5 vortex force "
                        "and tracer, multigrid projection (mgfor V-cycles, tol 1e-6 squared)" % n,
            "grid": [n, n, n], "precision": args.precision, "dt": dt, "seed": 2,
            "mg_vcycles_per_step": round(head["nite"] / args.steps, 3),
            "mg_vcycles_per_solve": round(head["nite"] / (3 * args.steps), 3),
        },
        "roofline": {
            "bound": "hbm", "kernel": head["kclass"],
            "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": measured_traffic(kclass, args.precision, n), "launches": klaunch,
            "mean_launch_us": round(kms / klaunch * 1e3, 2) if klaunch else None,
            "alg_bytes_per_launch": kbytes / klaunch if klaunch else None,
        },
    }
    out["stage_kernel"] = head["stage"]
    if oth is not None:
        out[other] = {
            "value": round(n ** 3 * args.steps / oth["el"] / 1e6, 3),
            "ms_per_step": round(oth["el"] / args.steps * 1e3, 4),
            "mg_vcycles_per_step": round(oth["nite"] / args.steps, 3),
            "stage_kernel": oth["stage"],
        }
        out["speedup_mixed_vs_fp64"] = round(runs["fp64"]["el"] / runs["mixed"]["el"], 3)
    if args.breakdown:
        out["breakdown"] = {p: runs[p]["breakdown"] for p in runs}
    if args.err_steps > 0:
        print("error check, %d steps" % args.err_steps, file=sys.stderr, flush=True)
        err = {"tolerance": "KE, velocity and buoyancy within 1e-5 relative of FP64",
               "taylor_green_%d" % n: error_check(n, dt, args.err_steps),
               "lock_exchange_64x32x32": lock_exchange_error(args.err_steps)}
        out["error_vs_fp64"] = err
    if args.ensemble_members > 0:
        print("ensemble, %d members" % args.ensemble_members, file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_run(args.ensemble_members, args.ensemble_steps,
                                       args.ensemble_seq_steps)
    if args.serving_jobs > 0:
        print("serving, %d jobs" % args.serving_jobs, file=sys.stderr, flush=True)
        out["serving"] = serving_run(args.serving_jobs, args.serving_rate)
    if args.rt_steps > 0:
        print("rayleigh-taylor, %d steps" % args.rt_steps, file=sys.stderr, flush=True)
        out["rayleigh_taylor"] = rt_run(max(args.rt_steps - 1, 1))
        if "error_vs_fp64" in out:
            e = dict(out["rayleigh_taylor"]["error_vs_fp64"])
            e["steps"] = out["rayleigh_taylor"]["steps"]
            out["error_vs_fp64"]["rayleigh_taylor_256x256x512"] = e
    if not args.no_cpu_baseline:
        print("cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps, args.cpu_steps_1thread)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
