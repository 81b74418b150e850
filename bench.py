"""Benchmark of the Nyles time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: Taylor-Green vortex on a triply periodic 256^3 grid (the grid the
headline metric is quoted on), LES model with RK3_SSP stepping, WENO5
vortex force and tracer advection, mgfor-equivalent multigrid projection,
fixed dt = 0.5 / max|U| at t = 0.  One step = one full RK3 step (3
right-hand sides + 3 projections); one cell-step = one interior cell advanced
by one step.  The state is resident in HBM before the timed region starts.

Prints ONE JSON line.  ``roofline`` prices the dominant kernel: achieved =
its algorithmic HBM bytes per launch (each input field read once, each
output written once) divided by its mean launch time, both measured with HIP
events on the library's stream over the timed steps.  ``cpu_baseline`` times
the FP64 CPU oracle (C restatement of the reference kernels + numpy
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
    ap.add_argument("--precision", choices=["fp64"], default="fp64")
    ap.add_argument("--roofline-kernel", default="stage")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--breakdown", action="store_true",
                    help="also time every kernel class (one extra pass each)")
    return ap.parse_args()


def cpu_baseline(n, steps):
    """FP64 oracle on the host cores: initial projection untimed, then `steps`
    RK3 steps timed."""
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
    O.set_threads(cores)
    st, meta = workloads.taylor_green(n, seed=2)
    L = meta["L"]
    orc = O.OracleLES(n, n, n, L[0], L[1], L[2], "perio_xyz")
    orc.set_interior(orc.b, st["b"])
    orc.set_interior(orc.u["i"], st["u_i"])
    orc.set_interior(orc.u["j"], st["u_j"])
    orc.set_interior(orc.u["k"], st["u_k"])
    orc.diagnose_var()
    U = orc.interior(orc.U["i"]), orc.interior(orc.U["j"]), orc.interior(orc.U["k"])
    dt = 0.5 / float(np.sqrt(np.max(U[0] ** 2 + U[1] ** 2 + U[2] ** 2)))
    t0 = time.perf_counter()
    for _ in range(steps):
        orc.rk3_ssp(dt)
    el = time.perf_counter() - t0
    return {"value": round(n ** 3 * steps / el / 1e6, 4), "unit": "Mcell-steps/s",
            "cores": O.num_threads(), "kind": "port",
            "sample": "taylor-green %d^3 fp64, %d RK3 step(s) after the initial projection, "
                      "%.1f s wall on %d OpenMP threads" % (n, steps, el, O.num_threads())}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from nyles_amd import _lib as L
    from nyles_amd import workloads
    from nyles_amd.device import DeviceModel

    n = args.n
    st, meta = workloads.taylor_green(n, seed=2)
    Lx, Ly, Lz = meta["L"]
    dx, dy, dz = Lx / n, Ly / n, Lz / n
    dev = DeviceModel(n, n, n, (1, 1, 1), dx, dy, dz, idx2=1 / dx ** 2, idy2=1 / dy ** 2,
                      idz2=1 / dz ** 2, dx2=dx ** 2, precision=L.NYL_PREC_FP64)
    for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                   ("u_k", L.NYL_F_W)):
        dev.upload(fid, st[k])
    dev.diagnose()
    dt = 0.5 / float(np.sqrt(dev.reduce(L.NYL_RED_MAXU2)))
    for _ in range(args.warmup):
        dev.step_rk3(dt)
    torch.cuda.synchronize()

    dev.prof_start(args.roofline_kernel)
    nite = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nite += dev.step_rk3(dt)["nite"]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kms, klaunch, kbytes = dev.prof_stop()

    value = n ** 3 * args.steps / el / 1e6
    achieved = (kbytes / klaunch) / (kms / klaunch / 1e3) / 1e9 if klaunch else None
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mcell-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {
            "workload": "taylor-green vortex %d^3 perio_xyz: LES, RK3_SSP, WENO5 vortex force "
                        "and tracer, multigrid projection (mgfor V-cycles, tol 1e-6 squared)" % n,
            "grid": [n, n, n], "precision": args.precision, "dt": dt, "seed": 2,
            "mg_vcycles_per_step": round(nite / args.steps, 3),
            "mg_vcycles_per_solve": round(nite / (3 * args.steps), 3),
        },
        "roofline": {
            "bound": "hbm", "kernel": args.roofline_kernel,
            "achieved": round(achieved, 2) if achieved else None, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": None, "launches": klaunch,
            "mean_launch_us": round(kms / klaunch * 1e3, 2) if klaunch else None,
            "alg_bytes_per_launch": kbytes / klaunch if klaunch else None,
        },
    }
    if args.breakdown:
        br = {}
        for kc in L.KERNEL_CLASSES:
            dev.prof_start(kc)
            dev.step_rk3(dt)
            ms, nl, by = dev.prof_stop()
            br[kc] = {"ms_per_step": round(ms, 4), "launches": nl,
                      "GB/s": round(by / (ms / 1e3) / 1e9, 1) if ms > 0 else None}
        out["breakdown"] = br
    dev.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
