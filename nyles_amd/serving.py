"""Serving: a request loop that runs simulation jobs of different sizes on one
GPU with continuous batching.

Jobs arrive over time (``Job.arrival``, seconds after the session starts).
Between batched steps the server admits the jobs that have arrived while
slots and HBM allow (admission by the memory its resident jobs hold against
a budget, checked against the device's free memory), keeps every admitted
job's state resident on the device, advances all resident jobs by one RK3
step in one batch (``DeviceBatch``: one set of launches for all of them, so
their work overlaps on the device), and retires each job when it has done
its steps.  A job's result does not depend on the other jobs it was batched
with (the batched kernels are per-member computations).

The report gives p50/p99 time-to-completion (arrival to last step) beside
the throughput in cell-steps per second.
"""
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib as L
from . import workloads
from .device import DeviceBatch, DeviceModel, parse_tile, tile_name

GEOMETRY_XYZ = {"closed": (0, 0, 0), "perio_x": (1, 0, 0), "perio_y": (0, 1, 0),
                "perio_xy": (1, 1, 0), "perio_xyz": (1, 1, 1)}
PRECISION = {"fp64": L.NYL_PREC_FP64, "mixed": L.NYL_PREC_MIXED}


@dataclass
class Job:
    """One simulation request.

    kind: "lock" (lock exchange, n x n/2 x n/2, perio_x), "tgv" (Taylor-Green
    n^3, perio_xyz) or "shear" (stratified shear layer n^3, perio_xy); the
    time step is fixed at admission (CFL 0.5 of the initial flow, or of the
    buoyancy wave speed for the lock exchange, which starts at rest).
    """
    kind: str
    n: int
    steps: int
    seed: int = 0
    arrival: float = 0.0
    id: int = -1
    admitted_at: float = math.nan
    done_at: float = math.nan
    dt: float = math.nan
    steps_done: int = 0
    cells: int = 0
    ws_bytes: int = 0
    vcycles: int = 0
    result: dict = field(default_factory=dict)
    model: object = None
    failed: str = ""
    bypassed: int = 0

    @property
    def latency(self):
        return self.done_at - self.arrival


def job_grid(job):
    """(nx, ny, nz, geometry, (Lx, Ly, Lz)) of a job"""
    n = job.n
    if job.kind == "lock":
        return n, n // 2, n // 2, "perio_x", (16.0, 8.0, 8.0)
    if job.kind == "tgv":
        return n, n, n, "perio_xyz", (2 * np.pi,) * 3
    if job.kind == "shear":
        return n, n, n, "perio_xy", (1.0, 1.0, 1.0)
    raise ValueError("unknown job kind %r" % job.kind)


def initial_state(job):
    nx, ny, nz, geo, Ls = job_grid(job)
    if job.kind == "lock":
        st, _ = workloads.lock_exchange(nx, ny, nz, *Ls, seed=job.seed)
    elif job.kind == "tgv":
        st, _ = workloads.taylor_green(nx, seed=job.seed)
    else:
        st, _ = workloads.shear_layer(nx, seed=job.seed)
    return st


def workspace_bytes(job, precision):
    """device bytes a job holds while resident (queried from the library)"""
    import ctypes
    lib = L.load()
    nx, ny, nz, geo, Ls = job_grid(job)
    d = L.NylDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.periodic_x, d.periodic_y, d.periodic_z = GEOMETRY_XYZ[geo]
    dx, dy, dz = Ls[0] / nx, Ls[1] / ny, Ls[2] / nz
    d.dx, d.dy, d.dz = dx, dy, dz
    d.idx2, d.idy2, d.idz2 = 1 / dx ** 2, 1 / dy ** 2, 1 / dz ** 2
    d.dx2 = dx ** 2
    d.precision = PRECISION[precision]
    d.mg_maxite, d.mg_tol, d.mg_omega = 20, 1e-6, 0.9
    ctx = ctypes.c_void_p()
    L.check(lib.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)))
    nb = ctypes.c_uint64()
    L.check(lib.nyl_workspace_bytes(ctx, ctypes.byref(nb)))
    lib.nyl_ctx_destroy(ctx)
    return nb.value


def make_model(job, precision):
    """device model of a job with its initial state projected, and its dt"""
    nx, ny, nz, geo, Ls = job_grid(job)
    dx, dy, dz = Ls[0] / nx, Ls[1] / ny, Ls[2] / nz
    m = DeviceModel(nx, ny, nz, GEOMETRY_XYZ[geo], dx, dy, dz, idx2=1 / dx ** 2,
                    idy2=1 / dy ** 2, idz2=1 / dz ** 2, dx2=dx ** 2,
                    precision=PRECISION[precision])
    st = initial_state(job)
    for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                   ("u_k", L.NYL_F_W)):
        m.upload(fid, st[k])
    m.diagnose()
    if job.kind == "lock":
        dt = 0.5 * dx / math.sqrt(1.0 * Ls[2])     # buoyancy jump 1, depth Lz
    else:
        dt = 0.5 / math.sqrt(m.reduce(L.NYL_RED_MAXU2))
    return m, dt


def final_result(model, keep_state):
    r = {"ke": model.reduce(L.NYL_RED_KE), "bsum": model.reduce(L.NYL_RED_BSUM)}
    if keep_state:
        r["state"] = {k: model.download(f) for k, f in (("b", L.NYL_F_B), ("u", L.NYL_F_U),
                                                         ("v", L.NYL_F_V), ("w", L.NYL_F_W))}
    return r


class Admission:
    """Which arrived jobs to admit between two batched steps (host logic).

    Jobs are considered in arrival order.  A job is admitted when a slot is
    free, the resident jobs' bytes plus its own stay within the budget and
    it fits in the device's free memory.  A job that does not fit is passed
    over by later, smaller jobs at most ``max_bypass`` times; after that no
    later job is admitted before it (the memory that frees up is kept for
    it).  A job that can never fit (more bytes than the budget, or more than
    the free memory while none of the server's jobs is resident) is
    rejected instead of blocking the queue.
    """

    def __init__(self, budget, max_slots, max_bypass=4):
        self.budget = int(budget)
        self.max_slots = int(max_slots)
        self.max_bypass = int(max_bypass)

    def select(self, arrived, n_active, used, free, need):
        """arrived: jobs that have arrived and wait, in arrival order;
        n_active: resident jobs; used: bytes they hold; free: free device
        bytes; need(job): bytes the job would hold.
        Returns (admit, reject): lists of jobs, in arrival order."""
        admit, reject, waiting = [], [], []
        for j in arrived:
            nb = need(j)
            if nb > self.budget:
                j.failed = "needs %d bytes, above the HBM budget of %d" % (nb, self.budget)
                reject.append(j)
                continue
            if nb > free and n_active == 0 and not admit:
                j.failed = "needs %d bytes, above the free device memory (%d) " \
                           "with no job resident" % (nb, free)
                reject.append(j)
                continue
            if (n_active + len(admit) < self.max_slots and used + nb <= self.budget
                    and nb <= free):
                admit.append(j)
                used += nb
                free -= nb
                for w in waiting:
                    w.bypassed += 1
                continue
            waiting.append(j)
            if j.bypassed >= self.max_bypass or n_active + len(admit) >= self.max_slots:
                break
        return admit, reject


class Server:
    """Continuous-batching request loop on one GPU.

    max_slots: most jobs resident at once; hbm_fraction: share of the device
    memory the resident jobs may hold; tile: stage tile of the batches (None:
    autotune whenever the number of resident jobs doubles or halves);
    graphs: True/False, or "auto": a batch replays its plans from hipGraphs
    once its membership has stayed the same for graph_after steps (capturing
    the plans of a batch that changes every few steps costs more than the
    replays save).
    """

    def __init__(self, precision="mixed", max_slots=64, hbm_fraction=0.8, graphs="auto",
                 tile=None, keep_state=False, graph_after=8):
        if precision not in PRECISION:
            raise ValueError("precision must be fp64 or mixed")
        self.precision = precision
        self.max_slots = max_slots
        if graphs not in (True, False, "auto"):
            raise ValueError('graphs must be True, False or "auto"')
        self.graphs = graphs
        self.graph_after = graph_after
        self.tile = tile
        self.keep_state = keep_state
        props = torch.cuda.get_device_properties(0)
        self.budget = int(hbm_fraction * props.total_memory)
        self.max_bypass = 4

    def run(self, jobs, log=None):
        for q, j in enumerate(jobs):
            j.id = q
        pending = sorted(jobs, key=lambda j: j.arrival)
        active, done, rejected = [], [], []
        policy = Admission(self.budget, self.max_slots, self.max_bypass)
        needs = {}

        def need(j):
            if j.id not in needs:
                needs[j.id] = workspace_bytes(j, self.precision)
            return needs[j.id]
        used = 0
        batch = None
        dirty = False
        stable = 0
        tuned_for, tile = 0, self.tile
        rebuilds, steps, slot_steps, cellsteps = 0, 0, 0, 0
        peak = 0
        t0 = time.perf_counter()
        while pending or active:
            now = time.perf_counter() - t0
            arrived = [j for j in pending if j.arrival <= now]
            admit, reject = [], []
            if arrived:
                free, _ = torch.cuda.mem_get_info()
                free += torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
                admit, reject = policy.select(arrived, len(active), used, free, need)
            taken = set(id(j) for j in admit + reject)
            if taken:
                pending = [j for j in pending if id(j) not in taken]
            for j in reject:
                j.done_at = now
                rejected.append(j)
                if log:
                    log("job %d %s n=%d rejected: %s" % (j.id, j.kind, j.n, j.failed))
            for j in admit:
                j.model, j.dt = make_model(j, self.precision)
                j.ws_bytes = need(j)
                nx, ny, nz, _, _ = job_grid(j)
                j.cells = nx * ny * nz
                j.admitted_at = time.perf_counter() - t0
                used += j.ws_bytes
                peak = max(peak, used)
                active.append(j)
                dirty = True
            if not active:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            if dirty:
                if batch is not None:
                    batch.close()
                n = len(active)
                retune = tile is None or n >= 2 * tuned_for or 2 * n <= tuned_for
                g0 = self.graphs is True
                if self.tile is None and retune:
                    batch = DeviceBatch([j.model for j in active], graphs=g0, autotune=True)
                    tile, tuned_for = batch.tile, n
                else:
                    batch = DeviceBatch([j.model for j in active], graphs=g0, tile=tile)
                rebuilds += 1
                dirty = False
                stable = 0
            if self.graphs == "auto":
                stable += 1
                if stable == self.graph_after:
                    batch.set_graphs(True)
            st = batch.step_rk3([j.dt for j in active])
            steps += 1
            slot_steps += len(active)
            for j, s in zip(active, st):
                j.steps_done += 1
                j.vcycles += s["nite"]
                cellsteps += j.cells
            finished = [j for j in active if j.steps_done >= j.steps]
            if finished:
                t = time.perf_counter() - t0
                for j in finished:
                    j.done_at = t
                    j.result = final_result(j.model, self.keep_state)
                    j.model.close()
                    j.model = None
                    used -= j.ws_bytes
                    active.remove(j)
                    done.append(j)
                    if log:
                        log("job %d %s n=%d steps=%d latency %.3f s" % (
                            j.id, j.kind, j.n, j.steps, j.latency))
                dirty = True
        if batch is not None:
            batch.close()
        elapsed = time.perf_counter() - t0
        lat = np.array([j.latency for j in done]) if done else np.array([math.nan])
        return {
            "jobs": len(done), "rejected": len(rejected), "precision": self.precision,
            "elapsed_s": elapsed,
            "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()),
            "Mcell_steps_per_s": cellsteps / elapsed / 1e6,
            "batched_steps": steps, "mean_slots": slot_steps / max(steps, 1),
            "batch_rebuilds": rebuilds,
            "tile": tile_name(parse_tile(tile)) if tile is not None else None,
            "peak_resident_bytes": peak,
            "budget_bytes": self.budget,
        }


def poisson_jobs(count, rate, seed=5, sizes=(32, 64, 128), steps=(50, 200),
                 kinds=("lock", "tgv", "shear"), size_range=None):
    """`count` jobs arriving as a Poisson process of `rate` per second; kind
    uniform, steps uniform in [steps[0], steps[1]]; n uniform over `sizes`,
    or, with size_range = (lo, hi), log-uniform in [lo, hi] rounded to a
    multiple of 8"""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for q in range(count):
        t += rng.exponential(1.0 / rate)
        kind = str(rng.choice(kinds))
        if size_range is None:
            n = int(rng.choice(sizes))
        else:
            v = math.exp(rng.uniform(math.log(size_range[0]), math.log(size_range[1])))
            n = int(8 * round(v / 8))
        jobs.append(Job(kind=kind, n=n, steps=int(rng.integers(steps[0], steps[1] + 1)),
                        seed=int(rng.integers(0, 2 ** 31)), arrival=t))
    return jobs
