"""One Nyles model instance resident on the GPU.

``DeviceModel`` owns a libnyles_hip context and the device workspace it
computes in.  PyTorch is used only to allocate that workspace and to move
arrays between host and device; every arithmetic operation on the step path
runs in the library's HIP kernels.
"""
import ctypes

import numpy as np
import torch

from . import _lib as L

STATE_FIELDS = {"b": L.NYL_F_B, "u_i": L.NYL_F_U, "u_j": L.NYL_F_V, "u_k": L.NYL_F_W}
DIAG_FIELDS = {"ke": L.NYL_F_KE, "vor_i": L.NYL_F_VORX, "vor_j": L.NYL_F_VORY,
               "vor_k": L.NYL_F_VORZ, "U_i": L.NYL_F_UCX, "U_j": L.NYL_F_UCY,
               "U_k": L.NYL_F_UCZ, "p": L.NYL_F_P, "div": L.NYL_F_DIV}
_DT = {L.NYL_DT_F64: torch.float64, L.NYL_DT_F32: torch.float32}


class DeviceModel:
    """Device state and time step of one LES instance.

    Parameters mirror what core/model_les.py:33-95 reads from ``param`` and
    ``grid``; the spacing-derived constants are passed as the caller computed
    them (core/grid.py:30-44, core/mgfordriver.py:78) so the device uses the
    same doubles as the reference.
    """

    def __init__(self, nx, ny, nz, periodic, dx, dy, dz, idx2=None, idy2=None, idz2=None,
                 dx2=None, fparameter=0.0, precision=L.NYL_PREC_FP64, mg_maxite=20,
                 mg_tol=1e-6, mg_omega=0.9, device=None, viscosity=0.0, n_tracers=0,
                 model=0, upwind_a=0, upwind_vf=0):
        lib = L.load()
        if not torch.cuda.is_available():
            raise L.NylesHipError("no GPU visible: the Nyles time step runs only on the GPU")
        self.device = torch.device(device if device is not None else "cuda:0")
        self.shape = (nz, ny, nx)
        d = L.NylDesc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.periodic_x, d.periodic_y, d.periodic_z = (int(bool(p)) for p in periodic)
        d.dx, d.dy, d.dz = dx, dy, dz
        d.idx2 = idx2 if idx2 is not None else 1 / dx ** 2
        d.idy2 = idy2 if idy2 is not None else 1 / dy ** 2
        d.idz2 = idz2 if idz2 is not None else 1 / dz ** 2
        d.dx2 = dx2 if dx2 is not None else dx ** 2
        d.fparameter = fparameter
        d.precision = precision
        d.mg_maxite = mg_maxite
        d.mg_tol = mg_tol
        d.mg_omega = mg_omega
        d.viscosity = viscosity
        d.n_tracers = n_tracers
        d.model = model
        d.upwind_a, d.upwind_vf = upwind_a, upwind_vf
        self.n_tracers = n_tracers
        self.model = model
        self.desc = d
        ctx = ctypes.c_void_p()
        L.check(lib.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)))
        self.ctx = ctx
        nbytes = ctypes.c_uint64()
        L.check(lib.nyl_workspace_bytes(ctx, ctypes.byref(nbytes)))
        self.workspace_bytes = nbytes.value
        with torch.cuda.device(self.device):
            self.ws = torch.empty(self.workspace_bytes, dtype=torch.uint8, device=self.device)
            # a stream of its own (default priority): the library captures the
            # coarse multigrid levels into a hipGraph, which the null stream
            # does not allow
            self.stream = torch.cuda.Stream(device=self.device)
        torch.cuda.current_stream(self.device).synchronize()   # ws allocated
        L.check(lib.nyl_set_stream(ctx, ctypes.c_void_p(self.stream.cuda_stream)))
        L.check(lib.nyl_bind_workspace(ctx, ctypes.c_void_p(self.ws.data_ptr()),
                                       self.workspace_bytes))
        self.lib = lib
        self.last_stats = L.NylMGStats()

    # -- fields -------------------------------------------------------------
    def field(self, fid):
        """torch view (nz, ny, nx) of a workspace field (state views move
        between the two state buffers after every stage)."""
        off, n, dt = ctypes.c_uint64(), ctypes.c_int64(), ctypes.c_int32()
        L.check(self.lib.nyl_field_offset(self.ctx, fid, ctypes.byref(off), ctypes.byref(n),
                                          ctypes.byref(dt)))
        tdt = _DT[dt.value]
        esz = torch.tensor([], dtype=tdt).element_size()
        flat = self.ws[off.value: off.value + n.value * esz].view(tdt)
        return flat.view(self.shape)

    def upload(self, fid, array):
        a = np.ascontiguousarray(array)
        assert a.shape == self.shape, (a.shape, self.shape)
        dst = self.field(fid)
        with torch.cuda.stream(self.stream):
            dst.copy_(torch.from_numpy(a).to(dst.dtype), non_blocking=False)

    def download(self, fid):
        src = self.field(fid)
        with torch.cuda.stream(self.stream):
            out = src.to("cpu")
        self.stream.synchronize()
        return out.numpy().astype(np.float64, copy=False)

    # -- operations ---------------------------------------------------------
    def rhs(self, slot, last=False):
        L.check(self.lib.nyl_rhs(self.ctx, slot, int(bool(last))))

    def rk3_stage(self, stage, coef):
        L.check(self.lib.nyl_rk3_stage(self.ctx, stage, float(coef)))

    def stage_update(self, mode, c, c2=0.0):
        """stage combination of mode 1..6 from the stored tendencies
        (nyl_stage_update)"""
        L.check(self.lib.nyl_stage_update(self.ctx, int(mode), float(c), float(c2)))

    def diagnose(self):
        st = L.NylMGStats()
        L.check(self.lib.nyl_diagnose(self.ctx, ctypes.byref(st)))
        self.last_stats = st
        return st.as_dict()

    def step_rk3(self, dt):
        st = L.NylMGStats()
        L.check(self.lib.nyl_step_rk3(self.ctx, float(dt), ctypes.byref(st)))
        self.last_stats = st
        return st.as_dict()

    def step_ef(self, dt):
        st = L.NylMGStats()
        L.check(self.lib.nyl_step_ef(self.ctx, float(dt), ctypes.byref(st)))
        self.last_stats = st
        return st.as_dict()

    def step_lfam3(self, dt):
        st = L.NylMGStats()
        L.check(self.lib.nyl_step_lfam3(self.ctx, float(dt), ctypes.byref(st)))
        self.last_stats = st
        return st.as_dict()

    def lfam3_reset(self):
        L.check(self.lib.nyl_lfam3_reset(self.ctx))

    def mg_solve(self):
        st = L.NylMGStats()
        L.check(self.lib.nyl_mg_solve(self.ctx, ctypes.byref(st)))
        self.last_stats = st
        return st.as_dict()

    def apply(self, op):
        L.check(self.lib.nyl_apply(self.ctx, op))

    def reduce(self, what):
        out = ctypes.c_double()
        L.check(self.lib.nyl_reduce(self.ctx, what, ctypes.byref(out)))
        return out.value

    def prof_start(self, kernel_class):
        L.check(self.lib.nyl_prof_start(self.ctx, L.KERNEL_CLASSES[kernel_class]))

    def prof_stop(self):
        """(total kernel ms, launches, algorithmic bytes) since prof_start."""
        ms, n, by = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
        L.check(self.lib.nyl_prof_stop(self.ctx, ctypes.byref(ms), ctypes.byref(n),
                                       ctypes.byref(by)))
        return ms.value, n.value, by.value

    def debug_mg(self, what):
        """Level-1 multigrid kernel check (see nyl_debug_mg); returns the
        reduced residual norm for what = 2, 3, 5."""
        L.check(self.lib.nyl_debug_mg(self.ctx, what))
        if what in (2, 3, 5):
            out = ctypes.c_double()
            L.check(self.lib.nyl_reduce_last(self.ctx, ctypes.byref(out)))
            return out.value
        return None

    def sync(self):
        L.check(self.lib.nyl_sync(self.ctx))

    def set_graphs(self, on):
        """replay the coarse multigrid levels from a hipGraph (default) or not"""
        L.check(self.lib.nyl_set_graphs(self.ctx, int(bool(on))))

    def mg_levels(self):
        n = ctypes.c_int32()
        L.check(self.lib.nyl_mg_levels(self.ctx, ctypes.byref(n)))
        shapes = []
        for l in range(n.value):
            a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
            L.check(self.lib.nyl_mg_level_shape(self.ctx, l, ctypes.byref(a), ctypes.byref(b),
                                                ctypes.byref(c)))
            shapes.append((c.value, b.value, a.value))
        return shapes

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.sync()
            self.lib.nyl_ctx_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tile_candidates():
    """the stage tiles the batch autotuner tries, as (cells, tx, ty)"""
    lib = L.load()
    out = []
    for q in range(L.NYL_TILE_CANDIDATES):
        t = (ctypes.c_int32 * 3)()
        L.check(lib.nyl_batch_tile_candidate(q, t))
        out.append((t[0], t[1], t[2]))
    return out


def tile_name(t):
    """"256" for a row tile of 256 cells, "32x4x2" for a brick"""
    cells, tx, ty = t
    return str(cells) if tx == 0 else "%dx%dx%d" % (tx, ty, cells // (tx * ty))


def parse_tile(t):
    """(cells, tx, ty) from an int (row tile), a tile name or a triple"""
    if isinstance(t, int):
        return (t, 0, 0)
    if isinstance(t, str):
        parts = [int(v) for v in t.split("x")]
        if len(parts) == 1:
            return (parts[0], 0, 0)
        tx, ty, tz = parts
        return (tx * ty * tz, tx, ty)
    return tuple(int(v) for v in t)


class DeviceBatch:
    """An ensemble of DeviceModel members advanced together (nyl_batch_*).

    Each RK stage is one persistent stage-kernel launch per periodicity class
    and one launch per kernel class of the projection for all members, with
    every member's multigrid stop rule evaluated on the device; the plans are
    replayed from hipGraphs.  The batch runs on its own stream (default
    priority); members keep their workspaces and may be used alone between
    batch steps.
    """

    def __init__(self, members, graphs=True, autotune=True, tile=None):
        if not members:
            raise ValueError("a batch needs at least one member")
        self.lib = L.load()
        self.members = list(members)
        dev = self.members[0].device
        arr = (ctypes.c_void_p * len(self.members))(*[m.ctx.value for m in self.members])
        b = ctypes.c_void_p()
        L.check(self.lib.nyl_batch_create(arr, len(self.members), ctypes.byref(b)))
        self.batch = b
        nbytes = ctypes.c_uint64()
        L.check(self.lib.nyl_batch_workspace_bytes(b, ctypes.byref(nbytes)))
        with torch.cuda.device(dev):
            self.stream = torch.cuda.Stream(device=dev)
            self.ws = torch.empty(max(nbytes.value, 256), dtype=torch.uint8, device=dev)
        # the members' data must be ready before the batch stream reads it
        for m in self.members:
            m.stream.synchronize()
        L.check(self.lib.nyl_batch_set_stream(b, ctypes.c_void_p(self.stream.cuda_stream)))
        L.check(self.lib.nyl_batch_bind_workspace(b, ctypes.c_void_p(self.ws.data_ptr()),
                                                  nbytes.value))
        L.check(self.lib.nyl_batch_set_graphs(b, int(bool(graphs))))
        self.tile = (256, 0, 0)
        self.tune_ms = None
        if tile is not None:
            self.set_tile(tile)
        elif autotune:
            self.autotune()

    def autotune(self):
        """time RK stage 1 for every candidate stage tile, keep the fastest"""
        t = (ctypes.c_int32 * 3)()
        ms = (ctypes.c_double * L.NYL_TILE_CANDIDATES)()
        L.check(self.lib.nyl_batch_autotune(self.batch, t, ms))
        self.tile = (t[0], t[1], t[2])
        self.tune_ms = {tile_name(c): ms[q] for q, c in enumerate(tile_candidates())}
        return self.tile

    def set_tile(self, tile):
        """tile: cells of a row tile (int), a name ("256", "32x4x2") or (cells, tx, ty)"""
        t = parse_tile(tile)
        L.check(self.lib.nyl_batch_set_tile(self.batch, *t))
        self.tile = t

    def set_graphs(self, on):
        """replay the step plans from hipGraphs (captured on next use) or not"""
        L.check(self.lib.nyl_batch_set_graphs(self.batch, int(bool(on))))

    def step_rk3(self, dts):
        """one RK3_SSP step of every member (member q with dts[q])"""
        n = len(self.members)
        if len(dts) != n:
            raise ValueError("one dt per member")
        arr = (ctypes.c_double * n)(*[float(d) for d in dts])
        st = (L.NylMGStats * n)()
        for m in self.members:          # the batch stream reads what they wrote
            m.stream.synchronize()
        L.check(self.lib.nyl_batch_step_rk3(self.batch, arr, st))
        return [s.as_dict() for s in st]

    def close(self):
        if getattr(self, "batch", None) is not None and self.batch.value:
            self.stream.synchronize()
            self.lib.nyl_batch_destroy(self.batch)
            self.batch = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
