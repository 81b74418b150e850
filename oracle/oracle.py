"""CPU oracle of the Nyles time step (TEST INFRASTRUCTURE ONLY).

This module is the checker the HIP library is compared with.  Only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it; the product path (``nyles_amd``) never does.

It restates the reference's Python orchestration around the C restatement
of the Fortran kernels (``nyles_oracle.c``):

* ``OracleLES.rhs``            core/model_les.py:129-144
* ``OracleLES.diagnose_var``   core/model_les.py:97-127
* ``OracleLES.rk3_ssp``        core/timescheme.py:180-221
* ``OracleLES.compute_dt``     core/nyles.py:227-260
* ``halo_fill``                core/mpi/halo.py:123-188 (one rank)
* ``OracleMG.solve_directly``  core/mgfordriver.py:66-78

Arrays keep the reference layout: the 'i' orientation ``[k, j, i]`` with a
halo of ``nh`` cells only along periodic axes (core/mpi/topology.py:253-304).
The other orientations are transposed *views* (core/variables.py:146-150,
404-411); the C kernels take strides, so no copies are needed.
"""
import ctypes
import itertools
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnyles_oracle.so")

_lib = None


def build(force=False):
    """Compile the C restatement (gcc, no FMA contraction, OpenMP)."""
    import subprocess
    src = os.path.join(_HERE, "nyles_oracle.c")
    if not force and os.path.exists(LIB_PATH) and \
            os.path.getmtime(LIB_PATH) >= os.path.getmtime(src):
        return LIB_PATH
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
           "-fno-fast-math", "-std=c99", "-o", LIB_PATH + ".tmp", src, "-lm"]
    subprocess.check_call(cmd)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        ci = ctypes.c_int
        L.orc_weno3.restype = ctypes.c_double
        L.orc_weno3.argtypes = [ctypes.c_double] * 3
        L.orc_weno5.restype = ctypes.c_double
        L.orc_weno5.argtypes = [ctypes.c_double] * 5
        L.orc_flux1d.argtypes = [dp, dp, dp, ci]
        for name in ("orc_upwind", "orc_vortex_force_direc", "orc_vortex_force_flip",
                     "orc_vorticity"):
            getattr(L, name).argtypes = [dp, dp, dp, ci, ci, ci, i64, i64, i64]
        for name in ("orc_upwind_linear", "orc_vortex_force_direc_linear",
                     "orc_vortex_force_flip_linear"):
            getattr(L, name).argtypes = [dp, dp, dp, ci, ci, ci, ci, i64, i64, i64]
        L.orc_interp_tracer.argtypes = [dp, dp, dp, ci, ci]
        L.orc_interp_vf.argtypes = [dp, dp, dp, ci, ci]
        L.orc_gradke.argtypes = [dp, dp, ci, ci, ci, i64, i64, i64]
        L.orc_gradkeandb.argtypes = [dp, dp, dp, ctypes.c_double, ci, ci, ci, i64, i64, i64]
        L.orc_div.argtypes = [dp, dp, ci, ci, ci, ci, i64, i64, i64]
        L.orc_kin.argtypes = [dp, dp, ctypes.c_double, ci, ci, ci, i64, i64, i64]
        L.orc_add_laplacian.argtypes = [dp, dp, ctypes.c_double, ci, ci, ci, i64, i64, i64]
        L.orc_mg_create.restype = ctypes.c_void_p
        L.orc_mg_create.argtypes = [ci, ci, ci, ci, ci, ctypes.c_double, ctypes.c_double]
        L.orc_mg_destroy.argtypes = [ctypes.c_void_p]
        L.orc_mg_solve.restype = ci
        L.orc_mg_solve.argtypes = [ctypes.c_void_p]
        L.orc_mg_vcycle.argtypes = [ctypes.c_void_p]
        for name in ("orc_mg_smooth", "orc_mg_residual", "orc_mg_restrict", "orc_mg_prolong"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ci]
        L.orc_mg_norm.restype = ctypes.c_double
        L.orc_mg_norm.argtypes = [ctypes.c_void_p, ci, ci]
        L.orc_mg_nlevels.restype = ci
        L.orc_mg_nlevels.argtypes = [ctypes.c_void_p]
        L.orc_mg_shape.argtypes = [ctypes.c_void_p, ci, ctypes.POINTER(ci)]
        L.orc_mg_array.restype = dp
        L.orc_mg_array.argtypes = [ctypes.c_void_p, ci, ci]
        L.orc_mg_msk.restype = ctypes.POINTER(ci)
        L.orc_mg_msk.argtypes = [ctypes.c_void_p, ci]
        L.orc_mg_last_nite.restype = ci
        L.orc_mg_last_nite.argtypes = [ctypes.c_void_p]
        L.orc_mg_last_res.restype = ctypes.c_double
        L.orc_mg_last_res.argtypes = [ctypes.c_void_p]
        L.orc_mg_last_normb.restype = ctypes.c_double
        L.orc_mg_last_normb.argtypes = [ctypes.c_void_p]
        L.orc_mg_set_params.argtypes = [ctypes.c_void_p, ci, ctypes.c_double, ctypes.c_double]
        L.orc_mg_hierarchy.restype = ci
        L.orc_mg_hierarchy.argtypes = [ci, ci, ci, ci, ctypes.POINTER(ci), ctypes.POINTER(ci),
                                       ctypes.POINTER(ci), ci]
        L.orc_num_threads.restype = ci
        L.orc_set_num_threads.argtypes = [ci]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _view_args(a):
    """(l, m, n, s0, s1, s2) of a 3-D float64 view, strides in elements."""
    assert a.dtype == np.float64 and a.ndim == 3
    s = [st // 8 for st in a.strides]
    return (a.shape[0], a.shape[1], a.shape[2], s[0], s[1], s[2])


def _same_layout(*arrays):
    ref = arrays[0]
    for a in arrays[1:]:
        assert a.shape == ref.shape and a.strides == ref.strides


# --------------------------------------------------------------------------
# kernel wrappers (f2py call signatures of the reference)

def weno3(qm, q0, qp):
    return lib().orc_weno3(qm, q0, qp)


def weno5(qmm, qm, q0, qp, qpp):
    return lib().orc_weno5(qmm, qm, q0, qp, qpp)


def flux1d(u, q):
    u = np.ascontiguousarray(u, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    flux = np.empty_like(u)
    lib().orc_flux1d(_ptr(u), _ptr(q), _ptr(flux), len(u))
    return flux


def upwind(trac, u, dtrac):
    """core/fortran_upwind.f90:3 upwind(trac, u, dtrac, order)."""
    _same_layout(trac, u, dtrac)
    lib().orc_upwind(_vptr(trac), _vptr(u), _vptr(dtrac), *_view_args(dtrac))


def _vptr(a):
    """Pointer to element (0,0,0) of a (possibly transposed) view."""
    return ctypes.cast(a.__array_interface__["data"][0], ctypes.POINTER(ctypes.c_double))


def vortex_force_direc(U, vort, res):
    """core/fortran_vortex_force.f90:10 vortex_force_direc(U, vort, res, order)."""
    _same_layout(U, vort, res)
    lib().orc_vortex_force_direc(_vptr(U), _vptr(vort), _vptr(res), *_view_args(res))


def interp_tracer(q, order):
    """core/interpolate_tracer.f90, odd orders: (qp, qm) on a line"""
    q = np.ascontiguousarray(q, dtype=np.float64)
    qp, qm = np.empty_like(q), np.empty_like(q)
    lib().orc_interp_tracer(_ptr(q), _ptr(qp), _ptr(qm), int(order), q.size)
    return qp, qm


def interp_vf(v, order):
    """core/interpolate.f90, odd orders, on v(1..n): (qp(0..n-1), qm(1..n))"""
    n = len(v)
    v1 = np.zeros(n + 1)
    v1[1:] = v
    qp, qm = np.zeros(n + 1), np.zeros(n + 1)
    lib().orc_interp_vf(_ptr(v1), _ptr(qp), _ptr(qm), int(order), n)
    return qp[:n].copy(), qm[1:].copy()


def upwind_linear(trac, u, dtrac, order):
    """core/fortran_upwind.f90:34-67 (the linear branch), odd orders"""
    _same_layout(trac, u, dtrac)
    lib().orc_upwind_linear(_vptr(trac), _vptr(u), _vptr(dtrac), int(order),
                            *_view_args(dtrac))


def vortex_force_direc_linear(U, vort, res, order):
    """core/fortran_vortex_force.f90:43-65 (the linear branch), odd orders"""
    _same_layout(U, vort, res)
    lib().orc_vortex_force_direc_linear(_vptr(U), _vptr(vort), _vptr(res), int(order),
                                        *_view_args(res))


def vortex_force_flip_linear(U, vort, res, order):
    """core/fortran_vortex_force.f90:121-144 (the linear branch), odd orders"""
    _same_layout(U, vort, res)
    lib().orc_vortex_force_flip_linear(_vptr(U), _vptr(vort), _vptr(res), int(order),
                                       *_view_args(res))


def vortex_force_flip(U, vort, res):
    """core/fortran_vortex_force.f90:87 vortex_force_flip(U, vort, res, order)."""
    _same_layout(U, vort, res)
    lib().orc_vortex_force_flip(_vptr(U), _vptr(vort), _vptr(res), *_view_args(res))


def vorticity_k(ui, uj, wk):
    """core/fortran_vorticity.f90:2 vorticity(ui, uj, wk)."""
    _same_layout(ui, uj, wk)
    lib().orc_vorticity(_vptr(ui), _vptr(uj), _vptr(wk), *_view_args(wk))


def gradke(ke, du):
    """core/fortran_bernoulli.f90:2 gradke(ke, du)."""
    _same_layout(ke, du)
    lib().orc_gradke(_vptr(ke), _vptr(du), *_view_args(du))


def gradkeandb(ke, b, du, dz):
    """core/fortran_bernoulli.f90:29 gradkeandb(ke, b, du, dz)."""
    _same_layout(ke, b, du)
    lib().orc_gradkeandb(_vptr(ke), _vptr(b), _vptr(du), float(dz), *_view_args(du))


def div(d, u, iflag):
    """core/fortran_bernoulli.f90:61 div(d, u, iflag)."""
    _same_layout(d, u)
    lib().orc_div(_vptr(d), _vptr(u), int(iflag), *_view_args(d))


def kin(u, ke, ds2):
    """core/fortran_kinenergy.f90:3 kin(u, v, ke, ds2, order) (v unused)."""
    _same_layout(u, ke)
    lib().orc_kin(_vptr(u), _vptr(ke), float(ds2), *_view_args(ke))


def add_laplacian(phi, dphi, coef):
    """core/fortran_dissipation.f90:2 add_laplacian(phi, dphi, coef)."""
    _same_layout(phi, dphi)
    lib().orc_add_laplacian(_vptr(phi), _vptr(dphi), float(coef), *_view_args(dphi))


# --------------------------------------------------------------------------
# orientations (core/variables.py:146-207)

def view(base, d):
    """base is the 'i' array [k, j, i]; return the orientation with d innermost."""
    if d == "i":
        return base
    if d == "j":           # [i, k, j]  (rightswap, core/variables.py:404-407)
        return base.transpose(2, 0, 1)
    if d == "k":           # [j, i, k]
        return base.transpose(1, 2, 0)
    raise ValueError(d)


def flipview(base, d):
    """core/variables.py:184-203: i -> view('j'), j -> view('k'), k -> view('i')."""
    return view(base, {"i": "j", "j": "k", "k": "i"}[d])


GEOMETRY_PERIODIC = {       # (z, y, x) periodicity, core/mpi/topology.py:84-157
    "closed": (False, False, False),
    "perio_x": (False, False, True),
    "perio_y": (False, True, False),
    "perio_xy": (False, True, True),
    "perio_xyz": (True, True, True),
}
# mgfor topology enum (core/mgfor/mg_enums.f90:6); perio_x / perio_y are the
# mappings core/model_les.py:80-88 lacks (it asserts), mgfor supports them.
GEOMETRY_MG_TOPOLOGY = {"closed": 1, "perio_x": 2, "perio_y": 3, "perio_xy": 5, "perio_xyz": 6}


def neighbours_of(geometry):
    """Keys of the one-rank neighbour dict (core/mpi/topology.py:72-157)."""
    perio = GEOMETRY_PERIODIC[geometry]
    keys = []
    for d in itertools.product([-1, 0, 1], repeat=3):
        if d == (0, 0, 0):
            continue
        if all(perio[a] or d[a] == 0 for a in range(3)):
            keys.append(d)
    return keys


def halo_fill(x, geometry, nh):
    """core/mpi/halo.py:140-178 on one rank: every halo cell takes the value
    of the interior band on the opposite side (send-to-self)."""
    perio = GEOMETRY_PERIODIC[geometry]
    size = x.shape
    for d in neighbours_of(geometry):
        oidx, iidx = [], []
        for a in range(3):
            if d[a] == -1:
                oidx.append(slice(0, nh))
                iidx.append(slice(size[a] - 2 * nh, size[a] - nh))
            elif d[a] == 1:
                oidx.append(slice(size[a] - nh, size[a]))
                iidx.append(slice(nh, 2 * nh))
            else:
                lo = nh if perio[a] else 0
                hi = size[a] - nh if perio[a] else size[a]
                oidx.append(slice(lo, hi))
                iidx.append(slice(lo, hi))
        x[tuple(oidx)] = x[tuple(iidx)]


# --------------------------------------------------------------------------

class OracleMG:
    """core/mgfordriver.py:5-105 on top of the C restatement of mgfor."""

    def __init__(self, nx, ny, nz, geometry, nh=3, maxite=20, tol=1e-6, omega=0.9):
        L = lib()
        self.geometry = geometry
        self.nh = nh
        self.handle = L.orc_mg_create(nx, ny, nz, GEOMETRY_MG_TOPOLOGY[geometry],
                                      maxite, tol, omega)
        if not self.handle:
            raise ValueError("mgfor coarsening never reaches a 2-cell level for "
                             "(%d, %d)" % (nx, ny))
        self.nlevels = L.orc_mg_nlevels(self.handle)
        self.shape = self.level_shape(0)
        perio = GEOMETRY_PERIODIC[geometry]
        # get_idx_from_neighbours, core/mgfordriver.py:36-48
        self.idx = tuple(slice(0, None) if perio[a] else slice(nh, -nh) for a in range(3))
        self.stats = {"normb": 0, "res": [0], "blowup": False, "nite": []}

    def level_shape(self, lev):
        s = (ctypes.c_int * 3)()
        lib().orc_mg_shape(self.handle, lev, s)
        return (s[0], s[1], s[2])

    def array(self, lev, ivar):
        """numpy view of an mgfor array (ivar numbering of core/mgfor/pytools.f90)."""
        shape = self.level_shape(lev)
        if ivar == 7:
            p = lib().orc_mg_msk(self.handle, lev)
            return np.ctypeslib.as_array(p, shape=shape)
        p = lib().orc_mg_array(self.handle, lev, ivar)
        return np.ctypeslib.as_array(p, shape=shape)

    @property
    def x(self):
        return self.array(0, 1)

    @property
    def b(self):
        return self.array(0, 2)

    def solve(self):
        nite = lib().orc_mg_solve(self.handle)
        self.stats["nite"].append(nite)
        return nite

    def solve_directly(self, p, divarr, dx, halo=None):
        """core/mgfordriver.py:66-78."""
        if halo is not None:
            halo(divarr)
        self.b[self.idx] = divarr
        self.solve()
        p[:, :, :] = self.x[self.idx] * dx ** 2

    def last(self):
        L = lib()
        return (L.orc_mg_last_nite(self.handle), L.orc_mg_last_res(self.handle),
                L.orc_mg_last_normb(self.handle))

    def __del__(self):
        try:
            if self.handle:
                lib().orc_mg_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class OracleLES:
    """The LES model of core/model_les.py with the RK3_SSP scheme of
    core/timescheme.py, on one rank, in the reference array layout."""

    PROGNOSTIC = ("b", "u_i", "u_j", "u_k")

    def __init__(self, nx, ny, nz, Lx=1.0, Ly=1.0, Lz=1.0, geometry="closed", nh=3,
                 rotating=False, coriolis=1.0, mg_maxite=20, mg_tol=1e-6, mg_omega=0.9,
                 viscosity=0.0, n_tracers=0, model="LES", upwind_a=0, upwind_vf=0):
        self.nx, self.ny, self.nz, self.nh = nx, ny, nz, nh
        # param.model["modelname"] (core/nyles.py:92-100): "LES", "linear"
        # (LES(linear=True), core/model_les.py:33-34) or "Euler3d"
        # (core/model_les_euler.py)
        if model not in ("LES", "linear", "Euler3d"):
            raise ValueError("model must be LES, linear or Euler3d")
        self.model = model
        self.nonlinear = model != "linear"
        # the reference's linear upwind branches (linear=.false. there:
        # core/fortran_upwind.f90:30, core/fortran_vortex_force.f90:28,104),
        # of order orderA for the tracers and orderVF for the vortex force;
        # 0 is the WENO branch the reference runs
        if upwind_a not in (0, 1, 3, 5) or upwind_vf not in (0, 1, 3, 5):
            raise ValueError("linear upwind orders are 1, 3 or 5 (0: WENO)")
        self.upwind_a, self.upwind_vf = upwind_a, upwind_vf
        # core/model_les.py:36-43: passive tracers t0, t1, ... after b
        self.traclist = ["b"] + ["t%d" % q for q in range(n_tracers)]
        self.PROGNOSTIC = tuple(self.traclist) + ("u_i", "u_j", "u_k")
        # core/model_les.py:54-57: diff_coef['u'] is the viscosity
        self.viscosity = viscosity
        self.geometry = geometry
        perio = GEOMETRY_PERIODIC[geometry]
        self.perio = perio
        self.shape = tuple(n + 2 * nh if p else n for n, p in zip((nz, ny, nx), perio))
        # core/grid.py:30-53
        self.dx = Lx / nx
        self.dy = Ly / ny
        self.dz = Lz / nz
        self.idx2 = 1 / self.dx ** 2
        self.idy2 = 1 / self.dy ** 2
        self.idz2 = 1 / self.dz ** 2
        self.ids2 = {"i": self.idx2, "j": self.idy2, "k": self.idz2}
        # core/model_les.py:61-67
        self.fparameter = coriolis * (self.dx * self.dy) if rotating else 0.0
        z = lambda: np.zeros(self.shape)
        self.b, self.p, self.ke, self.div = z(), z(), z(), z()
        self.tracers = {t: z() for t in self.traclist[1:]}
        self.u = {d: z() for d in "ijk"}
        self.U = {d: z() for d in "ijk"}
        self.vor = {d: z() for d in "ijk"}
        self.ds = [self._new_tendency() for _ in range(3)]
        self.mg = OracleMG(nx, ny, nz, geometry, nh, mg_maxite, mg_tol, mg_omega)
        self.nite_log = []
        # user forcing: forcing(model, dstate, t) adds to the tendencies, as
        # forcing.add(state, dstate, t) does in core/model_les.py:143-144
        self.forcing = None

    def _new_tendency(self):
        t = {name: np.zeros(self.shape) for name in self.traclist}
        t["u"] = {d: np.zeros(self.shape) for d in "ijk"}
        return t

    # -- state access ----------------------------------------------------
    def get(self, name, ds=None):
        if ds is None:
            if name == "b":
                return self.b
            if name in self.tracers:
                return self.tracers[name]
            return self.u[name[-1]]
        if name in self.traclist:
            return ds[name]
        return ds["u"][name[-1]]

    def fill(self, x):
        halo_fill(x, self.geometry, self.nh)

    # -- core/model_les.py:129-144 ----------------------------------------
    def rhs(self, ds, last=False, forcing=None, t=0.0):
        for name in self.traclist:               # reset_state, core/model_les.py:167-175
            ds[name][...] = 0.0
        for d in "ijk":
            ds["u"][d][...] = 0.0
        # Tracer_numerics.rhstrac, core/tracer.py:59-72 (no diffusion: LES calls
        # it without `last`, core/model_les.py:133); Euler3d advects nothing
        # (the call is commented out, core/model_les_euler.py:131)
        for name in (self.traclist if self.model != "Euler3d" else ()):
            for d in "ijk":
                velocity = view(self.U[d], d)
                field = view(self.get(name), d)
                dfield = view(ds[name], d)
                if d == "i":
                    dfield[...] = 0.0
                if self.upwind_a:
                    upwind_linear(field, velocity, dfield, self.upwind_a)
                else:
                    upwind(field, velocity, dfield)
        # vortex_force, core/vortex_force.py:69-81 (not in the linear model,
        # core/model_les.py:135-136)
        for k, j, i in (["ikj", "jik", "kji"] if self.nonlinear else []):
            u_i = flipview(ds["u"][i], j)
            u_k = flipview(ds["u"][k], j)
            U_i = flipview(self.U[i], j)
            U_k = flipview(self.U[k], j)
            w_j = flipview(self.vor[j], j)
            if self.upwind_vf:
                vortex_force_direc_linear(U_k, w_j, u_i, self.upwind_vf)
                vortex_force_flip_linear(U_i, w_j, u_k, self.upwind_vf)
            else:
                vortex_force_direc(U_k, w_j, u_i)
                vortex_force_flip(U_i, w_j, u_k)
        # bernoulli, core/bernoulli.py:14-23; Euler3d: bernoulli_euler, the
        # gradient of ke only (core/bernoulli.py:26-32)
        for d in "ijk":
            du = view(ds["u"][d], d)
            ke = view(self.ke, d)
            if d in "ij" or self.model == "Euler3d":
                gradke(ke, du)
            else:
                gradkeandb(ke, view(self.b, d), du, self.dz)
        # viscosity at the last stage, core/model_les.py:140-141 with
        # core/viscosity.py:3-9
        if last and self.viscosity != 0.0:
            for d in "ijk":
                coef = self.viscosity * self.ids2[d]
                for comp in "ijk":
                    add_laplacian(view(self.u[comp], d), view(ds["u"][comp], d), coef)
        # user forcing, core/model_les.py:143-144
        if forcing is not None:
            forcing(self, ds, t)

    # -- core/cov_to_contra.py:4-20 ----------------------------------------
    def U_from_u(self):
        self.U["i"][:] = self.u["i"] * self.idx2
        self.U["j"][:] = self.u["j"] * self.idy2
        self.U["k"][:] = self.u["k"] * self.idz2

    # -- core/projection.py:15-87 ------------------------------------------
    def compute_div(self):
        for count, d in enumerate("ijk"):
            div(view(self.div, d), view(self.U[d], d), count)

    def compute_p(self):
        self.compute_div()
        self.mg.solve_directly(self.p, self.div, self.dx, halo=self.fill)
        self.nite_log.append(self.mg.last()[0])
        for d in "ijk":
            gradke(view(self.p, d), view(self.u[d], d))

    # -- core/vorticity.py:7-34 --------------------------------------------
    def vorticity(self):
        perm = {"i": ("k", "j"), "j": ("i", "k"), "k": ("j", "i")}
        for dirk in "ijk":
            dirj, diri = perm[dirk]
            ui = flipview(self.u[diri], dirk)
            uj = flipview(self.u[dirj], dirk)
            wk = flipview(self.vor[dirk], dirk)
            vorticity_k(ui, uj, wk)
            if (self.fparameter > 0.0) and (dirk == "k"):
                wk[:, :-1, :-1] += self.fparameter

    # -- core/kinenergy.py:7-24 --------------------------------------------
    def kinenergy(self):
        for d in "ijk":
            u = view(self.u[d], d)
            ke = view(self.ke, d)
            if d == "i":
                ke[...] = 0.0
            kin(u, ke, self.ids2[d])

    # -- core/model_les.py:97-127 ------------------------------------------
    def diagnose_var(self):
        self.fill(self.b)
        # the reference refills only b and u here (core/model_les.py:100-101),
        # so a passive tracer's periodic halo is never refreshed; the oracle
        # refills it like b's (the periodic image), the device's behaviour
        for t in self.tracers.values():
            self.fill(t)
        for d in "ijk":
            self.fill(self.u[d])
        self.U_from_u()
        self.compute_p()
        for d in "ijk":
            self.fill(self.u[d])
        self.U_from_u()
        if self.nonlinear:                       # core/model_les.py:122-127
            self.vorticity()
            self.kinenergy()
            for d in "ijk":
                self.fill(self.vor[d])
            self.fill(self.ke)

    # -- core/timescheme.py:180-221 ----------------------------------------
    def _pairs(self, *ds):
        out = []
        for name in self.PROGNOSTIC:
            out.append((self.get(name),) + tuple(self.get(name, d) for d in ds))
        return out

    def rk3_ssp(self, dt, t=0.0):
        ds0, ds1, ds2 = self.ds
        self.rhs(ds0, last=False, forcing=self.forcing, t=t)
        for s, d0 in self._pairs(ds0):
            s += dt * d0
        self.diagnose_var()
        self.rhs(ds1, last=False, forcing=self.forcing, t=t + dt)
        for s, d0, d1 in self._pairs(ds0, ds1):
            s += (dt / 4.) * (d1 - 3 * d0)
        self.diagnose_var()
        self.rhs(ds2, last=True, forcing=self.forcing, t=t + dt * 0.5)
        for s, d0, d1, d2 in self._pairs(ds0, ds1, ds2):
            s += (dt / 12.) * (8 * d2 - d0 - d1)
        self.diagnose_var()

    # -- core/timescheme.py:113-122 ----------------------------------------
    def euler_forward(self, dt, t=0.0):
        ds = self.ds[0]
        self.rhs(ds, last=True, forcing=self.forcing, t=t)
        for s, d in self._pairs(ds):
            s += dt * d
        self.diagnose_var()

    # -- core/timescheme.py:125-177 ----------------------------------------
    def lfam3(self, dt, t=0.0):
        if not hasattr(self, "lf_first"):
            self.lf_first = True
            self.stateb = self._new_tendency()
            self.staten = self._new_tendency()
        ds = self.ds[0]
        self.rhs(ds, forcing=self.forcing, t=t)
        if self.lf_first:
            for s, d, sb, sn in self._pairs(ds, self.stateb, self.staten):
                sn[:] = s
                sb[:] = s
                s += dt * d
            self.lf_first = False
            self.diagnose_var()
        else:
            for s, d, sb, sn in self._pairs(ds, self.stateb, self.staten):
                sn[:] = s
                s[:] = sb + (2. * dt) * d
                s[:] = (1. / 12.) * (5. * s + 8. * sn - sb)
                sb[:] = sn
            self.diagnose_var()
            self.rhs(ds, last=True, forcing=self.forcing, t=t + dt * .5)
            for s, d, sn in self._pairs(ds, self.staten):
                s[:] = sn + dt * d
            self.diagnose_var()

    # -- core/nyles.py:227-260 ---------------------------------------------
    def compute_dt(self, cfl, dt_max):
        U, V, W = self.U["i"], self.U["j"], self.U["k"]
        U_max = np.sqrt(np.max(U ** 2 + V ** 2 + W ** 2))
        if U_max == 0.0:
            return dt_max
        return min(cfl / U_max, dt_max)

    # -- helpers -----------------------------------------------------------
    def interior(self, a):
        nh = self.nh
        sl = tuple(slice(nh, -nh) if p else slice(None) for p in self.perio)
        return a[sl]

    def set_interior(self, a, values):
        self.interior(a)[...] = values


def set_threads(n):
    lib().orc_set_num_threads(int(n))


def num_threads():
    return lib().orc_num_threads()
