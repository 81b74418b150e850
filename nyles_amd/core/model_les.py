"""LES model (interface of core/model_les.py) running on one GPU.

``LES(param, grid)`` builds the host State (views for scripts), a device
context holding the same fields, and the time scheme bound to it.
``rhs``, ``diagnose_var`` and ``forward`` run in the HIP library; the
host mirrors are synchronised lazily (see variables.py).

Physics hooks, as in the reference: ``diff_coef['u']`` is a viscosity added
at the last stage of a step (core/model_les.py:140-141); ``n_tracers``
passive tracers t0, t1, ... are advected like b (core/model_les.py:36-43);
with ``forced`` the script attaches ``model.forcing``, whose
``add(state, dstate, t)`` runs after each right-hand side on host views of
the state and the tendencies (core/model_les.py:143-144), so the step runs
rhs, forcing and stage update as separate device calls.

Differences from the reference that a script can see:
* geometries perio_x and perio_y are accepted (the reference asserts,
  core/model_les.py:80-88, although its multigrid supports them);
* a passive tracer's periodic halo is refilled with the periodic image like
  b's (the reference refills only b and u, core/model_les.py:100-101, so its
  tracers read stale halo values on periodic axes); on wall axes there is no
  halo and the results are the reference's;
* diffusion coefficients of b and the tracers are accepted and, as in the
  reference's LES (rhstrac is called without ``last``, core/model_les.py:133),
  not applied;
* the Euler3d variant (``LES(..., euler=True)``, core/model_les_euler.py)
  skips the halo fill of b and u before the projection
  (core/model_les_euler.py:98-99), so in the reference the divergence at the
  first interior cell of a periodic axis reads the halo left by the previous
  diagnosis; here the wrapped neighbour is read, as in the LES.  Its missing
  perio_xy case in the multigrid topology (core/model_les_euler.py:84-86
  against core/model_les.py:85-86) is not reproduced either;
* ``forward`` returns True when b or u holds a non-finite value (the
  reference's blow-up flag is never set, core/mgfordriver.py:32).
"""
import os as _os
import sys as _sys
_ROOT = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
if _ROOT not in _sys.path:
    _sys.path.insert(0, _ROOT)
import numpy as np

from nyles_amd import _lib as L
from nyles_amd.device import DeviceModel

from nyles_amd.core import halo
from nyles_amd.core import timescheme as ts
from nyles_amd.core import variables as var
from nyles_amd.core.timing import timing

GEOMETRY_PERIODIC_XYZ = {
    "closed": (0, 0, 0), "perio_x": (1, 0, 0), "perio_y": (0, 1, 0),
    "perio_xy": (1, 1, 0), "perio_xyz": (1, 1, 1),
}


class MGInfo(object):
    """What the reference's MG driver object exposes (core/mgfordriver.py:32)."""

    def __init__(self):
        self.stats = {"normb": 0, "res": [0], "blowup": False, "nite": 0}


class LES(object):
    """linear=True: LES(linear=True) of core/model_les.py (no vortex force,
    vorticity and ke not diagnosed); euler=True: the LES of
    core/model_les_euler.py (b and the tracers not advected, Bernoulli
    without buoyancy)."""

    def __init__(self, param, grid, linear=False, euler=False):
        n_tracers = param["n_tracers"]
        if not 0 <= n_tracers <= L.NYL_MAX_TRACERS:
            raise ValueError("n_tracers must be 0..%d on the GPU path" % L.NYL_MAX_TRACERS)
        self.nonlinear = not linear
        self.euler = euler
        self.grid = grid
        self.param = param
        # core/model_les_euler.py:36 keeps '' in place of b
        self.traclist = ["" if euler else "b"] + ["t%d" % q for q in range(n_tracers)]
        self.state = var.get_state(param, [("t%d" % q, "tracer%d" % q)
                                           for q in range(n_tracers)])
        self.halo = halo.set_halo(param, self.state)
        self.neighbours = param["neighbours"]
        self.timescheme = ts.Timescheme(param, self.state)
        self.timescheme.set(self.rhs, self.diagnose_var)
        self.orderA = param["orderA"]
        self.orderVF = param["orderVF"]
        self.orderKE = param["orderKE"]
        self.rotating = param["rotating"]
        self.forced = param["forced"]
        self.diff_coef = param["diff_coef"]
        self.add_viscosity = "u" in self.diff_coef.keys()
        self.viscosity = self.diff_coef["u"] if self.add_viscosity else 0.0
        self.forcing = None
        if self.rotating:
            self.fparameter = param["coriolis"] * (grid.dx * grid.dy)
        else:
            self.fparameter = 0.0
        precision = {"fp64": L.NYL_PREC_FP64, "mixed": L.NYL_PREC_MIXED}[param["precision"]]
        # the multigrid stop test: mgfor's parameters (the reference's mgfor
        # ignores the multigrid block, core/mgfor/mg_types.f90:15-26), or
        # the multigrid block (tol on |r|/|b|, squared for the test)
        if param["mg_params"] == "multigrid":
            mg_tol, mg_maxite, mg_omega = param["tol"] ** 2, param["maxite"], param["omega"]
        else:
            mg_tol, mg_maxite, mg_omega = param["mg_tol"], param["mg_maxite"], param["mg_omega"]
        self.blowup_check = param["blowup_check"]
        # device.upwind = "linear": the reference's linear upwind branches
        # (core/fortran_upwind.f90:34-67, core/fortran_vortex_force.f90:43-65,
        # 121-144; compiled out there) of orders orderA and orderVF
        upwind_a = upwind_vf = 0
        if param["upwind"] == "linear":
            upwind_a, upwind_vf = self.orderA, self.orderVF
            if upwind_a not in (1, 3, 5) or upwind_vf not in (1, 3, 5):
                raise ValueError("the linear upwind branches run odd orders 1, 3 or 5")
        self.device = DeviceModel(
            param["nx"], param["ny"], param["nz"], GEOMETRY_PERIODIC_XYZ[param["geometry"]],
            grid.dx, grid.dy, grid.dz, idx2=grid.idx2, idy2=grid.idy2, idz2=grid.idz2,
            dx2=grid.dx ** 2, fparameter=self.fparameter, precision=precision,
            mg_maxite=mg_maxite, mg_tol=mg_tol, mg_omega=mg_omega,
            viscosity=float(self.viscosity), n_tracers=n_tracers,
            model=L.NYL_MODEL_EULER if euler else (L.NYL_MODEL_LINEAR if linear
                                                   else L.NYL_MODEL_LES),
            upwind_a=upwind_a, upwind_vf=upwind_vf)
        self.timescheme.bind(self.device, self)
        self.mg = MGInfo()
        st = self.state
        pairs = [(st.b, L.NYL_F_B), (st.u["i"], L.NYL_F_U), (st.u["j"], L.NYL_F_V),
                 (st.u["k"], L.NYL_F_W), (st.ke, L.NYL_F_KE), (st.vor["i"], L.NYL_F_VORX),
                 (st.vor["j"], L.NYL_F_VORY), (st.vor["k"], L.NYL_F_VORZ),
                 (st.div, L.NYL_F_DIV)]
        pairs += [(st.get("t%d" % q), L.NYL_F_TRACER(q)) for q in range(n_tracers)]
        for scalar, fid in pairs:
            scalar.attach(self.device, fid)
        # U is recomputed inside the stage kernel, not stored: fill it on demand
        for d, fid in (("i", L.NYL_F_UCX), ("j", L.NYL_F_UCY), ("k", L.NYL_F_UCZ)):
            st.U[d].attach(self.device, fid,
                           before_pull=lambda: self.device.apply(L.NYL_OP_U_FROM_U))
        st.p.attach(self.device, L.NYL_F_P,
                    before_pull=lambda: self.device.apply(L.NYL_OP_PRESSURE))
        for slot, ds in enumerate(self._tendency_states()):
            base = L.NYL_F_DS0_B + 4 * slot
            for k, name in enumerate(("b", "u_i", "u_j", "u_k")):
                ds.get(name).attach(self.device, base + k)
            for q in range(n_tracers):
                ds.get("t%d" % q).attach(self.device, L.NYL_F_DS_TRACER(slot, q))
        self.stats = []

    # -- host <-> device ------------------------------------------------------
    def _tendency_states(self):
        """the host States of the device's tendency slots 0.."""
        ts_ = self.timescheme
        if ts_.timestepping == "RK3_SSP":
            return [ts_.ds0, ts_.ds1, ts_.ds2]
        return [ts_.dstate]

    def _attached(self):
        out = self.state.scalars()
        for ds in self._tendency_states():
            out += [s for s in ds.scalars() if s._dev is not None]
        return out

    def push(self):
        for s in self._attached():
            s.push()

    def mark_device_newer(self):
        for s in self._attached():
            s.mark_device_newer()

    def _update_stats(self, st):
        self.mg.stats["nite"] = st["nite"]
        self.mg.stats["res"] = [st["res"]]
        self.mg.stats["normb"] = st["normb"]

    # -- reference interface --------------------------------------------------
    @timing
    def diagnose_var(self, state):
        self.push()
        st = self.device.diagnose()
        self.mark_device_newer()
        self._update_stats(st)

    @timing
    def rhs(self, state, t, dstate, last=False):
        """Tendencies of the current state (viscosity when last); into the RK3
        slot that ``dstate`` is, slot 0 otherwise; then the user forcing on
        the host views (core/model_les.py:129-144)."""
        slot = 0
        for k, name in enumerate(("ds0", "ds1", "ds2")):
            if getattr(self.timescheme, name, None) is dstate:
                slot = k
        self.push()
        self.device.rhs(slot, last)
        self.mark_device_newer()
        if self.forced:
            if self.forcing is None:
                raise RuntimeError("forced = True: attach the forcing to model.forcing")
            self.forcing.add(state, dstate, t)

    @timing
    def forward(self, t, dt):
        self.push()
        self.timescheme.forward(self.state, t, dt)
        self.mark_device_newer()
        self._update_stats(self.timescheme.stats)
        if self.blowup_check and self.device.reduce(L.NYL_RED_NONFINITE) > 0:
            self.mg.stats["blowup"] = True
        return self.mg.stats["blowup"]

    def max_U2(self):
        """max over cells of U**2+V**2+W**2 (core/nyles.py:249), on the device."""
        self.push()
        return self.device.reduce(L.NYL_RED_MAXU2)

    def update_stats(self):
        self.stats.append(dict(self.mg.stats))

    def write_stats(self, path):
        import pickle
        with open("%s/stats.pkl" % path, "wb") as fid:
            pickle.dump(self.stats, fid)


def reset_state(state):
    for s in state.scalars():
        s.view()[...] = 0.0
