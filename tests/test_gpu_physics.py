"""Physics hooks on the device against the oracle (GPU), bit for bit in FP64:

* viscosity (diff_coef['u']) added at the last stage of each scheme
  (core/model_les.py:140-141, core/viscosity.py:3-9,
  core/fortran_dissipation.f90:2-34), on every geometry;
* passive tracers advected like b (core/model_les.py:36-43), on every
  geometry (the oracle refills a tracer's periodic halo like b's, as the
  device does);
* the user forcing hook (core/model_les.py:143-144) through an unchanged
  script that attaches ``nyles.model.forcing`` (as
  experiments/forced_convection/rayleighbenard.py does): its host-side
  ``add`` sees each stage's time and its additions reach the state.

Mixed precision with viscosity and tracers stays within 1e-5 (relative L2)
of the FP64 oracle over three steps.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOMETRIES = ["closed", "perio_x", "perio_y", "perio_xy", "perio_xyz"]
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORE = os.path.join(ROOT, "nyles_amd", "core")


@pytest.fixture(scope="module")
def cmp():
    import compare
    return compare


def _pair(cmp, geometry, precision=0, viscosity=0.0, n_tracers=0, seed=7, n=(32, 16, 16)):
    orc, dev = cmp.make_pair(*n, geometry, Lx=2.0, Ly=1.0, Lz=1.0, precision=precision,
                             viscosity=viscosity, n_tracers=n_tracers)
    st = cmp.random_state(orc, seed=seed, amp_u=3e-3)
    rng = np.random.default_rng(seed + 100)
    for q in range(n_tracers):
        st["t%d" % q] = rng.standard_normal(st["b"].shape)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    return orc, dev


def _same(cmp, dev, orc, what):
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in want:
        bad = cmp.mismatches(got[k], want[k])
        assert bad == 0, "%s %s: %d elements differ" % (what, k, bad)


@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_viscosity_rk3_bitwise(cmp, geometry):
    print("case viscosity rk3", geometry, flush=True)
    orc, dev = _pair(cmp, geometry, viscosity=2e-3)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
        _same(cmp, dev, orc, "step %d" % n)
    dev.close()


@pytest.mark.parametrize("scheme", ["ef", "lfam3"])
def test_viscosity_and_tracers_other_schemes_bitwise(cmp, scheme):
    print("case viscosity + tracers", scheme, flush=True)
    orc, dev = _pair(cmp, "perio_x", viscosity=1e-3, n_tracers=2)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        if scheme == "ef":
            orc.euler_forward(dt)
            dev.step_ef(dt)
        else:
            orc.lfam3(dt)
            dev.step_lfam3(dt)
        _same(cmp, dev, orc, "%s step %d" % (scheme, n))
    dev.close()


@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_tracers_rk3_bitwise(cmp, geometry):
    print("case tracers rk3", geometry, flush=True)
    orc, dev = _pair(cmp, geometry, n_tracers=3)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
    _same(cmp, dev, orc, "after 3 steps")
    dev.close()


def test_unfused_stages_match_fused_step(cmp):
    """rhs into the slots + nyl_stage_update per stage (the forced path with
    no forcing) is bitwise the fused step, tracers and viscosity included."""
    from nyles_amd import _lib as L
    print("case unfused stages", flush=True)
    _, fused = _pair(cmp, "perio_xy", viscosity=1e-3, n_tracers=1)
    _, split = _pair(cmp, "perio_xy", viscosity=1e-3, n_tracers=1)
    dt = 0.05
    for n in range(2):
        fused.step_rk3(dt)
        for k, (c, last) in enumerate(((dt, False), (dt / 4., False), (dt / 12., True))):
            split.rhs(k, last)
            split.stage_update(k + 1, c)
            split.diagnose()
    for fid in (L.NYL_F_B, L.NYL_F_U, L.NYL_F_V, L.NYL_F_W, L.NYL_F_TRACER(0), L.NYL_F_KE):
        assert cmp.mismatches(fused.download(fid), split.download(fid)) == 0, fid
    fused.close()
    split.close()


def test_mixed_with_viscosity_and_tracers_close_to_oracle(cmp):
    from nyles_amd import _lib as L
    print("case mixed viscosity + tracers", flush=True)
    orc, dev = _pair(cmp, "perio_xy", precision=L.NYL_PREC_MIXED, viscosity=1e-3, n_tracers=2)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in ("b", "u_i", "u_j", "u_k", "t0", "t1"):
        e = cmp.rel_l2(got[k], want[k])
        print("  %s rel err %.2e" % (k, e), flush=True)
        assert e < 1e-5, (k, e)
    dev.close()


def _forced_script(tmp_path, scheme, steps):
    """a flux-forced buoyancy run written against the reference API: the
    script sets forced = True, a viscosity and one passive tracer, and
    attaches a forcing object whose add() changes the tendencies"""
    if CORE not in sys.path:
        sys.path.insert(0, CORE)
    import parameters
    import nyles as nyles_module
    parameters.InextensibleDict.frozen = False
    param = parameters.UserParameters()
    param.model["modelname"] = "LES"
    param.model["geometry"] = "perio_xy"
    param.model["Lx"] = 2.0
    param.model["Ly"] = 1.0
    param.model["Lz"] = 1.0
    param.model["n_tracers"] = 1
    param.physics["forced"] = True
    param.physics["diff_coef"] = {"u": 1e-3}
    param.IO["datadir"] = str(tmp_path)
    param.IO["expname"] = "forced"
    param.IO["variables_in_history"] = ["b"]
    param.IO["timestep_history"] = 0.1
    param.time["timestepping"] = scheme
    param.time["auto_dt"] = False
    param.time["dt"] = 0.02
    param.time["tend"] = steps * 0.02 - 1e-9
    param.discretization["global_nx"] = 32
    param.discretization["global_ny"] = 16
    param.discretization["global_nz"] = 16
    param.MPI["nh"] = 3
    nyl = nyles_module.Nyles(param)
    parameters.InextensibleDict.frozen = False

    class Forcing(object):
        """bottom heating / top cooling (as Forcing_flux of
        experiments/forced_convection/rayleighbenard.py:114-138), plus a
        source of the tracer at the bottom"""

        def __init__(self, dz, Q):
            self.dz, self.Q = dz, Q
            self.times = []

        def add(self, state, dstate, time):
            self.times.append(time)
            db = dstate.b.view("i")
            db[0, :, :] += self.Q / self.dz
            db[-1, :, :] -= self.Q / self.dz
            dt0 = dstate.t0.view("i")
            dt0[0, :, :] += 1.0

    forcing = Forcing(nyl.grid.dz, 0.1)
    nyl.model.forcing = forcing
    rng = np.random.default_rng(11)
    b = nyl.model.state.b.view("i")
    b[...] = 1e-3 * rng.standard_normal(b.shape)
    b0 = b.copy()
    nyl.model.diagnose_var(nyl.model.state)
    nyl.run()
    return nyl, forcing, b0


@pytest.mark.parametrize("scheme", ["RK3_SSP", "LFAM3"])
def test_forced_script_matches_oracle(tmp_path, scheme):
    import oracle as O
    print("case forced script", scheme, flush=True)
    nyl, forcing, b0 = _forced_script(tmp_path, scheme, steps=3)
    st = nyl.model.state
    assert nyl.n == 3
    orc = O.OracleLES(32, 16, 16, 2.0, 1.0, 1.0, "perio_xy", viscosity=1e-3, n_tracers=1)
    orc.set_interior(orc.b, st.b.interior(b0))
    seen = []

    def forcing_fn(model, ds, t):
        seen.append(t)
        ds["b"][0, :, :] += forcing.Q / forcing.dz
        ds["b"][-1, :, :] -= forcing.Q / forcing.dz
        ds["t0"][0, :, :] += 1.0

    orc.forcing = forcing_fn
    orc.diagnose_var()
    t = 0.0
    for n in range(3):
        if scheme == "RK3_SSP":
            orc.rk3_ssp(0.02, t=t)
        else:
            orc.lfam3(0.02, t=t)
        t += 0.02
    np.testing.assert_allclose(forcing.times, seen, rtol=0, atol=1e-12)
    for name, mine, ref in (("b", st.b, orc.b), ("u", st.u["i"], orc.u["i"]),
                            ("w", st.u["k"], orc.u["k"]), ("t0", st.t0, orc.tracers["t0"])):
        got = mine.interior(mine.view("i"))
        np.testing.assert_array_equal(got, orc.interior(ref), err_msg=name)
    # the forcing reached the state: the tracer source at the bottom
    assert np.all(st.t0.interior(st.t0.view("i"))[0] > 0.0)
