"""The model variants of core/nyles.py:92-100 on the device against the
oracle (GPU), bit for bit in FP64:

* "linear" (LES(linear=True), core/model_les.py:33-34,122-127,135-136): no
  vortex force, vorticity and ke not diagnosed;
* "Euler3d" (core/model_les_euler.py): b and the passive tracers are not
  advected (the rhstrac call is commented out, :131) and Bernoulli has no
  buoyancy term (core/bernoulli.py:26-32).

Both run each stage as tendencies + stage combination inside the library;
RK3_SSP on three geometries, LFAM3 and Euler forward with viscosity and
tracers, mixed precision within 1e-5 (relative L2) of the FP64 oracle, a
script through the reference's host API, and a batch refusing such members.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORE = os.path.join(ROOT, "nyles_amd", "core")


@pytest.fixture(scope="module")
def cmp():
    import compare
    return compare


def _pair(cmp, model, geometry, precision=0, viscosity=0.0, n_tracers=0, seed=9):
    orc, dev = cmp.make_pair(32, 16, 16, geometry, Lx=2.0, Ly=1.0, Lz=1.0,
                             precision=precision, viscosity=viscosity, n_tracers=n_tracers,
                             model=model)
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


@pytest.mark.parametrize("model", ["linear", "Euler3d"])
@pytest.mark.parametrize("geometry", ["closed", "perio_xy", "perio_xyz"])
def test_model_rk3_bitwise(cmp, model, geometry):
    print("case model rk3", model, geometry, flush=True)
    orc, dev = _pair(cmp, model, geometry)
    b0 = cmp.device_fields(dev)["b"].copy()
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
        _same(cmp, dev, orc, "%s step %d" % (model, n))
    got = cmp.device_fields(dev)
    if model == "Euler3d":
        assert np.array_equal(got["b"], b0)
    else:
        assert np.all(got["vor_k"] == 0.0) and np.all(got["ke"] == 0.0)
    dev.close()


@pytest.mark.parametrize("model", ["linear", "Euler3d"])
@pytest.mark.parametrize("scheme", ["ef", "lfam3"])
def test_model_other_schemes_with_viscosity_and_tracers(cmp, model, scheme):
    print("case model", model, scheme, flush=True)
    orc, dev = _pair(cmp, model, "perio_x", viscosity=1e-3, n_tracers=2)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        if scheme == "ef":
            orc.euler_forward(dt)
            dev.step_ef(dt)
        else:
            orc.lfam3(dt)
            dev.step_lfam3(dt)
        _same(cmp, dev, orc, "%s %s step %d" % (model, scheme, n))
    dev.close()


@pytest.mark.parametrize("model", ["linear", "Euler3d"])
def test_model_mixed_close_to_oracle(cmp, model):
    print("case model mixed", model, flush=True)
    orc, dev = _pair(cmp, model, "perio_xyz", precision=1)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    uw = np.concatenate([want[k].ravel() for k in ("u_i", "u_j", "u_k")])
    ug = np.concatenate([got[k].ravel() for k in ("u_i", "u_j", "u_k")])
    e = np.linalg.norm(ug - uw) / np.linalg.norm(uw)
    eb = cmp.rel_l2(got["b"], want["b"])
    print("  velocity rel %.2e  b rel %.2e" % (e, eb), flush=True)
    assert e < 1e-5 and eb < 1e-5
    dev.close()


def test_batch_refuses_model_variants(cmp):
    from nyles_amd import _lib as L
    from nyles_amd.device import DeviceBatch
    print("case model batch refusal", flush=True)
    _, dev = _pair(cmp, "linear", "closed")
    with pytest.raises(L.NylesHipError):
        DeviceBatch([dev])
    dev.close()


@pytest.mark.parametrize("name", ["Euler3d", "linear"])
def test_script_model_names_run_on_the_device(tmp_path, name):
    """The reference's experiment setup (as experiments/taylorgreen/tgv.py,
    modelname "Euler3d") runs unchanged and matches the oracle's model of
    the same name bit for bit; "advection" is refused (the reference cannot
    construct it)."""
    if CORE not in sys.path:
        sys.path.insert(0, CORE)
    import parameters
    import nyles as nyles_module
    import oracle as O
    print("case model script", name, flush=True)
    parameters.InextensibleDict.frozen = False
    param = parameters.UserParameters()
    param.model["modelname"] = name
    param.model["geometry"] = "perio_xyz"
    L_ = 2 * np.pi
    param.model["Lx"] = param.model["Ly"] = param.model["Lz"] = L_
    param.IO["datadir"] = str(tmp_path)
    param.IO["expname"] = "tgv"
    param.IO["variables_in_history"] = ["u", "vor"]
    param.IO["timestep_history"] = 1.0
    param.time["timestepping"] = "LFAM3"
    param.time["auto_dt"] = False
    param.time["dt"] = 0.05
    param.time["tend"] = 3 * 0.05 - 1e-9
    for k in ("global_nx", "global_ny", "global_nz"):
        param.discretization[k] = 32
    nyl = nyles_module.Nyles(param)
    u = nyl.model.state.u["i"].view("i")
    v = nyl.model.state.u["j"].view("i")
    x, y, z = (getattr(nyl.grid, c + "_b").view("i") for c in "xyz")
    u[:] = np.sin(x + 1.2) * np.cos(y + 1.8) * np.cos(z + 0.5)
    v[:] = -np.cos(x + 1.2) * np.sin(y + 1.8) * np.cos(z + 0.5)
    u0, v0 = u.copy(), v.copy()
    nyl.run()
    parameters.InextensibleDict.frozen = False
    st = nyl.model.state
    orc = O.OracleLES(32, 32, 32, L_, L_, L_, "perio_xyz", model=name)
    orc.set_interior(orc.u["i"], st.u["i"].interior(u0))
    orc.set_interior(orc.u["j"], st.u["j"].interior(v0))
    orc.diagnose_var()
    for n in range(nyl.n):
        orc.lfam3(0.05)
    for key, mine, ref in (("u", st.u["i"], orc.u["i"]), ("v", st.u["j"], orc.u["j"]),
                           ("w", st.u["k"], orc.u["k"]), ("ke", st.ke, orc.ke)):
        np.testing.assert_array_equal(mine.view("i"), ref, err_msg=key)
    parameters.InextensibleDict.frozen = False
    param.model["modelname"] = "advection"
    with pytest.raises(NotImplementedError):
        nyles_module.Nyles(param)
    parameters.InextensibleDict.frozen = False
