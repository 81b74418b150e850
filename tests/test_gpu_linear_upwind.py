"""The reference's linear upwind branches on the device (GPU), bit for bit
in FP64 against the oracle: core/fortran_upwind.f90:34-67 with
core/interpolate_tracer.f90 for b and the passive tracers (order orderA),
core/fortran_vortex_force.f90:43-65,121-144 with core/interpolate.f90 for
the vortex force (order orderVF).  The shipped reference compiles these
branches out (linear=.false.); here they are an option
(nyl_desc.upwind_a / upwind_vf, device.upwind = "linear" in the host API),
run as tendencies + stage combination.  Odd orders 1, 3, 5; wall closures
on closed axes, the interior stencils wrapped on periodic ones.
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


def _pair(cmp, geometry, oa, ov, precision=0, n_tracers=0, seed=13, n=(32, 16, 16)):
    orc, dev = cmp.make_pair(*n, geometry, Lx=2.0, Ly=1.0, Lz=1.0, precision=precision,
                             n_tracers=n_tracers, upwind_a=oa, upwind_vf=ov)
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


@pytest.mark.parametrize("orders", [(5, 5), (3, 1), (1, 3)], ids=["5-5", "3-1", "1-3"])
@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_xyz"])
def test_linear_upwind_rk3_bitwise(cmp, geometry, orders):
    print("case linear upwind rk3", geometry, orders, flush=True)
    orc, dev = _pair(cmp, geometry, *orders)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
        _same(cmp, dev, orc, "%s %s step %d" % (geometry, orders, n))
    dev.close()


def test_linear_upwind_lfam3_with_tracers_bitwise(cmp):
    print("case linear upwind lfam3 tracers", flush=True)
    orc, dev = _pair(cmp, "perio_xy", 5, 3, n_tracers=2)
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.lfam3(dt)
        dev.step_lfam3(dt)
        _same(cmp, dev, orc, "lfam3 step %d" % n)
    dev.close()


def test_linear_upwind_differs_from_weno(cmp):
    """the option changes the result (the branch is really taken)"""
    print("case linear upwind vs weno", flush=True)
    _, lin = _pair(cmp, "closed", 5, 5)
    _, weno = _pair(cmp, "closed", 0, 0)
    dt = 0.02
    lin.step_rk3(dt)
    weno.step_rk3(dt)
    a, b = cmp.device_fields(lin), cmp.device_fields(weno)
    assert not np.array_equal(a["b"], b["b"]) and not np.array_equal(a["u_i"], b["u_i"])
    lin.close()
    weno.close()


def test_linear_upwind_mixed_close_to_oracle(cmp):
    print("case linear upwind mixed", flush=True)
    orc, dev = _pair(cmp, "perio_xyz", 5, 5, precision=1)
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


def test_host_api_upwind_option(tmp_path):
    """device.upwind = "linear" takes orderA and orderVF; even orders and
    short wall axes are refused."""
    if CORE not in sys.path:
        sys.path.insert(0, CORE)
    import parameters
    import nyles as nyles_module
    from nyles_amd import _lib as L
    print("case linear upwind host api", flush=True)
    parameters.InextensibleDict.frozen = False
    param = parameters.UserParameters()
    param.IO["datadir"] = str(tmp_path)
    for k in ("global_nx", "global_ny", "global_nz"):
        param.discretization[k] = 16
    param.device["upwind"] = "linear"
    param.discretization["orderA"] = 3
    param.discretization["orderVF"] = 5
    nyl = nyles_module.Nyles(param)
    d = nyl.model.device.desc
    assert (d.upwind_a, d.upwind_vf) == (3, 5)
    nyl.model.device.close()
    parameters.InextensibleDict.frozen = False
    param.discretization["orderA"] = 4
    with pytest.raises(ValueError):
        nyles_module.Nyles(param)
    parameters.InextensibleDict.frozen = False
    from nyles_amd.device import DeviceModel
    with pytest.raises(L.NylesHipError):
        DeviceModel(32, 4, 16, (1, 0, 1), 0.1, 0.1, 0.1, upwind_a=5)
    parameters.InextensibleDict.frozen = False
