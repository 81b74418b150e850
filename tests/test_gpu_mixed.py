"""Mixed precision path (FP32 stencils and inner V-cycles, FP64 outer
residual) against the FP64 oracle (GPU).

Tolerances, stated per test: the mixed path's target is KE and buoyancy
within 1e-5 relative of FP64 after 200 steps of the 256^3 Taylor-Green
workload (measured by bench.py); here smaller grids and fewer steps check
the same bound and the solver's behaviour.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOMETRIES = ["closed", "perio_x", "perio_y", "perio_xy", "perio_xyz"]


@pytest.fixture(scope="module")
def cmp():
    import compare
    return compare


@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_defect_correction_matches_fp64_solve(cmp, geometry):
    """FP32 V-cycles inside an FP64 outer residual reproduce the FP64 V-cycle
    iteration to FP32 precision: same cycle count (+-1 at the stop test) and the same
    solution to 1e-5 relative (up to the additive constant of the singular
    problem)."""
    from nyles_amd import _lib as L
    print("case mixed mg", geometry, flush=True)
    nx, ny, nz = 128, 64, 32
    orc, dev64 = cmp.make_pair(nx, ny, nz, geometry)
    _, devm = cmp.make_pair(nx, ny, nz, geometry, precision=L.NYL_PREC_MIXED)
    rng = np.random.default_rng(3)
    for rep in range(2):
        r = rng.standard_normal((nz, ny, nx))
        r -= r.mean()
        for d in (dev64, devm):
            d.upload(L.NYL_F_MGB, r)
        s64 = dev64.mg_solve()
        sm = devm.mg_solve()
        assert abs(sm["nite"] - s64["nite"]) <= 1, (rep, sm, s64)
        assert sm["res"] < 1e-6
        x64 = dev64.download(L.NYL_F_MGX)
        xm = devm.download(L.NYL_F_MGX)
        err = cmp.rel_l2(xm - xm.mean(), x64 - x64.mean())
        print("  rep", rep, "nite", s64["nite"], sm["nite"], "x rel err %.2e" % err, flush=True)
        assert err < 1e-5
    dev64.close()
    devm.close()


@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_mixed_steps_close_to_oracle(cmp, geometry):
    """Three RK3 steps on a random state: state fields within 1e-5 relative
    (L2) of the FP64 oracle."""
    from nyles_amd import _lib as L
    print("case mixed rk3", geometry, flush=True)
    nx, ny, nz = 64, 32, 32
    orc, dev = cmp.make_pair(nx, ny, nz, geometry, Lx=4.0, Ly=2.0, Lz=2.0,
                             precision=L.NYL_PREC_MIXED)
    st = cmp.random_state(orc, seed=7, amp_u=3e-3)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in ("b", "u_i", "u_j", "u_k"):
        e = cmp.rel_l2(got[k], want[k])
        print("  %s rel err %.2e" % (k, e), flush=True)
        assert e < 1e-5, (k, e)
    ke_o = float(np.sum(want["ke"]))
    ke_d = dev.reduce(L.NYL_RED_KE)
    assert abs(ke_d - ke_o) <= 1e-5 * abs(ke_o)
    dev.close()


def _tgv_pair(n):
    from nyles_amd import _lib as L
    from nyles_amd import workloads
    from nyles_amd.device import DeviceModel
    st, meta = workloads.taylor_green(n, seed=2)
    Lx = meta["L"][0]
    dx = Lx / n
    devs = []
    for prec in (L.NYL_PREC_FP64, L.NYL_PREC_MIXED):
        d = DeviceModel(n, n, n, (1, 1, 1), dx, dx, dx, idx2=1 / dx ** 2, idy2=1 / dx ** 2,
                        idz2=1 / dx ** 2, dx2=dx ** 2, precision=prec)
        for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                       ("u_k", L.NYL_F_W)):
            d.upload(fid, st[k])
        d.diagnose()
        devs.append(d)
    return devs


def test_mixed_taylor_green_ke_error():
    """Taylor-Green 64^3, 100 RK3 steps at CFL 0.5: KE within 1e-5 relative and
    velocity within 1e-4 relative (L2) of the FP64 path (which equals the
    oracle bit for bit)."""
    from nyles_amd import _lib as L
    print("case mixed tgv 64", flush=True)
    d64, dm = _tgv_pair(64)
    dt = 0.5 / float(np.sqrt(d64.reduce(L.NYL_RED_MAXU2)))
    for n in range(100):
        d64.step_rk3(dt)
        dm.step_rk3(dt)
    k64, km = d64.reduce(L.NYL_RED_KE), dm.reduce(L.NYL_RED_KE)
    import compare
    eu = max(compare.rel_l2(dm.download(f), d64.download(f))
             for f in (L.NYL_F_U, L.NYL_F_V, L.NYL_F_W))
    print("  ke %.12e %.12e rel %.2e; u rel %.2e" % (k64, km, abs(km - k64) / k64, eu),
          flush=True)
    assert abs(km - k64) <= 1e-5 * k64
    assert eu < 1e-4
    assert dm.reduce(L.NYL_RED_NONFINITE) == 0
    d64.close()
    dm.close()


def test_mixed_lock_exchange_buoyancy_error(cmp):
    """Lock exchange 64x32x32 (perio_x, walls in y and z), 50 RK3 steps at a
    fixed dt: buoyancy within 1e-5 relative (L2) of FP64, and its domain sum
    conserved to FP32 rounding."""
    from nyles_amd import _lib as L
    print("case mixed lock exchange", flush=True)
    nx, ny, nz = 64, 32, 32
    _, d64 = cmp.make_pair(nx, ny, nz, "perio_x", Lx=16.0, Ly=8.0, Lz=8.0)
    _, dm = cmp.make_pair(nx, ny, nz, "perio_x", Lx=16.0, Ly=8.0, Lz=8.0,
                          precision=L.NYL_PREC_MIXED)
    rng = np.random.default_rng(1)
    x = (np.arange(nx) + 0.5) * 16.0 / nx
    b = np.where(x < 8.0, 0.5, -0.5)[None, None, :] + 1e-3 * rng.uniform(-1, 1, (nz, ny, nx))
    z = np.zeros((nz, ny, nx))
    for d in (d64, dm):
        d.upload(L.NYL_F_B, b)
        for f in (L.NYL_F_U, L.NYL_F_V, L.NYL_F_W):
            d.upload(f, z)
        d.diagnose()
    dt = 0.5 * (16.0 / nx) / np.sqrt(1.0 * 8.0)
    for n in range(50):
        d64.step_rk3(dt)
        dm.step_rk3(dt)
    b64, bm = d64.download(L.NYL_F_B), dm.download(L.NYL_F_B)
    e = cmp.rel_l2(bm, b64)
    s0 = float(np.sum(b))
    sm = dm.reduce(L.NYL_RED_BSUM)
    print("  b rel err %.2e, b sum %.6e -> %.6e" % (e, s0, sm), flush=True)
    assert e < 1e-5
    assert abs(sm - s0) <= 1e-4 * float(np.sum(np.abs(b)))
    d64.close()
    dm.close()


@pytest.mark.parametrize("batched", [False, True])
def test_mixed_tight_tolerance_rayleigh_taylor(cmp, batched):
    """Rayleigh-Taylor 64x64x128 (perio_xy) with the projection converged to
    |r|/|b| <= 1e-10 (tol 1e-20 on the squared ratio, 60 V-cycles at most):
    the mixed path's FP64 defect correction reaches the same tolerance as the
    FP64 solve at every step, alone and in a batch, and after 20 steps the
    buoyancy is within 1e-6 relative (L2) of FP64."""
    from nyles_amd import _lib as L
    from nyles_amd import workloads
    from nyles_amd.device import DeviceBatch, DeviceModel
    print("case mixed rt tight tol batched=%d" % batched, flush=True)
    st, meta = workloads.rayleigh_taylor(64, 64, 128, nmodes=8)
    Lx, Ly, Lz = meta["L"]
    dx, dy, dz = Lx / 64, Ly / 64, Lz / 128
    dt = 0.5 * dz / np.sqrt(Lz)
    devs = {}
    for prec in (L.NYL_PREC_FP64, L.NYL_PREC_MIXED):
        d = DeviceModel(64, 64, 128, (1, 1, 0), dx, dy, dz, precision=prec,
                        mg_tol=1e-20, mg_maxite=60)
        for k, fid in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V),
                       ("u_k", L.NYL_F_W)):
            d.upload(fid, st[k])
        s = d.diagnose()
        assert s["res"] < 1e-20, (prec, s)
        devs[prec] = d
    d64, dm = devs[L.NYL_PREC_FP64], devs[L.NYL_PREC_MIXED]
    if batched:    # a batch holds members of one precision
        bats = [DeviceBatch([d], autotune=False) for d in (d64, dm)]
    nite = {0: 0, 1: 0}
    for n in range(20):
        if batched:
            ss = [bt.step_rk3([dt])[0] for bt in bats]
        else:
            ss = [d64.step_rk3(dt), dm.step_rk3(dt)]
        for q, s in enumerate(ss):
            assert s["res"] < 1e-20 and s["nite"] < 3 * 60, (n, q, s)
            nite[q] += s["nite"]
    if batched:
        for bt in bats:
            bt.close()
    e = cmp.rel_l2(dm.download(L.NYL_F_B), d64.download(L.NYL_F_B))
    print("  V-cycles fp64 %d mixed %d, b rel err %.2e" % (nite[0], nite[1], e), flush=True)
    assert e < 1e-6
    d64.close()
    dm.close()
