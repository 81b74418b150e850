"""The five benchmark configurations at their stated sizes (GPU).

C1  lock exchange 64x32x32 perio_x, RK3, fixed dt, 100 steps: FP64 device
    path against the oracle, bit for bit, with the V-cycle count of every
    projection equal.
C2  Taylor-Green 128^3 perio_xyz: FP64 against the oracle bit for bit over
    3 steps; mixed against FP64 over 200 steps (KE, u and b within 1e-5
    relative).
C3  Rayleigh-Taylor 256x256x512 perio_xy, projection to |r|/|b| <= 1e-10
    (squared tolerance 1e-20): 10 steps in both modes, every solve at the
    tolerance, b of the two modes within 1e-6 relative (L2).
C4  shear-layer ensemble, sizes 32..96 in steps of 8 (log-uniform draw):
    FP64 batch bit for bit against the members stepped alone; mixed batch
    within 1e-5 of the single-member mixed path (b and the velocity in L2).
C5  served jobs of every kind and sizes up to 256 are covered by
    tests/test_gpu_serving.py and tests/test_gpu_mg_sizes.py.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

STATE_KEYS = ("b", "u_i", "u_j", "u_k")


@pytest.fixture(scope="module")
def cmp():
    import compare
    return compare


def _upload(dev, st):
    from nyles_amd import _lib as L
    for k, f in zip(STATE_KEYS, (L.NYL_F_B, L.NYL_F_U, L.NYL_F_V, L.NYL_F_W)):
        dev.upload(f, st[k])


def _device(n3, geometry, Lxyz, precision, **kw):
    import compare
    from nyles_amd.device import DeviceModel
    nx, ny, nz = n3
    return DeviceModel(nx, ny, nz, compare.GEOMETRY_XYZ[geometry], Lxyz[0] / nx, Lxyz[1] / ny,
                       Lxyz[2] / nz, precision=precision, **kw)


def test_c1_lock_exchange_100_steps_bitwise(cmp):
    from nyles_amd import workloads as W
    print("case C1 lock exchange 64x32x32, 100 steps", flush=True)
    st, meta = W.lock_exchange()
    Lx, Ly, Lz = meta["L"]
    orc, dev = cmp.make_pair(64, 32, 32, "perio_x", Lx=Lx, Ly=Ly, Lz=Lz)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    dt = 0.5 * orc.dx / np.sqrt(1.0 * Lz)    # 0.5 dx / sqrt(db Lz), db = 1
    for n in range(100):
        orc.rk3_ssp(dt)
        dn = dev.step_rk3(dt)
        # the V-cycle count of the step's three projections
        assert dn["nite"] == sum(orc.nite_log[-3:]), (n, dn, orc.nite_log[-3:])
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in want:
        bad = cmp.mismatches(got[k], want[k])
        assert bad == 0, "%s: %d elements differ after 100 steps" % (k, bad)
    # the gravity current has started: the state is not at rest
    assert np.abs(got["u_i"]).max() > 1e-3
    dev.close()


def test_c2_taylor_green_128_fp64_bitwise(cmp):
    from nyles_amd import workloads as W
    print("case C2 Taylor-Green 128^3 fp64 vs oracle, 3 steps", flush=True)
    st, meta = W.taylor_green(128, seed=2)
    L = meta["L"][0]
    orc, dev = cmp.make_pair(128, 128, 128, "perio_xyz", Lx=L, Ly=L, Lz=L)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    dt = orc.compute_dt(cfl=0.5, dt_max=1e30)
    assert dt == 0.5 / np.sqrt(dev.reduce(0))
    for n in range(3):
        orc.rk3_ssp(dt)
        dev.step_rk3(dt)
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in want:
        bad = cmp.mismatches(got[k], want[k])
        assert bad == 0, "%s: %d elements differ" % (k, bad)
    dev.close()


def _errors(a, r):
    """relative L2 errors of model a against model r: b, the velocity (all
    three components together) and KE"""
    from nyles_amd import _lib as L
    out = {}
    x, y = a.download(L.NYL_F_B), r.download(L.NYL_F_B)
    den = np.linalg.norm(y)
    out["b"] = float(np.linalg.norm(x - y) / den) if den > 0 else float(np.linalg.norm(x))
    du2 = u2 = 0.0
    for f in (L.NYL_F_U, L.NYL_F_V, L.NYL_F_W):
        x, y = a.download(f), r.download(f)
        du2 += float(np.sum((x - y) ** 2))
        u2 += float(np.sum(y ** 2))
    out["velocity"] = (du2 / u2) ** 0.5
    ka, kr = a.reduce(L.NYL_RED_KE), r.reduce(L.NYL_RED_KE)
    out["ke"] = abs(ka - kr) / abs(kr)
    return out


def test_c2_taylor_green_128_mixed_200_steps():
    from nyles_amd import _lib as L
    from nyles_amd import workloads as W
    print("case C2 Taylor-Green 128^3 mixed vs fp64, 200 steps", flush=True)
    st, meta = W.taylor_green(128, seed=2)
    devs = [_device((128,) * 3, "perio_xyz", meta["L"], p) for p in (0, 1)]
    for d in devs:
        _upload(d, st)
        d.diagnose()
    dt = 0.5 / np.sqrt(devs[0].reduce(L.NYL_RED_MAXU2))
    for n in range(200):
        s64 = devs[0].step_rk3(dt)
        sm = devs[1].step_rk3(dt)
        assert s64["res"] < 1e-6 and sm["res"] < 1e-6
    e = _errors(devs[1], devs[0])
    print("  errors after 200 steps", {k: "%.2e" % v for k, v in e.items()}, flush=True)
    for k in ("velocity", "ke"):
        assert e[k] < 1e-5, (k, e)
    assert devs[1].reduce(L.NYL_RED_NONFINITE) == 0
    for d in devs:
        d.close()


def test_c3_rayleigh_taylor_full_size():
    from nyles_amd import _lib as L
    from nyles_amd import workloads as W
    print("case C3 Rayleigh-Taylor 256x256x512, 10 steps at |r|/|b| <= 1e-10", flush=True)
    st, meta = W.rayleigh_taylor()
    devs = [_device((256, 256, 512), "perio_xy", meta["L"], p, mg_tol=1e-20, mg_maxite=50)
            for p in (0, 1)]
    for d in devs:
        _upload(d, st)
        s = d.diagnose()
        assert s["res"] <= 1e-20, s
    dz = meta["L"][2] / 512
    dt = 0.5 * dz / np.sqrt(meta["L"][2])
    for n in range(10):
        for d in devs:
            s = d.step_rk3(dt)
            # the last of the step's three solves; every solve stops only at
            # the tolerance (maxite 50 is never reached)
            assert s["res"] <= 1e-20, (n, s)
            assert s["nite"] < 3 * 50, (n, s)
    e = _errors(devs[1], devs[0])
    print("  mixed vs fp64 after 10 steps", {k: "%.2e" % v for k, v in e.items()}, flush=True)
    assert e["b"] < 1e-6, e
    assert e["ke"] < 1e-5, e
    for d in devs:
        d.close()


def _c4_sizes(count, seed):
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(32), np.log(96), count))
    return [int(8 * round(v / 8)) for v in n]


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_c4_ensemble_sizes_batched_vs_single(precision):
    from nyles_amd import _lib as L
    from nyles_amd import workloads as W
    from nyles_amd.device import DeviceBatch
    sizes = sorted(set(_c4_sizes(64, 4)) | {40, 48, 56, 72, 80, 88, 96})
    print("case C4 ensemble sizes", sizes, "precision", precision, flush=True)

    def member(n):
        st, meta = W.shear_layer(n, seed=1000 + n)
        d = _device((n,) * 3, "perio_xy", meta["L"], precision)
        _upload(d, st)
        d.diagnose()
        return d

    singles = [member(n) for n in sizes]
    batched = [member(n) for n in sizes]
    dts = [0.5 / np.sqrt(d.reduce(L.NYL_RED_MAXU2)) for d in singles]
    batch = DeviceBatch(batched)
    for step in range(3):
        for d, dt in zip(singles, dts):
            d.step_rk3(dt)
        sts = batch.step_rk3(dts)
        for s in sts:
            assert s["res"] < 1e-6, s
    for q, n in enumerate(sizes):
        if precision == 0:
            for f in (L.NYL_F_B, L.NYL_F_U, L.NYL_F_V, L.NYL_F_W):
                a, r = batched[q].download(f), singles[q].download(f)
                assert np.array_equal(a, r), (n, f)
        else:
            e = _errors(batched[q], singles[q])
            assert e["b"] < 1e-5 and e["velocity"] < 1e-5, (n, e)
    batch.close()
    for d in singles + batched:
        d.close()
