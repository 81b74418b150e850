"""Bitwise parity of the FP64 HIP time step with the oracle (GPU).

Every comparison is exact: the FP64 kernels reproduce the reference's
operation order, so any differing element is a bug.  MG norms are reduced in
a different order than mgfor's sequential loop, so the residual values are
compared to a relative 1e-12 and the V-cycle counts exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GEOMETRIES = ["closed", "perio_x", "perio_y", "perio_xy", "perio_xyz"]


@pytest.fixture(scope="module")
def cmp():
    import compare
    return compare


def _assert_same(name, got, want):
    import compare
    bad = compare.mismatches(got, want)
    assert bad == 0, "%s: %d elements differ (max rel %.3e)" % (
        name, bad, compare.max_rel(got, want))


@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_diagnose_parity(cmp, geometry):
    """LES.diagnose_var: projection (div, MG solve, u -= grad p), U, vor, ke."""
    print("case diagnose", geometry, flush=True)
    orc, dev = cmp.make_pair(32, 16, 16, geometry, Lx=2.0, Ly=1.0, Lz=1.0)
    st = cmp.random_state(orc, seed=11)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    s = dev.diagnose()
    nite, res, normb = orc.mg.last()
    assert s["nite"] == nite
    assert abs(s["res"] - res) <= 1e-12 * max(abs(res), 1e-300)
    assert abs(s["normb"] - normb) <= 1e-12 * normb
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in want:
        _assert_same(k, got[k], want[k])
    dev.close()


@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_rhs_parity_random_inputs(cmp, geometry):
    """LES.rhs on arbitrary (not mutually consistent) u, vor, ke, b fields; the
    oracle's U is u*idx2 (core/cov_to_contra.py:15-17), which the device
    recomputes in the stage kernel."""
    from nyles_amd import _lib as L
    import oracle as O
    print("case rhs", geometry, flush=True)
    orc, dev = cmp.make_pair(16, 16, 32, geometry, Lx=1.0, Ly=1.0, Lz=2.0)
    rng = np.random.default_rng(5)
    shape = (orc.nz, orc.ny, orc.nx)
    fields = {
        "b": (orc.b, L.NYL_F_B), "u_i": (orc.U["i"], L.NYL_F_U), "u_j": (orc.U["j"], L.NYL_F_V),
        "u_k": (orc.U["k"], L.NYL_F_W), "vor_i": (orc.vor["i"], L.NYL_F_VORX),
        "vor_j": (orc.vor["j"], L.NYL_F_VORY), "vor_k": (orc.vor["k"], L.NYL_F_VORZ),
        "ke": (orc.ke, L.NYL_F_KE),
    }
    metric = {"u_i": orc.idx2, "u_j": orc.idy2, "u_k": orc.idz2}
    for name, (arr, fid) in fields.items():
        vals = rng.standard_normal(shape)
        # exact zeros on a few cells exercise the u > 0 / u <= 0 split
        vals.flat[::7] = 0.0
        dev.upload(fid, vals)
        if name in metric:
            vals = vals * metric[name]
        orc.set_interior(arr, vals)
        O.halo_fill(arr, geometry, orc.nh)
    ds = orc.ds[1]
    orc.rhs(ds)
    dev.rhs(1)
    want = {"b": ds["b"], "u_i": ds["u"]["i"], "u_j": ds["u"]["j"], "u_k": ds["u"]["k"]}
    ids = {"b": L.NYL_F_DS1_B, "u_i": L.NYL_F_DS1_U, "u_j": L.NYL_F_DS1_V, "u_k": L.NYL_F_DS1_W}
    for k in want:
        _assert_same("d" + k, dev.download(ids[k]), orc.interior(want[k]))
    dev.close()


MG_SIZES = [(32, 32, 16), (128, 64, 32)]


@pytest.mark.parametrize("size", MG_SIZES, ids=["%dx%dx%d" % s for s in MG_SIZES])
@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_mg_solve_parity(cmp, geometry, size):
    """mgfor solve on a compatible random right-hand side: cold start, warm
    start on a new b, then the same b again (no V-cycle needed)."""
    from nyles_amd import _lib as L
    print("case mg", geometry, size, flush=True)
    nx, ny, nz = size
    orc, dev = cmp.make_pair(nx, ny, nz, geometry)
    mg = orc.mg
    rng = np.random.default_rng(3)
    r = None
    for rep in range(3):
        if rep < 2:
            r = rng.standard_normal((nz, ny, nx))
            r -= r.mean()
        sub = np.zeros(orc.shape)
        orc.set_interior(sub, r)
        orc.fill(sub)
        mg.b[...] = 0.0
        mg.b[mg.idx] = sub
        nite = mg.solve()
        dev.upload(L.NYL_F_MGB, r)
        s = dev.mg_solve()
        assert s["nite"] == nite, (rep, s, nite)
        if rep == 2:
            assert nite == 0
        x_or = mg.x[tuple(slice(3, -3) for _ in range(3))]
        _assert_same("x rep%d" % rep, dev.download(L.NYL_F_MGX), x_or)
    dev.close()


RK_CASES = [("closed", (32, 16, 16)), ("perio_x", (32, 16, 16)), ("perio_xyz", (32, 16, 16)),
            ("closed", (64, 16, 16)), ("perio_xyz", (128, 32, 16)), ("perio_xy", (64, 64, 32))]


@pytest.mark.parametrize("geometry,size", RK_CASES,
                         ids=["%s-%dx%dx%d" % ((g,) + s) for g, s in RK_CASES])
def test_rk3_steps_parity(cmp, geometry, size):
    """Three RK3_SSP steps of the whole model, bitwise."""
    print("case rk3", geometry, size, flush=True)
    nx, ny, nz = size
    orc, dev = cmp.make_pair(nx, ny, nz, geometry, Lx=nx / 16, Ly=ny / 16, Lz=nz / 16)
    st = cmp.random_state(orc, seed=7, amp_u=3e-3)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        from nyles_amd import _lib as L
        umax2 = dev.reduce(L.NYL_RED_MAXU2)
        U_max = np.sqrt(umax2)
        dt_dev = 0.1 if U_max == 0.0 else min(0.5 / U_max, 0.1)
        assert dt_dev == dt
        orc.rk3_ssp(dt)
        s = dev.step_rk3(dt)
        assert s["nite"] == sum(orc.nite_log[-3:])
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in want:
        _assert_same(k, got[k], want[k])
    dev.close()


def test_rotating_vorticity(cmp):
    """Coriolis term added to vor_k (core/vorticity.py:33-34)."""
    orc, dev = cmp.make_pair(16, 16, 16, "perio_xy", rotating=True, coriolis=2.5)
    st = cmp.random_state(orc, seed=2)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    _assert_same("vor_k", got["vor_k"], want["vor_k"])
    dev.close()


@pytest.mark.parametrize("scheme,geometry", [("EF", "perio_xy"), ("LFAM3", "closed"),
                                             ("LFAM3", "perio_xyz")])
def test_ef_lfam3_parity(cmp, scheme, geometry):
    """Euler forward and the LFAM3 predictor-corrector, three steps, bitwise."""
    print("case", scheme, geometry, flush=True)
    orc, dev = cmp.make_pair(32, 16, 16, geometry, Lx=2.0, Ly=1.0, Lz=1.0)
    st = cmp.random_state(orc, seed=9, amp_u=3e-3)
    cmp.load_state(orc, dev, st)
    orc.diagnose_var()
    dev.diagnose()
    for n in range(3):
        dt = orc.compute_dt(cfl=0.5, dt_max=0.1)
        if scheme == "EF":
            orc.euler_forward(dt)
            dev.step_ef(dt)
        else:
            orc.lfam3(dt)
            dev.step_lfam3(dt)
    want = cmp.oracle_fields(orc)
    got = cmp.device_fields(dev)
    for k in want:
        _assert_same(k, got[k], want[k])
    dev.close()


SWEEP_SIZES = [(64, 8, 8), (128, 64, 32), (64, 16, 8), (256, 32, 16)]


@pytest.mark.parametrize("size", SWEEP_SIZES, ids=["%dx%dx%d" % s for s in SWEEP_SIZES])
@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_tiled_sweep_matches_per_cell_sweep(cmp, geometry, size):
    """The 2.5-D tiled fsmoother3d sweep equals the per-cell sweep bitwise, and
    its fused residual norm equals the standalone one to rounding."""
    from nyles_amd import _lib as L
    print("case sweep", geometry, size, flush=True)
    nx, ny, nz = size
    orc, dev = cmp.make_pair(nx, ny, nz, geometry)
    rng = np.random.default_rng(12)
    dev.upload(L.NYL_F_MGX, rng.standard_normal((nz, ny, nx)))
    dev.upload(L.NYL_F_MGB, rng.standard_normal((nz, ny, nx)))
    dev.debug_mg(0)
    want = dev.download(L.NYL_F_MGY)
    dev.upload(L.NYL_F_MGY, np.zeros((nz, ny, nx)))
    dev.debug_mg(1)
    _assert_same("tiled sweep", dev.download(L.NYL_F_MGY), want)
    dev.upload(L.NYL_F_MGY, np.zeros((nz, ny, nx)))
    r2 = dev.debug_mg(2)
    _assert_same("tiled sweep + norm", dev.download(L.NYL_F_MGY), want)
    r3 = dev.debug_mg(3)
    assert abs(r2 - r3) <= 1e-12 * r3
    dev.close()


@pytest.mark.parametrize("size", SWEEP_SIZES, ids=["%dx%dx%d" % s for s in SWEEP_SIZES])
@pytest.mark.parametrize("geometry", GEOMETRIES)
def test_two_sweep_pass_matches_two_sweeps(cmp, geometry, size):
    """The two-sweep pass (k_mg_sweep2t: plain, with the residual norm of its
    input, and from a zero input) equals two per-cell sweeps bitwise; its
    residual norm equals the standalone one to rounding."""
    from nyles_amd import _lib as L
    print("case sweep2", geometry, size, flush=True)
    nx, ny, nz = size
    orc, dev = cmp.make_pair(nx, ny, nz, geometry)
    rng = np.random.default_rng(13)
    x0 = rng.standard_normal((nz, ny, nx))
    dev.upload(L.NYL_F_MGB, rng.standard_normal((nz, ny, nx)))

    def two_sweeps(x):
        dev.upload(L.NYL_F_MGX, x)
        dev.debug_mg(0)
        dev.upload(L.NYL_F_MGX, dev.download(L.NYL_F_MGY))
        dev.debug_mg(0)
        return dev.download(L.NYL_F_MGY)

    want = two_sweeps(x0)
    want0 = two_sweeps(np.zeros_like(x0))
    dev.upload(L.NYL_F_MGX, x0)
    dev.upload(L.NYL_F_MGY, np.zeros_like(x0))
    dev.debug_mg(4)
    _assert_same("two-sweep pass", dev.download(L.NYL_F_MGY), want)
    dev.upload(L.NYL_F_MGY, np.zeros_like(x0))
    r5 = dev.debug_mg(5)
    _assert_same("two-sweep pass + norm", dev.download(L.NYL_F_MGY), want)
    r3 = dev.debug_mg(3)
    assert abs(r5 - r3) <= 1e-12 * r3
    dev.upload(L.NYL_F_MGX, np.full_like(x0, np.nan))   # not read
    dev.upload(L.NYL_F_MGY, np.zeros_like(x0))
    dev.debug_mg(6)
    _assert_same("two-sweep pass from zero", dev.download(L.NYL_F_MGY), want0)
    dev.close()
