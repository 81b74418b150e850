"""The oracle's physics hooks, pinned by closed forms (CPU): viscosity
(add_laplacian, core/fortran_dissipation.f90:2-34, applied by
core/viscosity.py:3-9 at the last stage, core/model_les.py:140-141),
passive tracers (core/model_les.py:36-43, advected by core/tracer.py:59-72)
and the user forcing hook (core/model_les.py:143-144).

Each check is chosen so that a plausible slip (a wrong sign, a dropped
end-of-line closure, a shifted neighbour, the wrong stage) breaks it.
"""
import numpy as np
import pytest


def test_add_laplacian_of_a_quadratic_on_a_wall_line(oracle_mod):
    """phi(i) = i^2 on a line of n cells: interior cells gain 2*coef, the
    first cell coef*(phi1 - phi0) = coef (no inflow flux), the last
    -coef*(phi(n-1) - phi(n-2)) = -coef*(2n-3) (no outflow flux)."""
    O = oracle_mod
    n, coef = 9, 0.5
    phi = np.zeros((2, 3, n))
    phi[...] = (np.arange(n, dtype=float) ** 2)[None, None, :]
    d = np.full_like(phi, 1.0)
    O.add_laplacian(phi, d, coef)
    want = np.full(n, 1.0 + 2 * coef)
    want[0] = 1.0 + coef
    want[-1] = 1.0 - coef * (2 * n - 3)
    for k in range(2):
        for j in range(3):
            assert np.array_equal(d[k, j], want)


def test_add_laplacian_conserves_on_a_wall_line(oracle_mod):
    """Integer data and coef = 1: the added values telescope to exactly 0
    along each line (fluxes cancel between neighbours, none at the ends)."""
    O = oracle_mod
    rng = np.random.default_rng(4)
    phi = rng.integers(-20, 20, size=(3, 4, 11)).astype(float)
    d = np.zeros_like(phi)
    O.add_laplacian(phi, d, 1.0)
    assert np.all(d.sum(axis=2) == 0.0)
    # and the interior is the three-point second difference
    sec = phi[:, :, 2:] - 2 * phi[:, :, 1:-1] + phi[:, :, :-2]
    assert np.array_equal(d[:, :, 1:-1], sec)


def test_add_laplacian_strided_view_follows_the_line_axis(oracle_mod):
    """On a transposed view the line runs along the view's last axis: a
    field varying only along array axis 0 is untouched by the x pass and
    gets the second difference from the z pass."""
    O = oracle_mod
    m = O.OracleLES(8, 8, 10, geometry="closed")
    z = np.arange(m.nz, dtype=float)[:, None, None] ** 2 * np.ones(m.shape)
    d = np.zeros(m.shape)
    O.add_laplacian(O.view(z, "i"), O.view(d, "i"), 1.0)
    assert np.all(d == 0.0)
    O.add_laplacian(O.view(z, "k"), O.view(d, "k"), 1.0)
    assert np.all(d[1:-1] == 2.0)
    assert np.all(d[0] == 1.0) and np.all(d[-1] == -(2 * m.nz - 3))


def _shear_state(m, amp=1e-3, seed=3):
    rng = np.random.default_rng(seed)
    I = m.interior
    for d in "ijk":
        I(m.u[d])[...] = amp * rng.standard_normal(I(m.u[d]).shape)
    if not m.perio[0]:
        I(m.u["k"])[-1] = 0.0
    if not m.perio[1]:
        I(m.u["j"])[:, -1] = 0.0
    if not m.perio[2]:
        I(m.u["i"])[:, :, -1] = 0.0


@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_xyz"])
def test_viscosity_only_at_the_last_stage(oracle_mod, geometry):
    """rhs(last=False) ignores the viscosity; rhs(last=True) adds exactly
    sum_d nu*ids2[d]*delta_d^2(u) (checked against numpy on the periodic
    image of the state)."""
    O = oracle_mod
    nu = 0.01
    m = O.OracleLES(16, 12, 10, 2.0, 1.5, 1.0, geometry, viscosity=nu)
    ref = O.OracleLES(16, 12, 10, 2.0, 1.5, 1.0, geometry)
    for mm in (m, ref):
        _shear_state(mm)
        mm.diagnose_var()
    d0, d1 = m._new_tendency(), ref._new_tendency()
    m.rhs(d0, last=False)
    ref.rhs(d1, last=False)
    for c in "ijk":
        assert np.array_equal(d0["u"][c], d1["u"][c])
    m.rhs(d0, last=True)
    ref.rhs(d1, last=True)
    I = m.interior
    pz, py, px = m.perio
    for c in "ijk":
        u = I(m.u[c])
        lap = np.zeros_like(u)
        for ax, per, ids2 in ((2, px, m.idx2), (1, py, m.idy2), (0, pz, m.idz2)):
            if per:
                sec = np.roll(u, -1, ax) - 2 * u + np.roll(u, 1, ax)
            else:
                fx = np.diff(u, axis=ax)
                zero = np.zeros_like(np.take(u, [0], axis=ax))
                right = np.concatenate([fx, zero], axis=ax)
                left = np.concatenate([zero, fx], axis=ax)
                sec = right - left
            lap += nu * ids2 * sec
        got = I(d0["u"][c]) - I(d1["u"][c])
        np.testing.assert_allclose(got, lap, rtol=1e-9, atol=1e-18)


@pytest.mark.parametrize("geometry", ["closed", "perio_xy"])
def test_tracer_equals_b_advected_without_buoyancy(oracle_mod, geometry):
    """A passive tracer holding the same values as b gets exactly b's
    advective tendency; b then acts on the momentum (Bernoulli's buoyancy
    term) and the tracer does not."""
    O = oracle_mod
    m = O.OracleLES(16, 12, 10, 2.0, 1.5, 1.0, geometry, n_tracers=2)
    _shear_state(m, amp=1e-2)
    rng = np.random.default_rng(8)
    bi = rng.standard_normal(m.interior(m.b).shape)
    m.interior(m.b)[...] = bi
    m.interior(m.tracers["t0"])[...] = bi
    m.interior(m.tracers["t1"])[...] = 2.0 * bi
    m.diagnose_var()
    ds = m._new_tendency()
    m.rhs(ds)
    assert np.array_equal(m.interior(ds["t0"]), m.interior(ds["b"]))
    # upwind is linear in the tracer: 2b has twice the tendency (exactly,
    # the WENO weights being scale-invariant up to the epsilons)
    np.testing.assert_allclose(m.interior(ds["t1"]), 2 * m.interior(ds["b"]),
                               rtol=1e-6, atol=1e-12)
    # a tracer's value does not reach the momentum
    m2 = O.OracleLES(16, 12, 10, 2.0, 1.5, 1.0, geometry, n_tracers=2)
    _shear_state(m2, amp=1e-2)
    m2.interior(m2.b)[...] = bi
    m2.interior(m2.tracers["t0"])[...] = 0.0
    m2.diagnose_var()
    ds2 = m2._new_tendency()
    m2.rhs(ds2)
    for c in "ijk":
        assert np.array_equal(ds2["u"][c], ds["u"][c])


def test_forcing_hook_sees_each_stage_time(oracle_mod):
    """The forcing callback runs once per rhs with the stage times of RK3
    (t, t+dt, t+dt/2) and what it adds reaches the state: a constant
    forcing F on a passive tracer in a fluid at rest gives t0 = dt*F after
    one step (RK3 integrates a constant tendency exactly), while b and u
    stay 0."""
    O = oracle_mod
    m = O.OracleLES(16, 12, 10, 2.0, 1.5, 1.0, "closed", n_tracers=1)
    seen = []
    F = 0.25

    def forcing(model, ds, t):
        seen.append(t)
        ds["t0"][...] += F

    m.forcing = forcing
    m.diagnose_var()
    dt = 0.1
    m.rk3_ssp(dt, t=1.0)
    assert seen == [1.0, 1.1, 1.05]
    np.testing.assert_allclose(m.interior(m.tracers["t0"]), dt * F, rtol=1e-14)
    assert np.all(m.b == 0.0) and all(np.all(m.u[d] == 0.0) for d in "ijk")
    # Euler forward and LFAM3 pass their own stage times
    seen.clear()
    m.euler_forward(dt, t=2.0)
    m.lfam3(dt, t=3.0)
    m.lfam3(dt, t=3.1)
    assert seen == [2.0, 3.0, 3.1, 3.1 + dt * .5]


def _random_model(O, model, geometry, seed):
    m = O.OracleLES(16, 12, 10, 2.0, 1.5, 1.0, geometry, model=model, n_tracers=1)
    _shear_state(m, amp=1e-2, seed=seed)
    rng = np.random.default_rng(seed + 1)
    m.interior(m.b)[...] = rng.standard_normal(m.interior(m.b).shape)
    m.interior(m.tracers["t0"])[...] = rng.standard_normal(m.interior(m.b).shape)
    return m


@pytest.mark.parametrize("geometry", ["closed", "perio_xy", "perio_xyz"])
def test_linear_model_has_no_vortex_force(oracle_mod, geometry):
    """LES(linear=True) (core/model_les.py:33-34,122-127,135-136): the
    diagnosis leaves vorticity and ke as they were (zero here), so the
    momentum tendency is the buoyancy term alone: du_x = du_y = 0 and
    du_z = 0.5*dz*(b(k) + b(k+1)) on the faces inside the domain; b and the
    tracer get the LES advective tendency.  The vorticity never reaches the
    tendencies: planting one changes nothing."""
    O = oracle_mod
    lin = _random_model(O, "linear", geometry, 11)
    les = _random_model(O, "LES", geometry, 11)
    lin.diagnose_var()
    les.diagnose_var()
    for d in "ijk":
        assert np.all(lin.vor[d] == 0.0)
        assert np.array_equal(lin.u[d], les.u[d])
    assert np.all(lin.ke == 0.0) and np.any(les.ke != 0.0)
    ds, dl = lin._new_tendency(), les._new_tendency()
    lin.rhs(ds)
    les.rhs(dl)
    assert np.array_equal(ds["b"], dl["b"]) and np.array_equal(ds["t0"], dl["t0"])
    I = lin.interior
    assert np.all(I(ds["u"]["i"]) == 0.0) and np.all(I(ds["u"]["j"]) == 0.0)
    b = I(lin.b)
    want = np.zeros_like(b)
    if lin.perio[0]:
        want[...] = 0.5 * lin.dz * (b + np.roll(b, -1, 0))
    else:
        want[:-1] = 0.5 * lin.dz * (b[:-1] + b[1:])
    np.testing.assert_allclose(I(ds["u"]["k"]), want, rtol=1e-15, atol=0)
    # a planted vorticity does not change the linear tendencies
    for d in "ijk":
        lin.vor[d][...] = np.random.default_rng(5).standard_normal(lin.shape)
    ds2 = lin._new_tendency()
    lin.rhs(ds2)
    for c in "ijk":
        assert np.array_equal(ds2["u"][c], ds["u"][c])


@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_xyz"])
def test_euler3d_advects_nothing_and_has_no_buoyancy(oracle_mod, geometry):
    """core/model_les_euler.py: b and the tracers get zero tendencies (the
    rhstrac call is commented out, :131) and the momentum tendency does not
    depend on b (bernoulli_euler, core/bernoulli.py:26-32); with b = 0 it
    is the LES tendency bit for bit.  Over an RK3 step b and the tracer
    keep their values."""
    O = oracle_mod
    eu = _random_model(O, "Euler3d", geometry, 21)
    eu.diagnose_var()
    ds = eu._new_tendency()
    eu.rhs(ds)
    assert np.all(ds["b"] == 0.0) and np.all(ds["t0"] == 0.0)
    # b does not reach the momentum
    eu2 = _random_model(O, "Euler3d", geometry, 21)
    eu2.interior(eu2.b)[...] = 0.0
    eu2.diagnose_var()
    ds2 = eu2._new_tendency()
    eu2.rhs(ds2)
    les = _random_model(O, "LES", geometry, 21)
    les.interior(les.b)[...] = 0.0
    les.diagnose_var()
    dl = les._new_tendency()
    les.rhs(dl)
    for c in "ijk":
        assert np.array_equal(ds2["u"][c], ds["u"][c])
        assert np.array_equal(dl["u"][c], ds["u"][c])
    b0, t0 = eu.b.copy(), eu.tracers["t0"].copy()
    eu.rk3_ssp(0.01)
    assert np.array_equal(eu.b, b0) and np.array_equal(eu.tracers["t0"], t0)
