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
