"""The oracle's stencil operators, pinned by closed forms and invariants (CPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a shifted index, swapped operands) breaks it.
"""
import numpy as np
import pytest


def model(oracle_mod, geometry="closed", n=(16, 12, 8), L=(2.0, 1.5, 1.0)):
    nx, ny, nz = n
    return oracle_mod.OracleLES(nx, ny, nz, L[0], L[1], L[2], geometry)


def test_vorticity_of_solid_body_rotation(oracle_mod):
    """u = -Om*y, v = Om*x (covariant: times dx, dy): vor_k = 2*Om*dx*dy
    (the check of core/vorticity.py:104-106) on every cell where the curl is
    evaluated."""
    m = model(oracle_mod, "closed", n=(16, 16, 8), L=(1.0, 1.0, 1.0))
    Om = 0.7
    nz, ny, nx = m.shape
    yb = (np.arange(ny) + 0.5) * m.dy          # u lives at (x_{i+1/2}, y_j)
    xb = (np.arange(nx) + 0.5) * m.dx
    m.u["i"][:] = (-Om * yb)[None, :, None] * m.dx
    m.u["j"][:] = (Om * xb)[None, None, :] * m.dy
    m.vorticity()
    w = m.vor["k"][:, :-1, :-1]
    np.testing.assert_allclose(w, 2 * Om * m.dx * m.dy, rtol=1e-12)
    assert np.all(m.vor["k"][:, :, -1] == 0.0)


def test_curl_of_gradient_is_zero(oracle_mod):
    """u_d = delta_d(phi) with integer phi gives vor = 0 exactly."""
    m = model(oracle_mod, "closed")
    rng = np.random.default_rng(0)
    phi = rng.integers(-50, 50, size=(m.nz + 1, m.ny + 1, m.nx + 1)).astype(float)
    m.u["i"][:] = phi[:-1, :-1, 1:] - phi[:-1, :-1, :-1]
    m.u["j"][:] = phi[:-1, 1:, :-1] - phi[:-1, :-1, :-1]
    m.u["k"][:] = phi[1:, :-1, :-1] - phi[:-1, :-1, :-1]
    m.vorticity()
    for d in "ijk":
        assert np.all(m.vor[d] == 0.0)


def test_vorticity_components_orientation(oracle_mod):
    """Linear velocities with known cross derivatives give each component's
    sign and axes: u = a*z, v = b*x, w = c*y  ->  vor_i = c - 0, vor_j = a, vor_k = b."""
    m = model(oracle_mod, "closed", n=(10, 10, 10), L=(1.0, 1.0, 1.0))
    a, b, c = 2.0, -3.0, 5.0
    z = np.arange(m.nz, dtype=float)[:, None, None]
    y = np.arange(m.ny, dtype=float)[None, :, None]
    x = np.arange(m.nx, dtype=float)[None, None, :]
    m.u["i"][:] = a * z + 0 * x + 0 * y
    m.u["j"][:] = b * x + 0 * y + 0 * z
    m.u["k"][:] = c * y + 0 * x + 0 * z
    m.vorticity()
    assert np.all(m.vor["i"][:-1, :-1, :] == c)     # d_y w - d_z v
    assert np.all(m.vor["j"][:-1, :, :-1] == a)     # d_z u - d_x w
    assert np.all(m.vor["k"][:, :-1, :-1] == b)     # d_x v - d_y u


def test_kinetic_energy_of_constant_velocity(oracle_mod):
    """ke = 1/2 (u0^2/dx^2 + v0^2/dy^2 + w0^2/dz^2) (core/kinenergy.py:76-77)
    except the first cell along each wall axis, which misses that axis'
    contribution (the loop starts at i = 2)."""
    m = model(oracle_mod, "closed")
    u0, v0, w0 = 0.8, -4.0, 0.6
    m.u["i"][:] = u0
    m.u["j"][:] = v0
    m.u["k"][:] = w0
    m.kinenergy()
    cx, cy, cz = (0.5 * u0 ** 2 / m.dx ** 2, 0.5 * v0 ** 2 / m.dy ** 2,
                  0.5 * w0 ** 2 / m.dz ** 2)
    ke = m.ke
    np.testing.assert_allclose(ke[1:, 1:, 1:], cx + cy + cz, rtol=1e-14)
    np.testing.assert_allclose(ke[1:, 1:, 0], cy + cz, rtol=1e-14)
    np.testing.assert_allclose(ke[1:, 0, 1:], cx + cz, rtol=1e-14)
    np.testing.assert_allclose(ke[0, 1:, 1:], cx + cy, rtol=1e-14)
    assert ke[0, 0, 0] == 0.0


def test_div_of_gradient_is_the_7_point_laplacian(oracle_mod):
    """U = delta(p) (forward differences, zero normal flux at walls) then
    div (backward differences) equals the Neumann 7-point Laplacian."""
    m = model(oracle_mod, "closed", L=(16.0, 12.0, 8.0))   # dx = dy = dz = 1
    rng = np.random.default_rng(3)
    p = rng.standard_normal(m.shape)
    for d in "ijk":
        m.U[d][:] = 0.0
    m.U["i"][:, :, :-1] = p[:, :, 1:] - p[:, :, :-1]
    m.U["j"][:, :-1, :] = p[:, 1:, :] - p[:, :-1, :]
    m.U["k"][:-1, :, :] = p[1:, :, :] - p[:-1, :, :]
    m.compute_div()
    pp = np.pad(p, 1, mode="edge")
    lap = (pp[1:-1, 1:-1, 2:] + pp[1:-1, 1:-1, :-2] + pp[1:-1, 2:, 1:-1] + pp[1:-1, :-2, 1:-1]
           + pp[2:, 1:-1, 1:-1] + pp[:-2, 1:-1, 1:-1] - 6 * p)
    np.testing.assert_allclose(m.div, lap, atol=1e-12)


def test_gradient_correction_is_forward_difference(oracle_mod):
    """u_d -= delta_d p on every face but the last wall face
    (core/fortran_bernoulli.f90:18-24)."""
    m = model(oracle_mod, "closed")
    rng = np.random.default_rng(4)
    p = rng.standard_normal(m.shape)
    before = rng.standard_normal(m.shape)
    u = before.copy()
    oracle_mod.gradke(p, u)
    np.testing.assert_array_equal(u[:, :, :-1], before[:, :, :-1] - (p[:, :, 1:] - p[:, :, :-1]))
    np.testing.assert_array_equal(u[:, :, -1], before[:, :, -1])


def test_bernoulli_buoyancy_term(oracle_mod):
    """With ke = 0 and b linear in z, du_k = dz/2 (b_k + b_{k+1}) and the
    horizontal components are untouched (core/fortran_bernoulli.f90:29-58)."""
    m = model(oracle_mod, "closed")
    z = np.arange(m.nz, dtype=float)[:, None, None] * np.ones(m.shape)
    m.b[:] = 0.3 * z
    m.ke[:] = 0.0
    ds = m.ds[0]
    for d in "ijk":
        ds["u"][d][:] = 0.0
    from oracle import view
    for d in "ij":
        oracle_mod.gradke(view(m.ke, d), view(ds["u"][d], d))
    oracle_mod.gradkeandb(view(m.ke, "k"), view(m.b, "k"), view(ds["u"]["k"], "k"), m.dz)
    expect = 0.5 * m.dz * (m.b[1:] + m.b[:-1])
    np.testing.assert_allclose(ds["u"]["k"][:-1], expect, rtol=1e-15)
    assert np.all(ds["u"]["k"][-1] == 0.0)
    assert np.all(ds["u"]["i"] == 0.0) and np.all(ds["u"]["j"] == 0.0)


def test_vortex_force_vanishes_without_vorticity(oracle_mod):
    m = model(oracle_mod, "perio_xy")
    rng = np.random.default_rng(5)
    for d in "ijk":
        m.U[d][:] = rng.standard_normal(m.shape)
        m.vor[d][:] = 0.0
    m.ke[:] = 0.0
    m.b[:] = 0.0
    ds = m.ds[0]
    m.rhs(ds)
    for d in "ijk":
        assert np.all(ds["u"][d] == 0.0)


def test_vortex_force_uniform_flow_single_component(oracle_mod):
    """Uniform U^x = U0 > 0 with vor_k only: du_j = -U0 * WENO5(vor_k) upwinded
    along x, and du_i = 0 since U^y = 0 (core/fortran_vortex_force.f90:66-81).
    On a periodic x line every interior point takes the interior stencil."""
    m = model(oracle_mod, "perio_xyz", n=(16, 8, 8), L=(1.0, 1.0, 1.0))
    rng = np.random.default_rng(6)
    U0 = 0.37
    for d in "ijk":
        m.U[d][:] = 0.0
        m.vor[d][:] = 0.0
    m.U["i"][:] = U0
    w = rng.standard_normal(m.shape)
    m.vor["k"][:] = w
    m.ke[:] = 0.0
    m.b[:] = 0.0
    ds = m.ds[0]
    m.rhs(ds)
    assert np.all(ds["u"]["i"] == 0.0)
    nh = m.nh
    k, j = nh + 2, nh + 3
    for i in range(nh, nh + m.nx):
        q = w[k, j, i - 3:i + 2]                     # q(i) = vor(i-1): stencil i-3 .. i+1
        expect = -(U0 * oracle_mod.weno5(*q))
        assert ds["u"]["j"][k, j, i] == expect


def test_tracer_flux_form_conserves_on_closed_domain(oracle_mod):
    """Wall axes: flux(n) = 0 and no inflow, so sum(db) telescopes to zero
    (up to rounding) for any b and U."""
    m = model(oracle_mod, "closed")
    rng = np.random.default_rng(7)
    for d in "ijk":
        m.U[d][:] = rng.standard_normal(m.shape)
        m.vor[d][:] = 0.0
    m.b[:] = rng.standard_normal(m.shape)
    ds = m.ds[0]
    m.rhs(ds)
    assert abs(ds["b"].sum()) < 1e-12 * np.abs(ds["b"]).sum()


def test_uniform_tracer_in_divergence_free_flow(oracle_mod):
    """A uniform b is not changed by a discretely divergence-free periodic U,
    up to the float32 consistency factor of the WENO5 stencils."""
    m = model(oracle_mod, "perio_xyz", n=(16, 16, 16), L=(1.0, 1.0, 1.0))
    rng = np.random.default_rng(8)
    psi = rng.standard_normal((m.nz, m.ny, m.nx))
    # psi at the (x, y) edges: U = delta_y psi, V = -delta_x psi (backward
    # differences onto the faces) makes the backward-difference div zero
    Ux = psi - np.roll(psi, 1, axis=1)
    Uy = -(psi - np.roll(psi, 1, axis=2))
    m.set_interior(m.U["i"], Ux)
    m.set_interior(m.U["j"], Uy)
    m.set_interior(m.U["k"], 0.0 * psi)
    for d in "ijk":
        m.fill(m.U[d])
        m.vor[d][:] = 0.0
    m.b[:] = 1.0
    ds = m.ds[0]
    m.rhs(ds)
    assert np.max(np.abs(m.interior(ds["b"]))) < 1e-6 * np.max(np.abs(Ux))


def _flux_line(oracle_mod, UU1, q_shifted):
    """flux1d of a line whose face velocity UU1 is constant along the line:
    the velocity the vortex force upwinds with is the staggered average
    0.5*(UU_{i-1} + UU_i) with UU_0 = 0, i.e. UU1/2 at the first point and
    UU1 elsewhere."""
    n = len(q_shifted)
    u1d = np.full(n, 0.5 * (UU1 + UU1))
    u1d[0] = 0.5 * (0.0 + UU1)
    return oracle_mod.flux1d(u1d, q_shifted)


def _only(m, U=None, vor=None):
    for d in "ijk":
        m.U[d][:] = 0.0
        m.vor[d][:] = 0.0
    for d, a in (U or {}).items():
        m.U[d][:] = a
    for d, a in (vor or {}).items():
        m.vor[d][:] = a
    m.ke[:] = 0.0
    m.b[:] = 0.0


def test_vortex_force_flip_term_with_sheared_flow(oracle_mod):
    """core/fortran_vortex_force.f90:145-159 through the 'kji' triplet of
    core/vortex_force.py:69-81: with U^x = f(z) only and vor_j (omega_y) only,
    du_k += flux1d along x of U^x averaged in z onto the omega point,
    upwinding q(i) = omega_y(i-1), q(1) = 0; the last z face is untouched and
    du_i = du_j = 0.  A wrong averaging axis (x instead of z) or an unshifted
    q(i) = omega(i) changes du_k on most cells."""
    m = model(oracle_mod, "closed", n=(12, 10, 14), L=(1.0, 1.0, 1.0))
    rng = np.random.default_rng(61)
    nz, ny, nx = m.shape
    f = rng.uniform(-1.0, 1.0, nz)
    f[3] = 0.0                                   # exact zero: the u > 0 split
    om = rng.standard_normal(m.shape)
    _only(m, U={"i": f[:, None, None] * np.ones(m.shape)}, vor={"j": om})
    ds = m.ds[0]
    m.rhs(ds)
    assert np.all(ds["u"]["i"] == 0.0) and np.all(ds["u"]["j"] == 0.0)
    got = ds["u"]["k"]
    for k in range(nz - 1):
        UU1 = 0.5 * (f[k] + f[k + 1])
        for j in range(ny):
            q = np.concatenate(([0.0], om[k, j, :-1]))
            np.testing.assert_array_equal(got[k, j, :], 0.0 + _flux_line(oracle_mod, UU1, q))
    assert np.all(got[nz - 1] == 0.0)
    # the wrong axis or the unshifted q would differ
    k, j = 2, 4
    wrong_axis = _flux_line(oracle_mod, f[k], np.concatenate(([0.0], om[k, j, :-1])))
    unshifted = _flux_line(oracle_mod, 0.5 * (f[k] + f[k + 1]), om[k, j, :])
    assert np.count_nonzero(wrong_axis != got[k, j]) > nx // 2
    assert np.count_nonzero(unshifted != got[k, j]) > nx // 2


def test_vortex_force_flip_term_uniform_flow_periodic(oracle_mod):
    """'ikj' flip term on a periodic y line: uniform U^y = U0 > 0 with
    omega_z only gives du_x = +U0 * WENO5(omega_z(j-3 .. j+1)) at every
    interior cell, and du_y = du_z = 0."""
    m = model(oracle_mod, "perio_xyz", n=(8, 16, 8), L=(1.0, 1.0, 1.0))
    rng = np.random.default_rng(62)
    U0 = 0.61
    w = rng.standard_normal(m.shape)
    _only(m, U={"j": U0}, vor={"k": w})
    ds = m.ds[0]
    m.rhs(ds)
    assert np.all(ds["u"]["j"] == 0.0) and np.all(ds["u"]["k"] == 0.0)
    nh = m.nh
    k, i = nh + 1, nh + 5
    for j in range(nh, nh + m.ny):
        expect = 0.0 + U0 * oracle_mod.weno5(*w[k, j - 3:j + 2, i])
        assert ds["u"]["i"][k, j, i] == expect


def test_vortex_force_direc_term_with_sheared_flow(oracle_mod):
    """core/fortran_vortex_force.f90:66-81 through the 'ikj' triplet: with
    U^x = f(y) only and omega_z only, du_j -= flux1d along x of U^x averaged
    in y onto the omega point; the last y face is untouched; du_i = du_k = 0."""
    m = model(oracle_mod, "closed", n=(14, 10, 8), L=(1.0, 1.0, 1.0))
    rng = np.random.default_rng(63)
    nz, ny, nx = m.shape
    f = rng.uniform(-1.0, 1.0, ny)
    om = rng.standard_normal(m.shape)
    _only(m, U={"i": f[None, :, None] * np.ones(m.shape)}, vor={"k": om})
    ds = m.ds[0]
    m.rhs(ds)
    assert np.all(ds["u"]["i"] == 0.0) and np.all(ds["u"]["k"] == 0.0)
    got = ds["u"]["j"]
    for j in range(ny - 1):
        UU1 = 0.5 * (f[j] + f[j + 1])
        for k in range(nz):
            q = np.concatenate(([0.0], om[k, j, :-1]))
            np.testing.assert_array_equal(got[k, j, :], 0.0 - _flux_line(oracle_mod, UU1, q))
    assert np.all(got[:, ny - 1, :] == 0.0)


@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_y", "perio_xy", "perio_xyz"])
def test_halo_fill_is_periodic_image(oracle_mod, geometry):
    """core/mpi/halo.py:123-188 on one rank: after a fill every halo cell
    equals its periodic image (index wrapped into the interior), the
    interior is unchanged, and wall axes carry no halo
    (core/mpi/topology.py:253-304)."""
    nh = 3
    m = model(oracle_mod, geometry, n=(10, 9, 8))
    perio = oracle_mod.GEOMETRY_PERIODIC[geometry]
    for a, (n, p) in enumerate(zip((m.nz, m.ny, m.nx), perio)):
        assert m.shape[a] == (n + 2 * nh if p else n)
    rng = np.random.default_rng(64)
    x = rng.standard_normal(m.shape)
    before = x.copy()
    oracle_mod.halo_fill(x, geometry, nh)
    idx = []
    inner = []
    for a, (n, p) in enumerate(zip((m.nz, m.ny, m.nx), perio)):
        if p:
            h = np.arange(n + 2 * nh)
            idx.append(nh + (h - nh) % n)
            inner.append(slice(nh, nh + n))
        else:
            idx.append(np.arange(n))
            inner.append(slice(0, n))
    image = before[np.ix_(*idx)]
    np.testing.assert_array_equal(x, image)
    np.testing.assert_array_equal(x[tuple(inner)], before[tuple(inner)])
