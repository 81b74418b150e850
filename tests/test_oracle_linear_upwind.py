"""The reference's linear upwind branches (never run by the shipped
reference: linear=.false. at core/fortran_upwind.f90:30 and
core/fortran_vortex_force.f90:28,104) as the oracle restates them, pinned
without retyping their formulas (CPU):

* the interpolations (core/interpolate_tracer.f90, core/interpolate.f90)
  reproduce the face values of finite-volume data of a polynomial of the
  scheme's degree (5th order: degree 4; 3rd: degree 2), up to the REAL(4)
  rounding of their literals, with the documented closures at the line ends;
* on data linear along the line, WENO5-Z's weights are the ideal ones, so the
  WENO flux (the oracle's separately pinned flux1d) and the linear-5 flux
  agree: tracer and vortex-force tendencies match the WENO branch's away
  from wall closures;
* the upwind split: a negative velocity takes the right-biased value
  (qm(i+1)), a positive one the left-biased qp(i); on a periodic line the
  flux divergence telescopes to zero.
"""
import numpy as np
import pytest


def _fv(p, idx):
    """averages of x**p over [i - 1/2, i + 1/2]"""
    a, b = idx - 0.5, idx + 0.5
    return (b ** (p + 1) - a ** (p + 1)) / (p + 1)


@pytest.mark.parametrize("order,deg", [(5, 4), (3, 2), (1, 0)])
def test_tracer_interpolation_is_exact_to_its_degree(oracle_mod, order, deg):
    O = oracle_mod
    n = 16
    i = np.arange(n, dtype=float)
    for p in range(deg + 1):
        q = _fv(p, i - 7.0)               # centred so that the values stay moderate
        qp, qm = O.interp_tracer(q, order)
        face_r, face_l = (i - 7.0 + 0.5) ** p, (i - 7.0 - 0.5) ** p
        lo = {5: 2, 3: 1, 1: 0}[order]
        hi = n - 1 - lo
        for a, w in ((qp, face_r), (qm, face_l)):
            np.testing.assert_allclose(a[lo:hi + 1], w[lo:hi + 1], rtol=2e-7, atol=2e-7 * 7 ** p)
        # the line ends are first order: copies of the cell value
        assert qp[0] == q[0] and qm[0] == q[0] and qp[-1] == q[-1] and qm[-1] == q[-1]
    if order == 5:
        # next to the ends: third order (exact on quadratics, not on cubics)
        q = _fv(2, i)
        qp, qm = O.interp_tracer(q, 5)
        for j in (1, n - 2):
            np.testing.assert_allclose(qp[j], (j + 0.5) ** 2, rtol=2e-7)
            np.testing.assert_allclose(qm[j], (j - 0.5) ** 2, rtol=2e-7)
        q = _fv(3, i)
        qp, _ = O.interp_tracer(q, 5)
        assert abs(qp[1] - 1.5 ** 3) > 1e-3


def test_vf_interpolation_closures_and_exactness(oracle_mod):
    """core/interpolate.f90 (order 5): qp(0) = 0, first order at 1 and n-1,
    third at 2 and n-2, fifth in 3..n-3 (1-based), qm(n) = v(n)."""
    O = oracle_mod
    n = 14
    k = np.arange(1, n + 1, dtype=float)
    v = _fv(4, k - 7.0)
    qp, qm = O.interp_vf(v, 5)            # qp(0..n-1), qm(1..n)
    assert qp[0] == 0.0 and qp[1] == v[0] and qm[0] == v[0]
    assert qp[n - 1] == v[n - 2] and qm[n - 2] == v[n - 2] and qm[n - 1] == v[n - 1]
    for j in range(3, n - 2):             # 1-based fifth-order range
        np.testing.assert_allclose(qp[j], (j - 7.0 + 0.5) ** 4, rtol=2e-7, atol=1e-4)
        np.testing.assert_allclose(qm[j - 1], (j - 7.0 - 0.5) ** 4, rtol=2e-7, atol=1e-4)
    v2 = _fv(2, k)
    qp, qm = O.interp_vf(v2, 5)
    for j in (2, n - 2):
        np.testing.assert_allclose(qp[j], (j + 0.5) ** 2, rtol=2e-7)
        np.testing.assert_allclose(qm[j - 1], (j - 0.5) ** 2, rtol=2e-7)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_upwind_split_takes_the_upwind_side(oracle_mod, sign):
    """Uniform velocity u on a line: the flux of face i is u*qp(i) for u > 0
    and u*qm(i+1) for u < 0; d(i) = (d + fx(i-1)) - fx(i), 0 fluxes through
    the ends."""
    O = oracle_mod
    n = 12
    rng = np.random.default_rng(2)
    q = rng.standard_normal((1, 1, n))
    u = np.full_like(q, 0.7 * sign)
    d = np.zeros_like(q)
    O.upwind_linear(q, u, d, 5)
    qp, qm = O.interp_tracer(q[0, 0], 5)
    fx = np.zeros(n + 1)                  # fx[i + 1]: face after cell i
    for i in range(n - 1):
        fx[i + 1] = 0.7 * sign * (qp[i] if sign > 0 else qm[i + 1])
    want = fx[:-1] - fx[1:]
    np.testing.assert_allclose(d[0, 0], want, rtol=1e-14, atol=1e-15)


def test_periodic_line_telescopes(oracle_mod):
    O = oracle_mod
    m = O.OracleLES(16, 8, 8, geometry="perio_xyz", upwind_a=5)
    rng = np.random.default_rng(7)
    for d in "ijk":
        m.interior(m.u[d])[...] = 1e-2 * rng.standard_normal(m.interior(m.u[d]).shape)
    m.interior(m.b)[...] = rng.standard_normal(m.interior(m.b).shape)
    m.diagnose_var()
    ds = m._new_tendency()
    m.rhs(ds)
    assert abs(m.interior(ds["b"]).sum()) < 1e-13 * np.abs(m.interior(ds["b"])).sum()


@pytest.mark.parametrize("geometry", ["closed", "perio_xyz"])
def test_linear5_matches_weno_on_linear_data(oracle_mod, geometry):
    """b and the vorticity linear along each line: WENO5-Z's weights are the
    ideal ones and both branches reconstruct the same face values (to the
    REAL(4) literals), so the tendencies agree away from wall closures."""
    O = oracle_mod
    n = 16
    ms = [O.OracleLES(n, n, n, geometry=geometry, upwind_a=a, upwind_vf=a) for a in (0, 5)]
    rng = np.random.default_rng(3)
    uu = {d: 1e-2 * rng.standard_normal((n, n, n)) for d in "ijk"}
    z, y, x = np.meshgrid(*(np.arange(n, dtype=float),) * 3, indexing="ij")
    for m in ms:
        for d in "ijk":
            m.interior(m.u[d])[...] = uu[d]
        m.U_from_u()
        m.interior(m.b)[...] = 0.3 * x - 0.2 * y + 0.5 * z + 1.0
        for c, (ax, ay, az) in zip("ijk", ((0.1, 0.2, -0.3), (-0.2, 0.1, 0.4), (0.3, -0.1, 0.2))):
            m.interior(m.vor[c])[...] = ax * x + ay * y + az * z
        if geometry != "closed":
            for f in [m.b] + [m.vor[c] for c in "ijk"] + [m.U[c] for c in "ijk"]:
                m.fill(f)
    out = []
    for m in ms:
        ds = m._new_tendency()
        m.rhs(ds)
        out.append(ds)
    if geometry == "closed":
        sl = (slice(3, -3),) * 3
    else:
        # linear data is not periodic: compare away from the wrap
        sl = (slice(3 + 3, -3 - 3),) * 3
    I = ms[0].interior
    for key in ("b", "u_i", "u_j", "u_k"):
        a = I(ms[0].get(key, out[0]))[sl]
        b = I(ms[1].get(key, out[1]))[sl]
        np.testing.assert_allclose(b, a, rtol=1e-6, atol=1e-7 * np.abs(a).max())


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_vf_split_takes_the_upwind_side(oracle_mod, sign):
    """vortex_force_direc, linear branch, uniform U along the averaging
    axis: U_interp(1) = U/2 (the line starts from UU_0 = 0), U elsewhere;
    res(k) -= qp(k-1)*U for U > 0 and qm(k)*U for U < 0; flip adds."""
    O = oracle_mod
    m_, n_, l_ = 2, 3, 10
    rng = np.random.default_rng(4)
    vort = rng.standard_normal((m_, n_, l_))
    U = np.full_like(vort, 0.6 * sign)
    res = np.zeros_like(vort)
    O.vortex_force_direc_linear(U, vort, res, 5)
    for j in range(m_):
        for i in range(n_ - 1):           # the last i is skipped (wall)
            qp, qm = O.interp_vf(vort[j, i], 5)
            Ui = np.full(l_, 0.6 * sign)
            Ui[0] = 0.3 * sign
            want = -(qp * Ui) if sign > 0 else -(qm * Ui)
            np.testing.assert_allclose(res[j, i], want, rtol=1e-14, atol=1e-16)
        assert np.all(res[j, n_ - 1] == 0.0)
    res2 = np.zeros((m_, l_, n_))
    vt = np.ascontiguousarray(np.transpose(vort, (0, 2, 1)))
    Ut = np.full_like(vt, 0.6 * sign)
    O.vortex_force_flip_linear(Ut, vt, res2, 5)
    np.testing.assert_allclose(np.transpose(res2, (0, 2, 1))[:, :-1], -res[:, :-1],
                               rtol=1e-14, atol=1e-16)
