"""The oracle's WENO reconstructions and flux1d closures, pinned by
exact-rational fixtures, closed forms and symmetries (CPU)."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

# REAL(4) literals of core/weno.f90 widened to double (f2py build without
# -fdefault-real-8); the values a float32 literal takes
LITERALS = {"1/3": 0.3333333432674408, "7/6": 1.1666666269302368, "11/6": 1.8333333730697632,
            "1/6": 0.1666666716337204, "5/6": 0.8333333134651184, "13/12": 1.0833333730697632,
            "eps5": 1.0000000168623835e-16, "eps3": 9.9999998245167e-15}


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(HERE, "golden", "weno_fixtures.json")) as f:
        return json.load(f)


def test_literal_table(fixtures):
    assert fixtures["constants"] == LITERALS


def test_weno5_matches_exact_rational(oracle_mod, fixtures):
    for case in fixtures["weno5"]:
        q = case["q"]
        got = oracle_mod.weno5(*q)
        scale = max(1e-300, max(abs(v) for v in q))
        assert abs(got - case["out"]) <= 2e-14 * scale, (q, got, case["out"])


def test_weno3_matches_exact_rational(oracle_mod, fixtures):
    for case in fixtures["weno3"]:
        q = case["q"]
        got = oracle_mod.weno3(*q)
        scale = max(1e-300, max(abs(v) for v in q))
        assert abs(got - case["out"]) <= 2e-14 * scale, (q, got, case["out"])


def test_weno5_linear_data_is_upwind5(oracle_mod):
    """On linear data beta1=beta2=beta3, tau5=0, the weights are (1,6,3)/10
    and WENO5 is the linear upwind-5 value (2,-13,47,27,-3)/60 . q, i.e. the
    midpoint value a + b/2 up to the float32 literals."""
    for a, b in [(0.0, 1.0), (3.0, -2.0), (-1.0, 0.25)]:
        q = [a + b * k for k in (-2, -1, 0, 1, 2)]
        got = oracle_mod.weno5(*q)
        lin = np.dot(np.array([2, -13, 47, 27, -3]) / 60.0, q)
        assert abs(got - (a + b / 2)) < 1e-6 * (abs(a) + abs(b))
        assert abs(got - lin) < 1e-6 * (abs(a) + abs(b))


def test_weno5_constant_consistency_factor(oracle_mod):
    """Constant data c gives c times (w1 s1 + 6 s2 + 3 s3)/10 where s_i are
    the float32 coefficient sums of the candidate stencils."""
    L = LITERALS
    s1 = L["1/3"] - L["7/6"] + L["11/6"]
    s2 = -L["1/6"] + L["5/6"] + L["1/3"]
    s3 = L["1/3"] + L["5/6"] - L["1/6"]
    for c in (1.0, 0.7, -3.5):
        # betas are zero, tau5 = 0, weights (1, 6, 3)
        expect = (1 * (c * s1) + 6 * (c * s2) + 3 * (c * s3)) / 10.0
        assert abs(oracle_mod.weno5(c, c, c, c, c) - expect) < 1e-15 * abs(c)
        assert abs(oracle_mod.weno5(c, c, c, c, c) - c) < 1e-7 * abs(c)


def test_weno5_fifth_order_on_smooth_averages(oracle_mod):
    """Cell averages of sin(x): the interface value converges at 5th order
    (WENO-Z with p = 1) until the float32 literals floor the error."""
    x0 = 0.3
    errs = []
    for h in (0.2, 0.1, 0.05):
        edges = x0 + h * np.arange(-2, 4) - h / 2      # 5 cells around x0
        avg = (np.cos(edges[:-1]) - np.cos(edges[1:])) / h
        got = oracle_mod.weno5(*avg)
        errs.append(abs(got - np.sin(x0 + h / 2)))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert r1 > 20 and r2 > 20, errs


def test_weno5_nonoscillatory_at_a_step(oracle_mod):
    for q in ([0, 0, 0, 1, 1], [0, 0, 1, 1, 1], [1, 1, 1, 0, 0]):
        v = oracle_mod.weno5(*map(float, q))
        assert -1e-6 <= v <= 1 + 1e-6


def test_flux1d_closures(oracle_mod):
    """flux(n) = 0; u > 0 at i = 1 takes q(1); u <= 0 at n-1 takes q(n);
    interior uses weno5 of the upwind stencil (core/weno.f90:116-150)."""
    rng = np.random.default_rng(1)
    n = 12
    q = rng.standard_normal(n)
    up = np.abs(rng.standard_normal(n)) + 0.1
    f = oracle_mod.flux1d(up, q)
    assert f[-1] == 0.0
    assert f[0] == up[0] * q[0]
    assert f[1] == up[1] * oracle_mod.weno3(q[0], q[1], q[2])
    for i in range(2, n - 3):
        assert f[i] == up[i] * oracle_mod.weno5(*q[i - 2:i + 3])
    assert f[n - 3] == up[n - 3] * oracle_mod.weno5(*q[n - 5:n])
    assert f[n - 2] == up[n - 2] * oracle_mod.weno3(q[n - 3], q[n - 2], q[n - 1])
    um = -up
    g = oracle_mod.flux1d(um, q)
    assert g[-1] == 0.0
    assert g[n - 2] == um[n - 2] * q[n - 1]
    assert g[0] == um[0] * oracle_mod.weno3(q[2], q[1], q[0])
    assert g[1] == um[1] * oracle_mod.weno5(q[4], q[3], q[2], q[1], q[0])


def test_flux1d_mirror_symmetry(oracle_mod):
    """Reversing the line and the velocity sign mirrors the interior fluxes:
    the flux at face i with u > 0 equals minus the flux at the mirrored face
    with -u (both use the same five values in the same order)."""
    rng = np.random.default_rng(2)
    n = 16
    q = rng.standard_normal(n)
    u = np.abs(rng.standard_normal(n)) + 0.1
    f = oracle_mod.flux1d(u, q)
    # face i sits between cells i and i+1; its mirror is face n-2-i of the
    # reversed line, whose velocity there is -u[i]
    qr = q[::-1].copy()
    ur = np.empty(n)
    ur[:n - 1] = -u[:n - 1][::-1]
    ur[n - 1] = 1.0
    g = oracle_mod.flux1d(ur, qr)
    for i in range(2, n - 4):
        assert g[n - 2 - i] == -f[i]
