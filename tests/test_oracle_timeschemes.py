"""The oracle's time schemes on a linear test equation ds/dt = lam*s (CPU):
one step must multiply s by the scheme's amplification factor
(core/timescheme.py:113-221)."""
import numpy as np
import pytest


def _linear_model(oracle_mod, lam):
    m = oracle_mod.OracleLES(8, 8, 8, 1.0, 1.0, 1.0, "perio_xyz")

    def rhs(ds, last=False, forcing=None, t=0.0):
        ds["b"][...] = lam * m.b
        for d in "ijk":
            ds["u"][d][...] = lam * m.u[d]

    m.rhs = rhs
    m.diagnose_var = lambda: None
    rng = np.random.default_rng(0)
    m.b[...] = rng.standard_normal(m.shape)
    for d in "ijk":
        m.u[d][...] = rng.standard_normal(m.shape)
    return m


@pytest.mark.parametrize("z", [-0.3, 0.1, -1.2])
def test_rk3_ssp_amplification(oracle_mod, z):
    dt = 0.05
    m = _linear_model(oracle_mod, z / dt)
    s0 = m.b.copy()
    m.rk3_ssp(dt)
    G = 1 + z + z ** 2 / 2 + z ** 3 / 6
    np.testing.assert_allclose(m.b, G * s0, rtol=1e-13)


def test_euler_forward_amplification(oracle_mod):
    dt, z = 0.1, -0.4
    m = _linear_model(oracle_mod, z / dt)
    s0 = m.b.copy()
    m.euler_forward(dt)
    np.testing.assert_allclose(m.b, (1 + z) * s0, rtol=1e-14)


def test_lfam3_two_steps(oracle_mod):
    """First step Euler; second step LF predictor to n+1 blended by AM3 to
    n+1/2, corrector from n with the slope at n+1/2."""
    dt, z = 0.1, -0.2
    lam = z / dt
    m = _linear_model(oracle_mod, lam)
    s0 = m.b.copy()
    m.lfam3(dt)
    s1 = (1 + z) * s0
    np.testing.assert_allclose(m.b, s1, rtol=1e-14)
    m.lfam3(dt)
    lf = s0 + 2 * z * s1                    # leap-frog to n+1
    half = (5 * lf + 8 * s1 - s0) / 12      # AM3 interpolation at n+1/2
    s2 = s1 + z * half
    np.testing.assert_allclose(m.b, s2, rtol=1e-13)
