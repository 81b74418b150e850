"""The oracle's mgfor multigrid, pinned by library Poisson solvers, closed
forms of its coefficients and its hierarchy rule (CPU)."""
import ctypes

import numpy as np
import pytest
import scipy.fft as sfft

NH = 3


def _interior(a):
    return a[NH:-NH, NH:-NH, NH:-NH]


def _solve(oracle_mod, geometry, rhs, maxite=60, tol=1e-26):
    nz, ny, nx = rhs.shape
    mg = oracle_mod.OracleMG(nx, ny, nz, geometry, maxite=maxite, tol=tol)
    perio = oracle_mod.GEOMETRY_PERIODIC[geometry]
    sub = np.zeros(tuple(n + 2 * NH if p else n for n, p in zip(rhs.shape, perio)))
    sl = tuple(slice(NH, -NH) if p else slice(None) for p in perio)
    sub[sl] = rhs
    oracle_mod.halo_fill(sub, geometry, NH)
    mg.b[mg.idx] = sub
    nite = mg.solve()
    return mg, _interior(mg.x).copy(), nite


def test_periodic_solution_matches_fft(oracle_mod):
    """perio_xyz: sum6(x) - 6x = b solved by FFT (eigenvalues 2cos(k)-2)."""
    n = 32
    rng = np.random.default_rng(0)
    b = rng.standard_normal((n, n, n))
    b -= b.mean()
    mg, x, nite = _solve(oracle_mod, "perio_xyz", b)
    k = 2 * np.pi * np.fft.fftfreq(n)
    lam = (2 * np.cos(k)[:, None, None] - 2) + (2 * np.cos(k)[None, :, None] - 2) \
        + (2 * np.cos(k)[None, None, :] - 2)
    lam[0, 0, 0] = 1.0
    bh = sfft.fftn(b)
    bh[0, 0, 0] = 0.0
    xe = np.real(sfft.ifftn(bh / lam))
    x = x - x.mean()
    assert np.max(np.abs(x - xe)) < 1e-9 * np.max(np.abs(xe)), nite


def test_closed_solution_matches_dct(oracle_mod):
    """closed: Neumann 7-point Laplacian diagonalised by DCT-II."""
    nz, ny, nx = 16, 16, 32
    rng = np.random.default_rng(1)
    b = rng.standard_normal((nz, ny, nx))
    b -= b.mean()
    mg, x, nite = _solve(oracle_mod, "closed", b)
    lam = sum((2 * np.cos(np.pi * np.arange(m) / m) - 2).reshape(sh)
              for m, sh in ((nz, (-1, 1, 1)), (ny, (1, -1, 1)), (nx, (1, 1, -1))))
    lam[0, 0, 0] = 1.0
    bh = sfft.dctn(b, type=2, norm="ortho")
    bh[0, 0, 0] = 0.0
    xe = sfft.idctn(bh / lam, type=2, norm="ortho")
    x = x - x.mean()
    assert np.max(np.abs(x - xe)) < 1e-9 * np.max(np.abs(xe)), nite


def test_mixed_walls_match_dct_fft(oracle_mod):
    """perio_x: FFT along x, DCT-II along y and z."""
    nz, ny, nx = 16, 16, 32
    rng = np.random.default_rng(2)
    b = rng.standard_normal((nz, ny, nx))
    b -= b.mean()
    mg, x, nite = _solve(oracle_mod, "perio_x", b)
    lz = (2 * np.cos(np.pi * np.arange(nz) / nz) - 2)[:, None, None]
    ly = (2 * np.cos(np.pi * np.arange(ny) / ny) - 2)[None, :, None]
    lx = (2 * np.cos(2 * np.pi * np.fft.fftfreq(nx)) - 2)[None, None, :]
    lam = lz + ly + lx
    lam[0, 0, 0] = 1.0
    bh = sfft.fft(sfft.dctn(b, type=2, norm="ortho", axes=(0, 1)), axis=2)
    bh[0, 0, 0] = 0.0
    xe = np.real(sfft.idctn(sfft.ifft(bh / lam, axis=2), type=2, norm="ortho", axes=(0, 1)))
    x = x - x.mean()
    assert np.max(np.abs(x - xe)) < 1e-9 * np.max(np.abs(xe)), nite


@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_xy", "perio_xyz"])
def test_coefficient_closed_forms(oracle_mod, geometry):
    """diag = wet neighbours, idiag = 1/diag, Rcoef = 1/2, Pcoef =
    1/((3+mx)(3+my)(3+mz)) on every level (core/mgfor/operators.f90:299-457)."""
    n = (16, 16, 16)
    mg = oracle_mod.OracleMG(*n, geometry)
    perio = dict(zip("zyx", oracle_mod.GEOMETRY_PERIODIC[geometry]))
    for lev in range(mg.nlevels):
        nz, nyh, nxh = mg.level_shape(lev)
        shp = (nz - 2 * NH, nyh - 2 * NH, nxh - 2 * NH)
        diag = _interior(mg.array(lev, 5))
        idiag = _interior(mg.array(lev, 6))
        msk = _interior(mg.array(lev, 7))
        assert np.all(msk == 1)
        wet = np.zeros(shp)
        for ax, name in enumerate("zyx"):
            m = shp[ax]
            idx = np.arange(m)
            lo = np.ones(m) if perio[name] else (idx > 0).astype(float)
            hi = np.ones(m) if perio[name] else (idx < m - 1).astype(float)
            sh = [1, 1, 1]
            sh[ax] = m
            wet = wet + (lo + hi).reshape(sh)
        np.testing.assert_array_equal(diag, wet)
        np.testing.assert_array_equal(idiag, 1.0 / wet)
        if lev > 0:
            np.testing.assert_array_equal(_interior(mg.array(lev, 8)), 0.5)
        if lev < mg.nlevels - 1:
            P = _interior(mg.array(lev, 9))
            f = np.ones(shp)
            for ax, name in enumerate("zyx"):
                m = shp[ax]
                idx = np.arange(m)
                if perio[name]:
                    side = np.ones(m)
                else:
                    # the coarse neighbour on the child's side exists?
                    c = idx // 2
                    nb = np.where(idx % 2 == 0, c - 1, c + 1)
                    side = ((nb >= 0) & (nb < m // 2)).astype(float)
                sh = [1, 1, 1]
                sh[ax] = m
                f = f * (3 + side).reshape(sh)
            np.testing.assert_array_equal(P, 1.0 / f)


def test_hierarchy_rule(oracle_mod):
    """Levels halve x and y until one of them is 2; z (with its in-array
    halo) follows z -> z/2 + nh (core/mgfor/mg_setup.f90:262-303)."""
    L = oracle_mod.lib()
    lx, ly, lz = (ctypes.c_int * 50)(), (ctypes.c_int * 50)(), (ctypes.c_int * 50)()
    cases = {(256, 256, 256): 8, (128, 128, 128): 7, (256, 256, 512): 8, (64, 32, 32): 5}
    for (nx, ny, nz), nlev in cases.items():
        n = L.orc_mg_hierarchy(nx, ny, nz, NH, lx, ly, lz, 50)
        assert n == nlev
        assert min(lx[n - 1], ly[n - 1]) == 2
        for l in range(n):
            assert lx[l] == nx >> l and ly[l] == ny >> l
            assert lz[l] - 2 * NH == nz >> l
    # a size that never reaches 2 is refused (the reference loops forever)
    assert L.orc_mg_hierarchy(48, 48, 48, NH, lx, ly, lz, 50) == -1


def test_residual_decreases_every_cycle(oracle_mod):
    n = 32
    rng = np.random.default_rng(3)
    b = rng.standard_normal((n, n, n))
    b -= b.mean()
    mg = oracle_mod.OracleMG(n, n, n, "perio_xyz", maxite=1, tol=0.0)
    sub = np.zeros((n + 6,) * 3)
    sub[NH:-NH, NH:-NH, NH:-NH] = b
    oracle_mod.halo_fill(sub, "perio_xyz", NH)
    mg.b[...] = sub
    prev = np.inf
    for _ in range(8):
        mg.solve()
        res = mg.last()[1]
        assert res < prev
        prev = res


def test_warm_start_idempotence(oracle_mod):
    """Solving the same b again from the converged x needs no V-cycle."""
    n = 16
    rng = np.random.default_rng(4)
    b = rng.standard_normal((n, n, n))
    b -= b.mean()
    mg, x, nite = _solve(oracle_mod, "perio_xyz", b, maxite=40, tol=1e-6)
    assert nite > 0
    assert mg.solve() == 0
