"""Synthetic initial states of the benchmark and test workloads.

All fields are interior arrays ``[z, y, x]`` in FP64 generated from
``numpy.random.default_rng(seed)``.  Velocities are covariant (physical
velocity times the grid spacing), as the reference's experiment scripts set
them (experiments/taylorgreen/tgv.py:84-85,
experiments/shear_instabilities/kelvin_helmholtz.py:75).  Staggered
positions follow core/grid.py:60-124: b at cell centres, u_i on the right x
face, u_j on the right y face, u_k on the right z face.
"""
import numpy as np


def centres(n, d):
    return (np.arange(n) + 0.5) * d


def taylor_green(n, seed=2, noise=1e-4, L=2 * np.pi):
    """Taylor-Green vortex on a triply periodic n^3 box of side L:
    u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 (+ N(0, noise))."""
    d = L / n
    xb = centres(n, d)
    xf = xb + d / 2
    rng = np.random.default_rng(seed)
    z = xb[:, None, None]
    yb, yf = xb[None, :, None], xf[None, :, None]
    xb3, xf3 = xb[None, None, :], xf[None, None, :]
    u = np.sin(xf3) * np.cos(yb) * np.cos(z)
    v = -np.cos(xb3) * np.sin(yf) * np.cos(z)
    u = u + noise * rng.standard_normal((n, n, n))
    v = v + noise * rng.standard_normal((n, n, n))
    w = noise * rng.standard_normal((n, n, n))
    b = np.zeros((n, n, n))
    return {"b": b, "u_i": u * d, "u_j": v * d, "u_k": w * d}, {"L": (L, L, L),
                                                                "geometry": "perio_xyz"}


def lock_exchange(nx=64, ny=32, nz=32, Lx=16.0, Ly=8.0, Lz=8.0, seed=1, noise=1e-3):
    """Lock exchange: b = +0.5 for x < Lx/2, -0.5 otherwise, plus uniform
    noise; velocity at rest; periodic in x, walls in y and z."""
    rng = np.random.default_rng(seed)
    x = centres(nx, Lx / nx)[None, None, :]
    b = np.where(x < Lx / 2, 0.5, -0.5) * np.ones((nz, ny, nx))
    b = b + noise * rng.uniform(-1.0, 1.0, (nz, ny, nx))
    zero = np.zeros((nz, ny, nx))
    return {"b": b, "u_i": zero.copy(), "u_j": zero.copy(), "u_k": zero.copy()}, \
        {"L": (Lx, Ly, Lz), "geometry": "perio_x"}


def shear_layer(n, seed=1000, delta_frac=0.05, N2=0.1, noise=1e-3, L=1.0):
    """Stratified shear layer: u = tanh((z - Lz/2)/delta), b = N^2 z, plus
    noise; periodic in x and y, walls in z."""
    rng = np.random.default_rng(seed)
    d = L / n
    z = centres(n, d)[:, None, None]
    u = np.tanh((z - L / 2) / (delta_frac * L)) * np.ones((n, n, n))
    u = u + noise * rng.standard_normal((n, n, n))
    b = N2 * z * np.ones((n, n, n))
    w = noise * rng.standard_normal((n, n, n))
    w[-1] = 0.0
    return {"b": b, "u_i": u * d, "u_j": np.zeros((n, n, n)), "u_k": w * d}, \
        {"L": (L, L, L), "geometry": "perio_xy"}


def rayleigh_taylor(nx=256, ny=256, nz=512, seed=3, nmodes=16, amp_frac=0.01, L=1.0):
    """Rayleigh-Taylor: heavy fluid on top, b = -0.5 above z = Lz/2 + eta(x, y)
    and +0.5 below,
    eta a sum of nmodes seeded Fourier modes of amplitude amp_frac*Lz;
    periodic in x and y, walls in z, Lz = 2 L."""
    rng = np.random.default_rng(seed)
    Lx = Ly = L
    Lz = 2 * L
    x = centres(nx, Lx / nx)[None, :]
    y = centres(ny, Ly / ny)[:, None]
    eta = np.zeros((ny, nx))
    for _ in range(nmodes):
        kx, ky = rng.integers(1, 9, size=2)
        ph = rng.uniform(0, 2 * np.pi)
        eta += np.cos(2 * np.pi * (kx * x / Lx + ky * y / Ly) + ph)
    eta *= amp_frac * Lz / np.sqrt(nmodes)
    z = centres(nz, Lz / nz)[:, None, None]
    b = np.where(z > Lz / 2 + eta[None], -0.5, 0.5)
    zero = np.zeros((nz, ny, nx))
    return {"b": b, "u_i": zero.copy(), "u_j": zero.copy(), "u_k": zero.copy()}, \
        {"L": (Lx, Ly, Lz), "geometry": "perio_xy"}
