"""Grid spacings and staggered coordinates (interface of core/grid.py).

Spacings and metric terms are computed with the same Python expressions as
the reference (core/grid.py:30-58) so that the doubles handed to the device
are identical; coordinates follow the staggering of core/grid.py:60-158
(b at cell centres, u_d on the right d-face, vor_d on the d-edges).
"""
import os as _os
import sys as _sys
_ROOT = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
if _ROOT not in _sys.path:
    _sys.path.insert(0, _ROOT)
import numpy as np

from nyles_amd.core.variables import Scalar, Vector


class Grid(object):
    def __init__(self, param):
        self.nx, self.ny, self.nz = param["nx"], param["ny"], param["nz"]
        self.npx, self.npy, self.npz = param["npx"], param["npy"], param["npz"]
        self.Lx, self.Ly, self.Lz = param["Lx"], param["Ly"], param["Lz"]

        self.dx = self.Lx / (self.npx * self.nx)
        self.dy = self.Ly / (self.npy * self.ny)
        self.dz = self.Lz / (self.npz * self.nz)
        self.dx2 = self.dx ** 2
        self.dy2 = self.dy ** 2
        self.dz2 = self.dz ** 2
        self.idx = 1 / self.dx
        self.idy = 1 / self.dy
        self.idz = 1 / self.dz
        self.idx2 = 1 / self.dx ** 2
        self.idy2 = 1 / self.dy ** 2
        self.idz2 = 1 / self.dz ** 2
        self.vol = self.dx * self.dy * self.dz
        self.ids2 = {"i": self.idx2, "j": self.idy2, "k": self.idz2}
        self.vol_per_ds2 = {"i": self.vol / self.dx2, "j": self.vol / self.dy2,
                            "k": self.vol / self.dz2}

        self.x_b = Scalar(param, "x at cell centers", "x_b", "L")
        self.y_b = Scalar(param, "y at cell centers", "y_b", "L")
        self.z_b = Scalar(param, "z at cell centers", "z_b", "L")
        size = self.x_b.size
        self.size = size
        self.domainindices = self.x_b.domainindices
        k0, k1, j0, j1, i0, i1 = self.domainindices
        loc = param.get("loc", [0, 0, 0])
        x0 = loc[2] * self.nx * self.dx
        y0 = loc[1] * self.ny * self.dy
        z0 = loc[0] * self.nz * self.dz
        self.x_b_1D = (np.arange(size["i"]) + 0.5 - i0) * self.dx + x0
        self.y_b_1D = (np.arange(size["j"]) + 0.5 - j0) * self.dy + y0
        self.z_b_1D = (np.arange(size["k"]) + 0.5 - k0) * self.dz + z0
        self._mesh((self.z_b, self.y_b, self.x_b), self.z_b_1D, self.y_b_1D, self.x_b_1D)

        self.x_vel = Vector(param, "x at cell faces", "x_vel", "L", is_velocity=True)
        self.y_vel = Vector(param, "y at cell faces", "y_vel", "L", is_velocity=True)
        self.z_vel = Vector(param, "z at cell faces", "z_vel", "L", is_velocity=True)
        hx, hy, hz = self.dx / 2, self.dy / 2, self.dz / 2
        self.x_u_1D, self.y_u_1D, self.z_u_1D = self.x_b_1D + hx, self.y_b_1D.copy(), self.z_b_1D.copy()
        self.x_v_1D, self.y_v_1D, self.z_v_1D = self.x_b_1D.copy(), self.y_b_1D + hy, self.z_b_1D.copy()
        self.x_w_1D, self.y_w_1D, self.z_w_1D = self.x_b_1D.copy(), self.y_b_1D.copy(), self.z_b_1D + hz
        for d, (zz, yy, xx) in zip("ijk", ((self.z_u_1D, self.y_u_1D, self.x_u_1D),
                                           (self.z_v_1D, self.y_v_1D, self.x_v_1D),
                                           (self.z_w_1D, self.y_w_1D, self.x_w_1D))):
            self._mesh((self.z_vel[d], self.y_vel[d], self.x_vel[d]), zz, yy, xx)

        self.x_vor = Vector(param, "x at cell edges", "x_vor", "L", is_velocity=False)
        self.y_vor = Vector(param, "y at cell edges", "y_vor", "L", is_velocity=False)
        self.z_vor = Vector(param, "z at cell edges", "z_vor", "L", is_velocity=False)
        self.x_vor_i_1D, self.y_vor_i_1D, self.z_vor_i_1D = self.x_b_1D.copy(), self.y_b_1D + hy, self.z_b_1D + hz
        self.x_vor_j_1D, self.y_vor_j_1D, self.z_vor_j_1D = self.x_b_1D + hx, self.y_b_1D.copy(), self.z_b_1D + hz
        self.x_vor_k_1D, self.y_vor_k_1D, self.z_vor_k_1D = self.x_b_1D + hx, self.y_b_1D + hy, self.z_b_1D.copy()
        for d, (zz, yy, xx) in zip("ijk", ((self.z_vor_i_1D, self.y_vor_i_1D, self.x_vor_i_1D),
                                           (self.z_vor_j_1D, self.y_vor_j_1D, self.x_vor_j_1D),
                                           (self.z_vor_k_1D, self.y_vor_k_1D, self.x_vor_k_1D))):
            self._mesh((self.z_vor[d], self.y_vor[d], self.x_vor[d]), zz, yy, xx)

    @staticmethod
    def _mesh(targets, z1, y1, x1):
        zs, ys, xs = (t.view("i") for t in targets)
        zs[...], ys[...], xs[...] = np.meshgrid(z1, y1, x1, indexing="ij")
