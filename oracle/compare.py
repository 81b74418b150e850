"""Pair an oracle model with a device model on the same inputs
(TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py).
"""
import numpy as np

import oracle as O

# (x, y, z) periodicity of each geometry
GEOMETRY_XYZ = {g: tuple(reversed(p)) for g, p in O.GEOMETRY_PERIODIC.items()}


def make_pair(nx, ny, nz, geometry, Lx=1.0, Ly=1.0, Lz=1.0, rotating=False, coriolis=1.0,
              mg_maxite=20, mg_tol=1e-6, mg_omega=0.9, precision=0, viscosity=0.0,
              n_tracers=0, model="LES", upwind_a=0, upwind_vf=0):
    from nyles_amd import _lib as L
    from nyles_amd.device import DeviceModel
    orc = O.OracleLES(nx, ny, nz, Lx, Ly, Lz, geometry, rotating=rotating, coriolis=coriolis,
                      mg_maxite=mg_maxite, mg_tol=mg_tol, mg_omega=mg_omega,
                      viscosity=viscosity, n_tracers=n_tracers, model=model,
                      upwind_a=upwind_a, upwind_vf=upwind_vf)
    model_id = {"LES": L.NYL_MODEL_LES, "linear": L.NYL_MODEL_LINEAR,
                "Euler3d": L.NYL_MODEL_EULER}[model]
    dev = DeviceModel(nx, ny, nz, GEOMETRY_XYZ[geometry], orc.dx, orc.dy, orc.dz,
                      idx2=orc.idx2, idy2=orc.idy2, idz2=orc.idz2, dx2=orc.dx ** 2,
                      fparameter=orc.fparameter, mg_maxite=mg_maxite, mg_tol=mg_tol,
                      mg_omega=mg_omega, precision=precision, viscosity=viscosity,
                      n_tracers=n_tracers, model=model_id, upwind_a=upwind_a,
                      upwind_vf=upwind_vf)
    return orc, dev


def random_state(orc, seed, amp_u=1e-2, amp_b=1.0):
    """Seeded interior state: b, and u, v, w with the wall faces at zero
    (the wall boundary condition, core/fortran_vortex_force.f90:42)."""
    rng = np.random.default_rng(seed)
    shape = (orc.nz, orc.ny, orc.nx)
    b = amp_b * rng.standard_normal(shape)
    u = amp_u * rng.standard_normal(shape)
    v = amp_u * rng.standard_normal(shape)
    w = amp_u * rng.standard_normal(shape)
    pz, py, px = orc.perio
    if not px:
        u[:, :, -1] = 0.0
    if not py:
        v[:, -1, :] = 0.0
    if not pz:
        w[-1, :, :] = 0.0
    return {"b": b, "u_i": u, "u_j": v, "u_k": w}


def load_state(orc, dev, st):
    from nyles_amd import _lib as L
    for q in range(len(orc.tracers)):
        name = "t%d" % q
        if name in st:
            orc.set_interior(orc.tracers[name], st[name])
            dev.upload(L.NYL_F_TRACER(q), st[name])
    orc.set_interior(orc.b, st["b"])
    orc.set_interior(orc.u["i"], st["u_i"])
    orc.set_interior(orc.u["j"], st["u_j"])
    orc.set_interior(orc.u["k"], st["u_k"])
    dev.upload(L.NYL_F_B, st["b"])
    dev.upload(L.NYL_F_U, st["u_i"])
    dev.upload(L.NYL_F_V, st["u_j"])
    dev.upload(L.NYL_F_W, st["u_k"])


def oracle_fields(orc):
    I = orc.interior
    out = {"b": I(orc.b), "u_i": I(orc.u["i"]), "u_j": I(orc.u["j"]), "u_k": I(orc.u["k"]),
           "ke": I(orc.ke), "vor_i": I(orc.vor["i"]), "vor_j": I(orc.vor["j"]),
           "vor_k": I(orc.vor["k"]), "U_i": I(orc.U["i"]), "U_j": I(orc.U["j"]),
           "U_k": I(orc.U["k"])}
    for name, t in orc.tracers.items():
        out[name] = I(t)
    return out


def device_fields(dev):
    from nyles_amd import _lib as L
    dev.apply(L.NYL_OP_U_FROM_U)   # U is filled on demand
    ids = {"b": L.NYL_F_B, "u_i": L.NYL_F_U, "u_j": L.NYL_F_V, "u_k": L.NYL_F_W,
           "ke": L.NYL_F_KE, "vor_i": L.NYL_F_VORX, "vor_j": L.NYL_F_VORY,
           "vor_k": L.NYL_F_VORZ, "U_i": L.NYL_F_UCX, "U_j": L.NYL_F_UCY, "U_k": L.NYL_F_UCZ}
    for q in range(getattr(dev, "n_tracers", 0)):
        ids["t%d" % q] = L.NYL_F_TRACER(q)
    return {k: dev.download(v) for k, v in ids.items()}


def rel_l2(got, want):
    """||got - want|| / ||want|| (0 when both vanish)"""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    den = float(np.sqrt(np.sum(want ** 2)))
    num = float(np.sqrt(np.sum((got - want) ** 2)))
    return num / den if den > 0 else num


def mismatches(a, b):
    """Number of elements that differ (bitwise up to the sign of zero)."""
    a = np.asarray(a)
    b = np.asarray(b)
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    return int((~same).sum())


def max_rel(a, b):
    den = max(np.max(np.abs(a)), 1e-300)
    return float(np.max(np.abs(a - b)) / den)
