"""Subdomain geometry: neighbours and array shapes.

Same interface and semantics as core/mpi/topology.py of the reference:
the module-level ``topology`` string selects which axes are periodic, the
neighbour dictionary is keyed by direction tuples (dk, dj, di), and a field
carries a halo of ``nh`` cells only towards existing neighbours
(core/mpi/topology.py:253-304).  On one GPU there is one subdomain, so every
periodic neighbour is the subdomain itself.
"""
import itertools

import numpy as np

topology = "undefined"

GEOMETRIES = ["closed", "perio_x", "perio_xy", "perio_y", "perio_xyz"]


def rank2loc(rank, procs):
    """(k, j, i) location of a rank; procs = [npz, npy, npx]."""
    return [rank // (procs[2] * procs[1]), (rank // procs[2]) % procs[1], rank % procs[2]]


def loc2rank(loc, procs):
    return (loc[0] * procs[1] + loc[1]) * procs[2] + loc[2]


def periodic_axes(geometry=None):
    """(z, y, x) periodicity of a geometry name."""
    g = topology if geometry is None else geometry
    if g not in GEOMETRIES:
        raise AssertionError("you forgot to set the topology")
    return ("z" in g, "y" in g, "x" in g)


def get_neighbours(location, procs, incr=(1, 1, 1), extension=26):
    """Neighbour ranks keyed by direction (dk, dj, di); a direction that
    leaves a non-periodic axis has no entry."""
    perio = periodic_axes()
    if isinstance(location, int):
        loc = rank2loc(location, procs)
    elif isinstance(location, (list, tuple)):
        loc = list(location)
    else:
        raise ValueError(location)
    if extension == 6:
        keep = lambda d: sum(map(abs, d)) == 1
    elif extension == 18:
        keep = lambda d: sum(map(abs, d)) in (1, 2)
    elif extension == 26:
        keep = lambda d: sum(map(abs, d)) >= 1
    else:
        raise ValueError("neighbours should be 6, 18 or 26")
    ngs = {}
    for d in itertools.product([-1, 0, 1], repeat=3):
        if not keep(d):
            continue
        pos = [loc[a] + d[a] * incr[a] for a in range(3)]
        if any(not perio[a] and not (0 <= pos[a] < procs[a]) for a in range(3)):
            continue
        ngs[d] = loc2rank([pos[a] % procs[a] for a in range(3)], procs)
    return ngs


def get_variable_shape(innersize, ngbs, nh):
    """Extended shape of a field and its interior index range
    (k0, k1, j0, j1, i0, i1): halo only towards existing neighbours."""
    size = list(innersize)
    lo = []
    for a, (m, p) in enumerate((((-1, 0, 0), (1, 0, 0)), ((0, -1, 0), (0, 1, 0)),
                                ((0, 0, -1), (0, 0, 1)))):
        start = 0
        if m in ngbs:
            size[a] += nh
            start = nh
        end = size[a]
        if p in ngbs:
            size[a] += nh
        lo.append((start, end))
    domainindices = (lo[0][0], lo[0][1], lo[1][0], lo[1][1], lo[2][0], lo[2][1])
    return size, domainindices


def check_graph(allneighbours):
    """The neighbour graph of all ranks must be symmetric."""
    n = len(allneighbours)
    conn = np.zeros((n, n), dtype=int)
    for i, ngs in enumerate(allneighbours):
        for j in ngs.values():
            if j is not None:
                conn[i, j] += 1
    assert (conn == conn.T).all(), "neighbours are not connected symmetrically"
