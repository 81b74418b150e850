"""Halo filling of host arrays (interface of core/mpi/halo.py on one rank).

On the device the halo is never stored: kernels read periodic neighbours at
their wrapped index.  This class fills the halo of the host mirrors with
the same values (the periodic images), for scripts that work on
``view('i')`` arrays.
"""
import os as _os
import sys as _sys
_ROOT = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
if _ROOT not in _sys.path:
    _sys.path.insert(0, _ROOT)
import numpy as np

from nyles_amd.core.variables import fill_halo


class Halo(object):
    def __init__(self, grid):
        self.neighbours = grid["neighbours"]
        self.nh = grid["nh"]
        self.size = grid["size"]
        self.domainindices = grid["domainindices"]

    def fill(self, thing):
        nature = type(thing).__name__
        if nature == "Scalar":
            self.fillarray(thing.view("i"))
        elif nature == "ndarray":
            self.fillarray(thing)
        elif nature == "Vector":
            for d in "ijk":
                self.fillarray(thing[d].view("i"))
        else:
            raise ValueError("try to fill halo with unidentified object")

    def fillarray(self, x):
        fill_halo(x, self.neighbours, self.nh, self.domainindices)


def set_halo(param, state):
    return Halo({"neighbours": param["neighbours"], "nh": param["nh"],
                 "size": np.shape(state.b.view("i")),
                 "domainindices": state.b.domainindices})
