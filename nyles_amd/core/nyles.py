"""The Nyles main class (interface of core/nyles.py) on one GPU.

``Nyles(param)`` checks and freezes the parameters, derives the subdomain
size, neighbours and grid, creates the output folder and the model;
``run()`` is the reference's time loop (core/nyles.py:119-225):
compute_dt -> model.forward -> history output -> progress line -> blow-up
test.  The time step runs on the GPU; the CFL reduction is a device
reduction of the same quantity (core/nyles.py:249).

One GPU holds one subdomain: npx = npy = npz = 1.
"""
import os as _os
import sys as _sys
_ROOT = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
if _ROOT not in _sys.path:
    _sys.path.insert(0, _ROOT)
import pickle
import sys
from time import time

import numpy as np

from nyles_amd.core import grid as grid_mod
from nyles_amd.core import model_les
from nyles_amd.core import mpitools
from nyles_amd.core import nylesIO
from nyles_amd.core import timing
from nyles_amd.core import topology as topo
from nyles_amd.core.online_diag import VFwork
from nyles_amd.core.parameters import UserParameterError


class Nyles(object):
    def __init__(self, user_param):
        user_param.check()
        user_param.freeze()
        param = user_param.view_parameters()
        self.param = param
        npx, npy, npz = param["npx"], param["npy"], param["npz"]
        if (npx, npy, npz) != (1, 1, 1):
            raise UserParameterError(
                "one GPU runs one subdomain: set npx = npy = npz = 1 (got %d, %d, %d)"
                % (npx, npy, npz))
        param["nx"] = param["global_nx"] // npx
        param["ny"] = param["global_ny"] // npy
        param["nz"] = param["global_nz"] // npz
        topo.topology = param["geometry"]
        procs = [npz, npy, npx]
        myrank = mpitools.get_myrank(procs)
        loc = topo.rank2loc(myrank, procs)
        neighbours = topo.get_neighbours(loc, procs)
        param["procs"] = procs
        param["myrank"] = myrank
        param["neighbours"] = neighbours
        param["loc"] = loc
        self.myrank = myrank
        if self.myrank == 0:
            print("-" * 80)
        self.banner()
        self.grid = grid_mod.Grid(param)
        self.IO = nylesIO.NylesIO(param)
        with open(_os.path.join(self.IO.output_directory, "param.pkl"), "wb") as fid:
            pickle.dump(param, fid)
        self.initiate(param)

    def initiate(self, param):
        name = param["modelname"]
        if name == "LES":
            self.model = model_les.LES(param, self.grid)
        elif name == "Euler3d":
            self.model = model_les.LES(param, self.grid, euler=True)
        elif name == "linear":
            self.model = model_les.LES(param, self.grid, linear=True)
        else:
            # "advection" (core/model_advection.py): its constructor passes the
            # grid where Tracer_numerics expects param (core/model_advection.py:29-30
            # against core/tracer.py:10), so the reference cannot build it
            raise NotImplementedError("model %r is not available (the reference's "
                                      "Advection model cannot be constructed)" % name)
        self.diag = VFwork(self.model, self.grid)
        self.tend = param["tend"]
        self.auto_dt = param["auto_dt"]
        self.dt0 = param["dt"]
        self.cfl = param["cfl"]
        self.dt_max = param["dt_max"]
        if param["show"]:
            print("the live 3-D view is not available; continuing without it")
        self.plotting = None
        self.gridcellpersubdom = param["nx"] * param["ny"] * param["nz"]

    def run(self):
        t = 0.0
        n = 0
        self.model.diagnose_var(self.model.state)
        if self.myrank == 0:
            print("Creating output file:", self.IO.hist_path)
        self.IO.init(self.model.state, self.grid, t, n)
        if self.myrank == 0:
            self.IO.backup_scriptfile(sys.argv[0])
        time_length = len(str(int(self.tend))) + 3
        time_string = "\r" + ", ".join([
            "n = {:3d}",
            "t = {:" + str(time_length) + ".2f}/{:" + str(time_length) + ".2f}",
            "dt = {:.4f}",
            "perf = {:.2e}"])
        if self.myrank == 0:
            print("-" * 80)
        stop = False
        realtime0 = time()
        while not stop:
            dt = self.compute_dt()
            blowup = self.model.forward(t, dt)
            t += dt
            n += 1
            if t >= self.IO.t_next_hist:
                self.diag.compute()
            stop = self.IO.write(self.model.state, t, n)
            if self.myrank == 0:
                realtime = time()
                # wall time per step per grid cell (the reference's metric)
                perf = (realtime - realtime0) / self.gridcellpersubdom
                realtime0 = realtime
                print(time_string.format(n, t, self.tend, dt, perf), end="")
            globalmood = mpitools.global_sum(1.0 if blowup else 0.0)
            if globalmood > 0.0:
                if self.myrank == 0:
                    print("")
                    print("BLOW UP!")
                    print("write a last snapshot at blowup time", end="")
                self.IO.t_next_hist = t
                self.IO.write(self.model.state, t, n)
                stop = True
            if t >= self.tend or stop:
                break
            mpitools.barrier()
        if self.myrank == 0:
            print()
            print("-" * 80)
            print("Job is aborted" if stop else "Job completed as expected")
        self.IO.finalize(self.model.state, t, n)
        if self.myrank == 0:
            print("Output written to:", self.IO.output_directory)
        self.model.write_stats(self.IO.output_directory)
        timing.write_timings(self.IO.output_directory)
        self.t = t
        self.n = n

    def compute_dt(self):
        """dt = min(cfl / max|U|, dt_max), dt_max if the flow is at rest, or
        the fixed dt when auto_dt is False (core/nyles.py:227-260)."""
        if self.auto_dt:
            U_max = np.sqrt(self.model.max_U2())
            U_max = mpitools.global_max(U_max)
            if U_max == 0.0:
                return self.dt_max
            return min(self.cfl / U_max, self.dt_max)
        return self.dt0

    def banner(self):
        if self.myrank == 0:
            print("Welcome to Nyles (MI355X time step)")
