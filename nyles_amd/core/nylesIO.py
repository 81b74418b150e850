"""History output (interface of core/nylesIO.py).

The reference writes netCDF (netCDF4 is not installed here); this writer
keeps the same paths, modes and history schedule and stores each frame as a
compressed NumPy archive next to where the netCDF file would be:
``<datadir>/<expname>/<expname>_00_hist_<frame>.npz`` holding the selected
variables (interior cells, or the full arrays with ``include_halo``), t and
n.
"""
import os as _os
import sys as _sys
_ROOT = _os.path.dirname(_os.path.dirname(_os.path.dirname(_os.path.abspath(__file__))))
if _ROOT not in _sys.path:
    _sys.path.insert(0, _ROOT)
import os
import shutil

import numpy as np

from nyles_amd.core import mpitools


class NylesIO(object):
    def __init__(self, param):
        self.disk_limit = param["disk_space_warning"]
        self.simplified_grid = param["simplified_grid"]
        self.variables_in_history = param["variables_in_history"]
        self.dt_hist = param["timestep_history"]
        self.include_halo = param["include_halo"]
        self.t_next_hist = 0.0
        self.n_hist = 0
        self.last_saved_frame = None
        self.myrank = param["myrank"]
        self.n_tracers = param["n_tracers"]
        datadir = os.path.expanduser(param["datadir"])
        expname = param["expname"]
        out_dir = os.path.join(datadir, expname)
        if param["mode"] == "overwrite":
            pass
        elif param["mode"] == "count":
            counter = 0
            full = "{}_{:04d}".format(expname, counter)
            out_dir = os.path.join(datadir, full)
            while os.path.exists(out_dir):
                counter += 1
                full = "{}_{:04d}".format(expname, counter)
                out_dir = os.path.join(datadir, full)
            expname = full
        elif param["mode"] == "continue":
            if not os.path.exists(out_dir):
                raise ValueError("cannot load experiment to continue: no directory {!r}"
                                 .format(out_dir))
            raise NotImplementedError("continuing an experiment is not implemented")
        else:
            raise ValueError("unknown mode: {}".format(param["mode"]))
        self.expname = expname
        self.hist_path = os.path.join(out_dir, expname + "_%02i" % param["myrank"] + "_hist")
        self.script_path = os.path.join(out_dir, expname + ".py")
        self.output_directory = out_dir
        if not os.path.isdir(out_dir):
            os.makedirs(out_dir)

    def init(self, state, grid, t=0.0, n=0):
        v = self.variables_in_history
        if v == "all":
            v = list(state.toc.keys())
        elif v == "prognostic":
            v = state.get_prognostic_variables()
        elif v == "p+p":
            v = state.get_prognostic_variables() + ["p"]
        elif any(x not in state.toc for x in v):
            raise ValueError("a variable chosen for the history file does not exist in this "
                             "model.  The available variables are " + str(state.toc))
        if len(v) == 0:
            raise ValueError("no variables are selected for the history file")
        self.variables_in_history = list(v)
        mpitools.barrier()
        self.write_history_file(state, t, n)
        self.t_next_hist = t + self.dt_hist

    def write(self, state, t, n):
        if t < self.t_next_hist:
            return False
        self.write_history_file(state, t, n)
        self.t_next_hist += self.dt_hist
        if self.disk_limit > 0:
            try:
                free = shutil.disk_usage(self.output_directory).free / 1024 ** 3
            except OSError:
                self.disk_limit = 0
                return False
            if free < self.disk_limit:
                print("\nWarning, low disk space: {:.2f} GB left in {}".format(
                    free, self.output_directory))
                self.disk_limit = 0
        return False

    def finalize(self, state, t, n):
        if n != self.last_saved_frame:
            self.write_history_file(state, t, n)

    def _arrays(self, state):
        out = {}
        for name in self.variables_in_history:
            obj = state.get(name)
            comps = [(name, obj)] if state.toc[name] == "scalar" else \
                [("%s_%s" % (name, d), obj[d]) for d in "ijk"]
            for key, s in comps:
                a = s.view("i")
                out[key] = np.array(a if self.include_halo else s.interior(a))
        return out

    def write_history_file(self, state, t, n):
        path = "%s_%05d.npz" % (self.hist_path, self.n_hist)
        np.savez_compressed(path, t=t, n=n, **self._arrays(state))
        self.n_hist += 1
        self.last_saved_frame = n

    def backup_scriptfile(self, filename):
        if filename and os.path.isfile(filename):
            shutil.copyfile(filename, self.script_path)

    def write_githashnumber(self):
        pass
