"""Wall-clock timers per function (interface of core/timing.py).

``@timing`` records the host wall time of each call for the first 100
``forward`` calls; device kernels are asynchronous, so these times measure
the host side of each call (use rocprofv3 for kernel times).
"""
import os
import pickle
from functools import wraps
from time import time

import numpy as np

stats = {}


def timing(f):
    @wraps(f)
    def wrap(*args, **kw):
        if len(stats.get("forward", [])) >= 100:
            return f(*args, **kw)
        ts = time()
        result = f(*args, **kw)
        te = time()
        if "timing" not in kw:
            stats.setdefault(f.__name__, []).append(te - ts)
        return result
    return wrap


def write_timings(path):
    with open(os.path.join(path, "timing.pkl"), "wb") as fid:
        pickle.dump(stats, fid)


def analyze_timing(path):
    """Print the mean time per call of every timed function."""
    fname = os.path.join(path, "timing.pkl")
    if not os.path.exists(fname):
        return
    with open(fname, "rb") as fid:
        data = pickle.load(fid)
    for name, vals in sorted(data.items(), key=lambda kv: -np.mean(kv[1])):
        print("%-20s %10.3e s  (%d calls)" % (name, float(np.mean(vals)), len(vals)))
