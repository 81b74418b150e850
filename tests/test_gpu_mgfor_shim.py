"""The mgfor ABI shim (GPU): the call sequence of the reference's MG driver
(core/mgfordriver.py:6-105: get_ptrmg, print_mginfos, get_pyshape,
set_pyarray, solve, get_pyarray), made the way core/build.py makes it
(symbol found by "MOD_" + name, every argument by reference, a double
result), runs on libnyles_hip.so and gives nyl_mg_solve's solution bit for
bit, warm start included; the halo of the returned x holds the periodic
images."""
import ctypes
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOPOLOGY = {"closed": 1, "perio_x": 2, "perio_y": 3, "perio_xy": 5, "perio_xyz": 6}
PERIODIC = {"closed": (0, 0, 0), "perio_x": (1, 0, 0), "perio_y": (0, 1, 0),
            "perio_xy": (1, 1, 0), "perio_xyz": (1, 1, 1)}
NH = 3


def _ptr(thing):
    """core/build.py:44-82: the by-reference conversion of each argument"""
    if isinstance(thing, (tuple, list)):
        return tuple(_ptr(t) for t in thing)
    if isinstance(thing, np.ndarray):
        return thing.ctypes.data_as(ctypes.POINTER(ctypes.c_int32 if thing.dtype == np.int32
                                                   else ctypes.c_double))
    if isinstance(thing, (int, np.int32)):   # bool too: isinstance(True, int)
        return ctypes.byref(ctypes.c_int32(thing))
    if isinstance(thing, np.int64):
        return ctypes.byref(ctypes.c_int64(thing))
    if isinstance(thing, float):
        return ctypes.byref(ctypes.c_double(thing))
    raise ValueError(thing)


def _function(lib_path, name):
    """core/build.py:107-121 + Function: the unique symbol holding
    "MOD_" + name, restype double"""
    syms = subprocess.check_output(["nm", "-D", "--defined-only", lib_path]).decode().split("\n")
    hits = [l.split()[-1] for l in syms if "MOD_" + name in l]
    assert len(hits) == 1, (name, hits)
    f = getattr(ctypes.CDLL(lib_path), hits[0])
    f.restype = ctypes.c_double
    return f


@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_xy", "perio_xyz"])
def test_mgfor_driver_sequence_matches_device_solve(geometry):
    from nyles_amd import _lib as L
    from nyles_amd.device import DeviceModel
    print("case mgfor shim", geometry, flush=True)
    nx, ny, nz = 64, 32, 32
    lib = L.LIB_PATH
    get_mg = _function(lib, "get_ptrmg")
    get_mg.restype = ctypes.POINTER(ctypes.c_int64)
    print_mg = _function(lib, "print_mginfos")
    libsolve = _function(lib, "solve")
    get_pyshape = _function(lib, "get_pyshape")
    set_pyarray = _function(lib, "set_pyarray")
    get_pyarray = _function(lib, "get_pyarray")
    mg = get_mg(*_ptr((1, 1, nx, ny, nz, False, False, True, TOPOLOGY[geometry])))
    assert mg
    print_mg(mg)
    shape = np.zeros((3,), dtype="i")
    get_pyshape(mg, _ptr(1), _ptr(shape))
    assert tuple(shape) == (nz + 2 * NH, ny + 2 * NH, nx + 2 * NH)
    x = np.zeros(tuple(shape))
    b = np.zeros(tuple(shape))
    n1, n2, n3 = x.shape       # core/mgfordriver.py:87-90
    rng = np.random.default_rng(TOPOLOGY[geometry])
    I = (slice(NH, -NH),) * 3
    dev = DeviceModel(nx, ny, nz, PERIODIC[geometry], 1.0, 1.0, 1.0)
    for rep in range(2):
        bi = rng.standard_normal((nz, ny, nx))
        bi -= bi.mean()
        b[I] = bi
        set_pyarray(mg, *_ptr((1, 2, n3, n2, n1, b)))
        libsolve(mg)
        get_pyarray(mg, *_ptr((1, 1, n3, n2, n1, x)))
        dev.upload(L.NYL_F_MGB, bi)
        st = dev.mg_solve()
        want = dev.download(L.NYL_F_MGX)
        assert np.array_equal(x[I], want), (rep, st)
        # periodic images in the halo, zeros on wall axes
        pz, py, px = PERIODIC[geometry][::-1]
        for ax, p in ((2, px), (1, py), (0, pz)):
            lo = [slice(NH, -NH)] * 3
            lo[ax] = slice(0, NH)
            src = [slice(NH, -NH)] * 3
            n = x.shape[ax] - 2 * NH
            src[ax] = slice(n, n + NH)
            if p:
                assert np.array_equal(x[tuple(lo)], x[tuple(src)])
            else:
                assert np.all(x[tuple(lo)] == 0.0)
    dev.close()
