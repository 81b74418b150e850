"""nyl_reduce_noalloc (GPU): the reductions of nyl_reduce on caller-owned
fields, with the partial sums in the library's static device array instead
of a context workspace.

The entry point runs the same kernels with the same number of partials as
nyl_reduce, so every result must be bit for bit the context's.  Two calls in
a row, the second on a smaller grid (fewer partials: it must not read the
first call's leftovers), in both precisions; plus the argument checks, which
return before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

WHATS = ("MAXU2", "KE", "BSUM", "NONFINITE")


def _model(n3, periodic, precision, seed):
    from nyles_amd import _lib as L
    from nyles_amd.device import DeviceModel
    nx, ny, nz = n3
    dev = DeviceModel(nx, ny, nz, periodic, 1.0 / nx, 1.0 / nx, 1.0 / nx, precision=precision)
    rng = np.random.default_rng(seed)
    for fid in (L.NYL_F_B, L.NYL_F_U, L.NYL_F_V, L.NYL_F_W):
        dev.upload(fid, rng.standard_normal((nz, ny, nx)) * 1e-3)
    return dev


def _noalloc(dev, what, out):
    from nyles_amd import _lib as L
    fields = [dev.field(f) for f in (L.NYL_F_B, L.NYL_F_U, L.NYL_F_V, L.NYL_F_W)]
    dtype = L.NYL_DT_F64 if fields[0].dtype == torch.float64 else L.NYL_DT_F32
    ptrs = [ctypes.c_void_p(t.data_ptr()) for t in fields]
    L.check(dev.lib.nyl_reduce_noalloc(ctypes.byref(dev.desc), dtype, *ptrs, what,
                                       ctypes.c_void_p(out.data_ptr()),
                                       ctypes.c_void_p(dev.stream.cuda_stream)))
    dev.stream.synchronize()
    return float(out.item())


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_noalloc_reduce_matches_nyl_reduce(precision):
    from nyles_amd import _lib as L
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda:0")
    # first a grid with many partials, then a smaller one with fewer
    cases = [((96, 64, 48), (1, 1, 0)), ((32, 16, 16), (1, 0, 0))]
    for call, (n3, periodic) in enumerate(cases):
        print("case noalloc", precision, n3, periodic, flush=True)
        dev = _model(n3, periodic, precision, seed=11 + call)
        if call == 1:
            # a non-finite value for the count
            with torch.cuda.stream(dev.stream):
                dev.field(L.NYL_F_V)[3, 5, 7] = float("inf")
            dev.stream.synchronize()
        for name in WHATS:
            what = getattr(L, "NYL_RED_" + name)
            want = dev.reduce(what)
            got = _noalloc(dev, what, out)
            assert got == want, (call, name, got, want)
        if call == 1:
            assert _noalloc(dev, L.NYL_RED_NONFINITE, out) == 1.0
        dev.close()


def test_noalloc_reduce_rejects_bad_arguments():
    from nyles_amd import _lib as L
    lib = L.load()
    out = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    d = L.NylDesc()
    d.nx, d.ny, d.nz = 2048, 2048, 1024          # 2^32 cells: beyond the 4 GiB field limit
    d.idx2 = d.idy2 = d.idz2 = 1.0
    p = ctypes.c_void_p(out.data_ptr())           # never dereferenced: rejected first
    rc = lib.nyl_reduce_noalloc(ctypes.byref(d), L.NYL_DT_F32, p, p, p, p, L.NYL_RED_BSUM, p,
                                None)
    assert rc == -4 and b"4 GiB" in lib.nyl_last_error()
    d.nx, d.ny, d.nz = 8, 8, 8
    rc = lib.nyl_reduce_noalloc(ctypes.byref(d), L.NYL_DT_F64, p, None, None, None,
                                L.NYL_RED_KE, p, None)
    assert rc == -1
    rc = lib.nyl_reduce_noalloc(ctypes.byref(d), 7, p, p, p, p, L.NYL_RED_KE, p, None)
    assert rc == -1
    rc = lib.nyl_reduce_noalloc(ctypes.byref(d), L.NYL_DT_F64, p, p, p, p, 9, p, None)
    assert rc == -1
    rc = lib.nyl_reduce_noalloc(ctypes.byref(d), L.NYL_DT_F64, p, p, p, p, L.NYL_RED_KE, None,
                                None)
    assert rc == -1
