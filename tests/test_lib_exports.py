"""libnyles_hip.so loads and exports every entry point include/nyles_hip.h
declares; no GPU call is made (CPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "nyles_hip.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(nyl_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding_list():
    from nyles_amd import _lib
    assert declared_functions() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    from nyles_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_binding_sets_signatures_and_abi_version():
    from nyles_amd import _lib
    L = _lib.load()
    assert L.nyl_abi_version() == 1
    assert L.nyl_last_error() is not None


def test_descriptor_layout_matches_header():
    """nyl_desc: six int32, eight doubles, two int32, two doubles (112 bytes)."""
    from nyles_amd import _lib
    assert ctypes.sizeof(_lib.NylDesc) == 112
    assert _lib.NylDesc.dx.offset == 24
    assert _lib.NylDesc.precision.offset == 88
    assert _lib.NylDesc.mg_tol.offset == 96
    assert ctypes.sizeof(_lib.NylMGStats) == 24


def test_context_rejects_bad_sizes_without_gpu():
    """Argument checks happen before any HIP call."""
    from nyles_amd import _lib
    L = _lib.load()
    d = _lib.NylDesc()
    d.nx, d.ny, d.nz = 48, 48, 48          # never halves down to 2
    d.dx = d.dy = d.dz = 1.0
    d.mg_maxite, d.mg_tol, d.mg_omega = 20, 1e-6, 0.9
    ctx = ctypes.c_void_p()
    rc = L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx))
    assert rc == -4 and b"halve" in L.nyl_last_error()
    d.nx = d.ny = d.nz = 4
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -4
    d.nx = d.ny = d.nz = 32
    rc = L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx))
    assert rc == 0
    nb = ctypes.c_uint64()
    assert L.nyl_workspace_bytes(ctx, ctypes.byref(nb)) == 0 and nb.value > 0
    # no workspace bound: compute entry points refuse
    assert L.nyl_rhs(ctx, 0, 0) == -2
    assert L.nyl_ctx_destroy(ctx) == 0


def test_context_rejects_fields_of_4gib_without_gpu():
    """The stage kernel addresses a field with 32-bit byte offsets: a grid
    whose field reaches 4 GiB is refused (2^29 cells in FP64, 2^30 in mixed
    precision), smaller grids are accepted."""
    from nyles_amd import _lib
    L = _lib.load()
    d = _lib.NylDesc()
    d.dx = d.dy = d.dz = 1.0
    d.mg_maxite, d.mg_tol, d.mg_omega = 20, 1e-6, 0.9
    ctx = ctypes.c_void_p()
    d.nx, d.ny, d.nz = 1024, 1024, 512        # 2^29 cells
    d.precision = _lib.NYL_PREC_FP64
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -4
    assert b"4 GiB" in L.nyl_last_error()
    d.precision = _lib.NYL_PREC_MIXED          # 2 GiB per FP32 field
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == 0
    assert L.nyl_ctx_destroy(ctx) == 0
    d.nx, d.ny, d.nz = 1024, 1024, 1024        # 2^30 cells: 4 GiB in FP32
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -4
    d.nx, d.ny, d.nz = 512, 512, 1024          # 2^28 cells: 2 GiB in FP64
    d.precision = _lib.NYL_PREC_FP64
    rc = L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx))
    assert rc == 0, L.nyl_last_error()
    assert L.nyl_ctx_destroy(ctx) == 0
