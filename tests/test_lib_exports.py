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
    assert L.nyl_abi_version() == 3
    assert L.nyl_last_error() is not None


def test_descriptor_layout_matches_header():
    """nyl_desc: six int32, eight doubles, two int32, three doubles, four
    int32 (136 bytes)."""
    from nyles_amd import _lib
    assert ctypes.sizeof(_lib.NylDesc) == 136
    assert _lib.NylDesc.dx.offset == 24
    assert _lib.NylDesc.precision.offset == 88
    assert _lib.NylDesc.mg_tol.offset == 96
    assert _lib.NylDesc.viscosity.offset == 112
    assert _lib.NylDesc.n_tracers.offset == 120
    assert _lib.NylDesc.model.offset == 124
    assert _lib.NylDesc.upwind_vf.offset == 132
    assert ctypes.sizeof(_lib.NylMGStats) == 24


def _desc(n3, periodic=(1, 1, 0), precision=0):
    from nyles_amd import _lib
    d = _lib.NylDesc()
    d.nx, d.ny, d.nz = n3
    d.periodic_x, d.periodic_y, d.periodic_z = periodic
    d.dx = d.dy = d.dz = 1.0
    d.precision = precision
    d.mg_maxite, d.mg_tol, d.mg_omega = 20, 1e-6, 0.9
    return d


def _levels(L, ctx):
    n = ctypes.c_int32()
    assert L.nyl_mg_levels(ctx, ctypes.byref(n)) == 0
    out = []
    for l in range(n.value):
        a, b, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        assert L.nyl_mg_level_shape(ctx, l, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) == 0
        out.append((a.value, b.value, c.value))
    return out


def test_context_rejects_bad_sizes_without_gpu():
    """Argument checks happen before any HIP call."""
    from nyles_amd import _lib
    L = _lib.load()
    ctx = ctypes.c_void_p()
    d = _desc((4, 4, 4))
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -4     # fewer than 8 cells
    d = _desc((40, 40, 41))                                                # no halving possible
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -4
    assert b"at least one halving" in L.nyl_last_error()
    d = _desc((1024, 1024, 24))            # coarsest 128x128x3: odd, too wide to solve directly
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == -4
    assert b"direct coarse solve" in L.nyl_last_error()
    d = _desc((32, 32, 32))
    rc = L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx))
    assert rc == 0
    nb = ctypes.c_uint64()
    assert L.nyl_workspace_bytes(ctx, ctypes.byref(nb)) == 0 and nb.value > 0
    # no workspace bound: compute entry points refuse
    assert L.nyl_rhs(ctx, 0, 0) == -2
    assert L.nyl_ctx_destroy(ctx) == 0


def test_hierarchy_of_grids_that_do_not_halve_to_2():
    """The reference's hierarchy (halve x, y and z until x or y is 2,
    core/mgfor/mg_setup.f90:262-303) where every level halves evenly; a grid
    whose halving reaches an odd size stops there (the level is solved
    directly), e.g. the ensemble sizes 40..96 in steps of 8 and the served
    sizes up to 256, in both precisions."""
    from nyles_amd import _lib
    L = _lib.load()
    ctx = ctypes.c_void_p()
    want = {64: [(64, 64, 64), (32, 32, 32), (16, 16, 16), (8, 8, 8), (4, 4, 4), (2, 2, 2)],
            48: [(48, 48, 48), (24, 24, 24), (12, 12, 12), (6, 6, 6), (3, 3, 3)],
            40: [(40, 40, 40), (20, 20, 20), (10, 10, 10), (5, 5, 5)],
            88: [(88, 88, 88), (44, 44, 44), (22, 22, 22), (11, 11, 11)],
            248: [(248, 248, 248), (124, 124, 124), (62, 62, 62), (31, 31, 31)]}
    for n, lev in want.items():
        for prec in (0, 1):
            d = _desc((n, n, n), precision=prec)
            assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == 0, L.nyl_last_error()
            assert _levels(L, ctx) == lev
            assert L.nyl_ctx_destroy(ctx) == 0
    for n in range(32, 257, 8):
        for prec in (0, 1):
            d = _desc((n, n, n), precision=prec)
            assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == 0, (n, L.nyl_last_error())
            assert L.nyl_ctx_destroy(ctx) == 0
    # every served shape: the lock exchange n x n/2 x n/2 (perio_x) as well,
    # whose coarsest odd level is 2m x m x m (x up to 62 cells)
    for n in range(32, 257, 8):
        for prec in (0, 1):
            d = _desc((n, n // 2, n // 2), periodic=(1, 0, 0), precision=prec)
            assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == 0, (n, L.nyl_last_error())
            assert L.nyl_ctx_destroy(ctx) == 0
    d = _desc((248, 124, 124), periodic=(1, 0, 0))
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == 0
    assert _levels(L, ctx)[-1] == (62, 31, 31)
    assert L.nyl_ctx_destroy(ctx) == 0
    # the lock-exchange shape keeps the reference's hierarchy down to y = 2
    d = _desc((64, 32, 32), periodic=(1, 0, 0))
    assert L.nyl_ctx_create(ctypes.byref(d), ctypes.byref(ctx)) == 0
    assert _levels(L, ctx) == [(64, 32, 32), (32, 16, 16), (16, 8, 8), (8, 4, 4), (4, 2, 2)]
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


def test_mgfor_shim_symbols_are_found_as_the_reference_finds_them():
    """core/build.py:107-121 looks each mgfor function up by the substring
    "MOD_" + name in the library's symbols and needs exactly one match."""
    import shutil
    import subprocess
    from nyles_amd import _lib
    nm = shutil.which("nm")
    if nm is None:
        import pytest
        pytest.skip("nm is not available")
    syms = subprocess.check_output([nm, "-D", "--defined-only", _lib.LIB_PATH]).decode()
    names = [l.split()[-1] for l in syms.split("\n") if l.strip()]
    for f in ("get_ptrmg", "print_mginfos", "solve", "get_pyshape", "set_pyarray",
              "get_pyarray", "get_param"):
        hits = [n for n in names if "MOD_" + f in n]
        assert len(hits) == 1, (f, hits)



def test_stage_tile_candidates_and_names():
    """The batch autotuner's tile list (a host-side table): 15 distinct
    shapes of 128, 256 and 512 cells, rows or bricks that multiply out;
    names round-trip; an index past the list is refused."""
    from nyles_amd import _lib
    from nyles_amd.device import parse_tile, tile_candidates, tile_name
    c = tile_candidates()
    assert len(c) == _lib.NYL_TILE_CANDIDATES == 15 and len(set(c)) == 15
    for cells, tx, ty in c:
        assert cells in (128, 256, 512)
        assert (tx, ty) == (0, 0) or (cells % (tx * ty) == 0 and tx * ty <= cells)
        assert parse_tile(tile_name((cells, tx, ty))) == (cells, tx, ty)
    assert sum(1 for t in c if t[1] == 0) == 3
    assert parse_tile(256) == (256, 0, 0) and tile_name((512, 32, 4)) == "32x4x4"
    L = _lib.load()
    t = (ctypes.c_int32 * 3)()
    assert L.nyl_batch_tile_candidate(15, t) == _lib.NYL_E_ARG
    assert L.nyl_batch_tile_candidate(-1, t) == _lib.NYL_E_ARG
