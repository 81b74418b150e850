"""Host-side API of the reference (parameters, topology, variables, grid) on
the CPU: the pieces an experiment script touches before the GPU runs."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CORE = os.path.join(ROOT, "nyles_amd", "core")
if CORE not in sys.path:
    sys.path.insert(0, CORE)


@pytest.fixture
def param_mod():
    import parameters
    parameters.InextensibleDict.frozen = False
    yield parameters
    parameters.InextensibleDict.frozen = False


def test_defaults_pass_the_check(param_mod):
    p = param_mod.UserParameters()
    p.check()
    assert p.time["timestepping"] == "LFAM3"
    assert p.discretization["global_nx"] == 64
    assert p.device["mg_tol"] == 1e-6 and p.device["mg_maxite"] == 20


def test_no_new_keys_and_freeze(param_mod):
    p = param_mod.UserParameters()
    p.IO["expname"] = "2nd_experiment"
    with pytest.raises(param_mod.UserParameterError):
        p.IO["epxname"] = "typo"
    with pytest.raises(param_mod.UserParameterError):
        p.model["expname"] = "wrong category"
    p.freeze()
    with pytest.raises(param_mod.UserParameterError):
        p.IO["expname"] = "after freeze"


@pytest.mark.parametrize("cat,key,bad,good", [
    ("model", "geometry", "open", "closed"),
    ("model", "Lx", -3, 3),
    ("model", "Ly", "5", 5),
    ("IO", "expname", "foo/bar", "foo-bar"),
    ("IO", "variables_in_history", "any", "all"),
    ("IO", "timestep_history", -1.0, 0.0),
    ("time", "cfl", 0, 1),
    ("discretization", "global_nx", 5, 3),
    ("discretization", "global_ny", 2 ** 1000, 64),
    ("MPI", "npx", 3, 1),
    ("MPI", "nh", -1, 3),
    ("animation", "iterations_per_frame", 0, 1),
    ("device", "precision", "fp16", "fp64"),
])
def test_check_rejects_bad_values(param_mod, cat, key, bad, good):
    p = param_mod.UserParameters()
    getattr(p, cat)[key] = bad
    with pytest.raises(param_mod.UserParameterError):
        p.check()
    getattr(p, cat)[key] = good
    p.check()


def test_more_cores_than_cells_is_rejected(param_mod):
    p = param_mod.UserParameters()
    p.discretization["global_nz"] = 2
    p.MPI["npz"] = 4
    with pytest.raises(param_mod.UserParameterError):
        p.check()


@pytest.mark.parametrize("geometry,expect", [
    ("closed", 0), ("perio_x", 2), ("perio_y", 2), ("perio_xy", 8), ("perio_xyz", 26)])
def test_one_rank_neighbours(geometry, expect):
    import topology as topo
    topo.topology = geometry
    ngs = topo.get_neighbours([0, 0, 0], [1, 1, 1])
    assert len(ngs) == expect
    assert all(v == 0 for v in ngs.values())


def test_neighbour_graph_is_symmetric_on_many_ranks():
    import topology as topo
    procs = [2, 2, 4]
    for geometry in ("closed", "perio_xy", "perio_xyz"):
        topo.topology = geometry
        allng = [topo.get_neighbours(r, procs) for r in range(16)]
        topo.check_graph(allng)


def _param(geometry, n=(16, 8, 12), nh=3):
    import topology as topo
    topo.topology = geometry
    return {"nx": n[0], "ny": n[1], "nz": n[2], "nh": nh,
            "neighbours": topo.get_neighbours([0, 0, 0], [1, 1, 1]),
            "npx": 1, "npy": 1, "npz": 1, "Lx": 2.0, "Ly": 1.0, "Lz": 1.5, "loc": [0, 0, 0]}


@pytest.mark.parametrize("geometry", ["closed", "perio_x", "perio_xy", "perio_xyz"])
def test_scalar_shapes_and_views(geometry):
    import variables as var
    param = _param(geometry)
    s = var.get_state(param)
    perio = {"closed": (0, 0, 0), "perio_x": (0, 0, 1), "perio_xy": (0, 1, 1),
             "perio_xyz": (1, 1, 1)}[geometry]
    shape = tuple(n + 6 * p for n, p in zip((12, 8, 16), perio))
    b = s.b.view("i")
    assert b.shape == shape
    rng = np.random.default_rng(0)
    b[...] = rng.standard_normal(shape)
    ref = b.copy()
    bj = s.b.view("j")
    assert bj.shape == (shape[2], shape[0], shape[1])
    np.testing.assert_array_equal(bj, np.transpose(ref, (2, 0, 1)))
    bk = s.b.view("k")
    np.testing.assert_array_equal(bk, np.transpose(ref, (1, 2, 0)))
    np.testing.assert_array_equal(s.b.view("i"), ref)
    assert s.b.flipview("k") is s.b.view("i")
    assert s.get_prognostic_scalars() == ["b", "u_i", "u_j", "u_k"]


def test_fill_halo_gives_periodic_images():
    import variables as var
    param = _param("perio_xyz", n=(8, 8, 8))
    s = var.get_state(param)
    a = s.b.view("i")
    rng = np.random.default_rng(1)
    inner = rng.standard_normal((8, 8, 8))
    s.b.interior(a)[...] = inner
    var.fill_halo(a, param["neighbours"], 3, s.b.domainindices)
    idx = (np.arange(14) - 3) % 8
    np.testing.assert_array_equal(a, inner[np.ix_(idx, idx, idx)])


def test_grid_coordinates_and_metric():
    import grid as g
    param = _param("perio_x")
    gr = g.Grid(param)
    assert gr.dx == 2.0 / 16 and gr.idx2 == 1 / gr.dx ** 2
    x = gr.x_b.view("i")
    i0 = gr.domainindices[4]
    np.testing.assert_allclose(x[0, 0, i0:i0 + 16], (np.arange(16) + 0.5) * gr.dx)
    xu = gr.x_vel["i"].view("i")
    np.testing.assert_allclose(xu - x, gr.dx / 2)
    zw = gr.z_vel["k"].view("i")
    np.testing.assert_allclose(zw - gr.z_b.view("i"), gr.dz / 2)


def test_history_frames_hold_the_state_on_schedule(tmp_path):
    """nylesIO: a frame at init, then one each time t reaches the next
    multiple of timestep_history (core/nylesIO.py's schedule), and one at
    finalize unless the last step was just written; every frame holds t, n
    and the chosen variables (vectors as name_i/_j/_k) equal to the state
    when it was written, interior cells or full arrays with include_halo."""
    import nylesIO
    import variables as var
    for halo in (False, True):
        param = _param("perio_x", n=(6, 4, 8))
        s = var.get_state(param)
        io_param = {"disk_space_warning": 0, "simplified_grid": True,
                    "variables_in_history": ["b", "u"], "timestep_history": 0.125,
                    "include_halo": halo, "myrank": 0, "n_tracers": 0,
                    "datadir": str(tmp_path), "expname": "h%d" % halo, "mode": "overwrite"}
        io = nylesIO.NylesIO(io_param)
        rng = np.random.default_rng(3)

        def perturb():
            for name in ("b", "u_i", "u_j", "u_k"):
                obj = s.b if name == "b" else s.u[name[-1]]
                obj.view("i")[...] = rng.standard_normal(obj.view("i").shape)

        def snap():
            out = {"b": s.b.view("i").copy()}
            for d in "ijk":
                out["u_" + d] = s.u[d].view("i").copy()
            return out

        perturb()
        want = [(0.0, 0, snap())]
        io.init(s, None, t=0.0, n=0)
        # dt = 1/16, history every 1/8 (exact binary): frames after steps 2, 4, 6
        for n in range(1, 8):
            t = n / 16
            perturb()
            io.write(s, t, n)
            if n in (2, 4, 6):
                want.append((t, n, snap()))
        perturb()
        io.finalize(s, 0.5, 8)
        want.append((0.5, 8, snap()))
        files = sorted(f for f in os.listdir(io.output_directory) if f.endswith(".npz"))
        assert len(files) == len(want) == 5
        for f, (tw, nw, st) in zip(files, want):
            with np.load(os.path.join(io.output_directory, f)) as z:
                assert float(z["t"]) == tw and int(z["n"]) == nw
                assert sorted(z.files) == ["b", "n", "t", "u_i", "u_j", "u_k"]
                for key, a in st.items():
                    obj = s.b if key == "b" else s.u[key[-1]]
                    np.testing.assert_array_equal(z[key], a if halo else obj.interior(a))
