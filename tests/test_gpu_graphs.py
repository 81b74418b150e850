"""The coarse multigrid levels replayed from a hipGraph give the same bits as
direct launches (GPU).  Each step replays the graph several times (once per
V-cycle) with direct launches and host synchronisations in between; every
step is compared."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
@pytest.mark.parametrize("geometry", ["closed", "perio_xy", "perio_xyz"])
def test_graph_replay_matches_direct_launches(precision, geometry):
    import compare as cmp
    from nyles_amd import _lib as L
    print("case graphs", geometry, precision, flush=True)
    orc, dg = cmp.make_pair(64, 64, 32, geometry, Lx=4.0, Ly=4.0, Lz=2.0, precision=precision)
    _, dd = cmp.make_pair(64, 64, 32, geometry, Lx=4.0, Ly=4.0, Lz=2.0, precision=precision)
    dd.set_graphs(False)
    st = cmp.random_state(orc, seed=17, amp_u=3e-3)
    for d in (dg, dd):
        for k, f in (("b", L.NYL_F_B), ("u_i", L.NYL_F_U), ("u_j", L.NYL_F_V), ("u_k", L.NYL_F_W)):
            d.upload(f, st[k])
    sg, sd = dg.diagnose(), dd.diagnose()
    assert sg["nite"] == sd["nite"]
    dt = 0.02
    for n in range(4):
        sg, sd = dg.step_rk3(dt), dd.step_rk3(dt)
        assert sg["nite"] == sd["nite"] and sg["nite"] > 0
        for f in (L.NYL_F_B, L.NYL_F_U, L.NYL_F_V, L.NYL_F_W, L.NYL_F_MGX):
            a, b = dg.download(f), dd.download(f)
            assert cmp.mismatches(a, b) == 0, (n, f)
    dg.close()
    dd.close()
