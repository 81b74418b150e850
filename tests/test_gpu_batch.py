"""Ensemble batches (nyl_batch_*): many members of uneven sizes and
geometries advanced by one set of launches per RK stage (GPU).

Tolerances: FP64 members reproduce the single-member FP64 path bit for bit
(same cell arithmetic; the multigrid norms are summed in another order, so a
V-cycle count could only differ if a residual fell within rounding of the
tolerance).  A member's result is bitwise the same alone and in any mix,
with or without hipGraphs, whatever tile the autotuner picks.  Mixed-
precision members (one FP32 V-cycle per FP64 outer residual in a batch) are
within 1e-5 relative (L2) of the single-member mixed path.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (nx, ny, nz, geometry, L)
MEMBERS = [(32, 32, 32, "perio_xy", (1.0, 1.0, 1.0)),
           (64, 32, 32, "perio_x", (16.0, 8.0, 8.0)),
           (32, 32, 16, "closed", (2.0, 2.0, 1.0)),
           (64, 64, 32, "perio_xyz", (2.0, 2.0, 1.0))]


@pytest.fixture(scope="module")
def cmp():
    import compare
    return compare


def _member(cmp, spec, seed, precision=0):
    nx, ny, nz, geo, L = spec
    orc, dev = cmp.make_pair(nx, ny, nz, geo, Lx=L[0], Ly=L[1], Lz=L[2], precision=precision)
    st = cmp.random_state(orc, seed=seed, amp_u=3e-3)
    cmp.load_state(orc, dev, st)
    dev.diagnose()
    return orc, dev


def _state(dev):
    from nyles_amd import _lib as L
    return {k: dev.download(f) for k, f in (("b", L.NYL_F_B), ("u", L.NYL_F_U),
                                             ("v", L.NYL_F_V), ("w", L.NYL_F_W),
                                             ("ke", L.NYL_F_KE), ("vor_k", L.NYL_F_VORZ))}


def _same(cmp, a, b, what):
    for k in a:
        bad = cmp.mismatches(a[k], b[k])
        assert bad == 0, "%s %s: %d elements differ" % (what, k, bad)


def _dt(dev):
    from nyles_amd import _lib as L
    return min(0.5 / np.sqrt(dev.reduce(L.NYL_RED_MAXU2)), 0.1)


def test_fp64_batch_matches_single_members(cmp):
    from nyles_amd.device import DeviceBatch
    print("case batch fp64 vs single", flush=True)
    singles = [_member(cmp, s, 20 + q)[1] for q, s in enumerate(MEMBERS)]
    batched = [_member(cmp, s, 20 + q)[1] for q, s in enumerate(MEMBERS)]
    dts = [_dt(d) for d in singles]
    batch = DeviceBatch(batched)
    print("  tile", batch.tile, batch.tune_ms, flush=True)
    for n in range(3):
        nites = [d.step_rk3(dt)["nite"] for d, dt in zip(singles, dts)]
        st = batch.step_rk3(dts)
        assert [s["nite"] for s in st] == nites, (st, nites)
        for q in range(len(MEMBERS)):
            _same(cmp, _state(batched[q]), _state(singles[q]), "step %d member %d" % (n, q))
    batch.close()
    for d in singles + batched:
        d.close()


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed"])
def test_member_same_alone_and_in_mix(cmp, precision):
    from nyles_amd.device import DeviceBatch
    print("case batch alone vs mix", precision, flush=True)
    _, alone = _member(cmp, MEMBERS[1], 5, precision)
    mix = [_member(cmp, s, 30 + q, precision)[1] for q, s in enumerate(MEMBERS)]
    _, inmix = _member(cmp, MEMBERS[1], 5, precision)
    mix[1] = inmix
    dt = _dt(alone)
    b1 = DeviceBatch([alone])
    bn = DeviceBatch(mix)
    dts = [_dt(d) for d in mix]
    dts[1] = dt
    for n in range(3):
        s1 = b1.step_rk3([dt])
        sn = bn.step_rk3(dts)
        assert s1[0]["nite"] == sn[1]["nite"]
        _same(cmp, _state(alone), _state(inmix), "step %d" % n)
    b1.close()
    bn.close()


def test_graph_replays_match_eager_launches(cmp):
    from nyles_amd.device import DeviceBatch
    print("case batch graphs", flush=True)
    ga = [_member(cmp, s, 40 + q, 1)[1] for q, s in enumerate(MEMBERS)]
    ea = [_member(cmp, s, 40 + q, 1)[1] for q, s in enumerate(MEMBERS)]
    dts = [_dt(d) for d in ga]
    bg = DeviceBatch(ga, graphs=True, autotune=False)
    be = DeviceBatch(ea, graphs=False, autotune=True)
    for n in range(4):
        sg = bg.step_rk3(dts)
        se = be.step_rk3(dts)
        assert [s["nite"] for s in sg] == [s["nite"] for s in se]
        for q in range(len(MEMBERS)):
            _same(cmp, _state(ga[q]), _state(ea[q]), "step %d member %d" % (n, q))
    bg.close()
    be.close()


def test_mixed_batch_close_to_single_mixed_path(cmp):
    from nyles_amd.device import DeviceBatch
    print("case batch mixed vs single", flush=True)
    singles = [_member(cmp, s, 50 + q, 1)[1] for q, s in enumerate(MEMBERS)]
    batched = [_member(cmp, s, 50 + q, 1)[1] for q, s in enumerate(MEMBERS)]
    dts = [_dt(d) for d in singles]
    batch = DeviceBatch(batched)
    for n in range(3):
        for d, dt in zip(singles, dts):
            d.step_rk3(dt)
        batch.step_rk3(dts)
    for q in range(len(MEMBERS)):
        a, b = _state(batched[q]), _state(singles[q])
        for k in ("b", "u", "v", "w"):
            e = cmp.rel_l2(a[k], b[k])
            print("  member %d %s rel %.2e" % (q, k, e), flush=True)
            assert e < 1e-5, (q, k, e)
    batch.close()
    for d in singles + batched:
        d.close()


def test_member_used_alone_between_batch_steps(cmp):
    """A member stepped on its own between batch steps (its state buffers
    swap) is picked up correctly by the next batch step."""
    from nyles_amd.device import DeviceBatch
    print("case batch interleaved", flush=True)
    ref = [_member(cmp, s, 60 + q)[1] for q, s in enumerate(MEMBERS[:2])]
    mem = [_member(cmp, s, 60 + q)[1] for q, s in enumerate(MEMBERS[:2])]
    dts = [_dt(d) for d in ref]
    batch = DeviceBatch(mem)
    for d, dt in zip(ref, dts):
        d.step_rk3(dt)
    batch.step_rk3(dts)
    for d, dt in zip(ref, dts):
        d.step_rk3(dt)
    mem[0].step_rk3(dts[0])     # alone
    mem[1].step_rk3(dts[1])
    for d, dt in zip(ref, dts):
        d.step_rk3(dt)
    batch.step_rk3(dts)
    for q in range(2):
        _same(cmp, _state(mem[q]), _state(ref[q]), "member %d" % q)
    batch.close()


def test_every_stage_tile_gives_the_same_bits(cmp):
    """Row tiles and bricks of every candidate shape (members of 16 to 64
    cells per axis, so some bricks overhang a member's edge) step the batch
    to the same bits as the single-member path; the V-cycle cap learnt from
    the previous step (V-cycles launched without a host read) changes
    nothing either."""
    from nyles_amd.device import DeviceBatch, tile_candidates, tile_name
    print("case batch tile shapes", flush=True)
    cands = tile_candidates()
    assert len(cands) == 15 and len(set(cands)) == 15
    for cells, tx, ty in cands:
        assert cells in (128, 256, 512)
        if tx:
            assert cells % (tx * ty) == 0
    singles = [_member(cmp, s, 60 + q)[1] for q, s in enumerate(MEMBERS)]
    batched = [_member(cmp, s, 60 + q)[1] for q, s in enumerate(MEMBERS)]
    dts = [_dt(d) for d in singles]
    batch = DeviceBatch(batched, autotune=False)
    for n, t in enumerate(cands):
        print("  tile", tile_name(t), flush=True)
        batch.set_tile(t)
        assert batch.tile == t
        nites = [d.step_rk3(dt)["nite"] for d, dt in zip(singles, dts)]
        st = batch.step_rk3(dts)
        assert [s["nite"] for s in st] == nites, (tile_name(t), st, nites)
        for q in range(len(MEMBERS)):
            _same(cmp, _state(batched[q]), _state(singles[q]),
                  "tile %s member %d" % (tile_name(t), q))
    batch.close()
    for d in singles + batched:
        d.close()
