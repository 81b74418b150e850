"""Continuous-batching request loop (GPU).

Each job's result is the same alone and in the mix: the served run (jobs
admitted and retired between batched steps, different companions at every
step) is compared bit for bit (tolerance 0) with the job run alone in a
batch of one.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs():
    from nyles_amd.serving import Job
    return [Job("lock", 32, 6, seed=1, arrival=0.0), Job("tgv", 32, 4, seed=2, arrival=0.0),
            Job("shear", 64, 5, seed=3, arrival=0.02), Job("tgv", 64, 3, seed=4, arrival=0.05),
            Job("lock", 64, 7, seed=5, arrival=0.1), Job("shear", 32, 2, seed=6, arrival=0.3)]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_results_same_alone_and_in_the_mix(precision):
    from nyles_amd import serving
    from nyles_amd.device import DeviceBatch
    print("case serving", precision, flush=True)
    jobs = _jobs()
    server = serving.Server(precision=precision, max_slots=4, keep_state=True, tile=256)
    rep = server.run(jobs)
    print("  report", rep, flush=True)
    assert rep["jobs"] == len(jobs)
    assert rep["p99_s"] >= rep["p50_s"] > 0
    assert rep["batch_rebuilds"] >= 3          # slots freed and refilled
    for j in jobs:
        m, dt = serving.make_model(j, precision)
        assert dt == j.dt
        b = DeviceBatch([m], tile=256)
        for n in range(j.steps):
            b.step_rk3([dt])
        alone = serving.final_result(m, True)
        b.close()
        m.close()
        for k, a in alone["state"].items():
            got = j.result["state"][k]
            assert np.array_equal(a, got), (j.id, j.kind, j.n, k)
        assert alone["ke"] == j.result["ke"]


def test_admission_waits_for_memory():
    """With an HBM budget that holds the largest job but not all of them, the
    resident jobs never exceed the budget, jobs wait for memory to free up,
    and all complete."""
    from nyles_amd import serving
    jobs = _jobs()
    needs = [serving.workspace_bytes(j, "mixed") for j in jobs]
    server = serving.Server(precision="mixed", tile=256)
    server.budget = max(needs) + min(needs) // 2
    rep = server.run(jobs)
    print("  report", rep, flush=True)
    assert rep["jobs"] == len(jobs)
    assert rep["peak_resident_bytes"] <= server.budget
    assert rep["mean_slots"] < len(jobs)
