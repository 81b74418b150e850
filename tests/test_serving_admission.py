"""The server's admission policy (host logic, CPU): jobs that can never fit
are rejected instead of blocking the queue, smaller jobs may pass a waiting
one a bounded number of times, and slots and the budget are respected."""
from nyles_amd.serving import Admission, Job


def _jobs(sizes):
    js = [Job("tgv", n, 10, arrival=0.01 * q) for q, n in enumerate(sizes)]
    for q, j in enumerate(js):
        j.id = q
    return js


def _need(j):
    return j.n        # bytes = n, for the test


def test_job_above_budget_is_rejected_not_blocking():
    pol = Admission(budget=100, max_slots=8)
    js = _jobs([500, 40, 30])
    admit, reject = pol.select(js, n_active=0, used=0, free=10 ** 9, need=_need)
    assert [j.id for j in reject] == [0] and "budget" in js[0].failed
    assert [j.id for j in admit] == [1, 2]


def test_job_above_free_memory_with_nothing_resident_is_rejected():
    pol = Admission(budget=1000, max_slots=8)
    js = _jobs([300, 50])
    admit, reject = pol.select(js, n_active=0, used=0, free=200, need=_need)
    assert [j.id for j in reject] == [0] and "free device memory" in js[0].failed
    assert [j.id for j in admit] == [1]


def test_job_above_free_memory_waits_while_jobs_are_resident():
    pol = Admission(budget=1000, max_slots=8)
    js = _jobs([300])
    admit, reject = pol.select(js, n_active=2, used=500, free=200, need=_need)
    assert admit == [] and reject == []


def test_budget_and_slots_are_respected():
    pol = Admission(budget=100, max_slots=3)
    js = _jobs([40, 40, 40, 10])
    admit, reject = pol.select(js, n_active=0, used=0, free=10 ** 9, need=_need)
    assert [j.id for j in admit] == [0, 1] or [j.id for j in admit] == [0, 1, 3]
    used = sum(_need(j) for j in admit)
    assert used <= 100 and len(admit) <= 3 and reject == []
    admit, _ = Admission(budget=1000, max_slots=2).select(_jobs([1, 1, 1]), 1, 0, 10 ** 9, _need)
    assert len(admit) == 1


def test_bypass_is_bounded():
    """A large job that waits is passed over at most max_bypass times; then
    later jobs wait behind it until it fits."""
    pol = Admission(budget=100, max_slots=64, max_bypass=2)
    big = _jobs([90])[0]
    passes = []
    used = 50                                       # resident jobs hold 50
    for rnd in range(4):
        small = Job("tgv", 5, 10, arrival=1.0 + rnd)
        small.id = 100 + rnd
        admit, reject = pol.select([big, small], n_active=1, used=used, free=10 ** 9, need=_need)
        assert reject == []
        passes.append([j.id for j in admit])
        used += sum(_need(j) for j in admit)
    assert passes[0] == [100] and passes[1] == [101]
    assert big.bypassed == 2
    assert passes[2] == [] and passes[3] == []     # reserved for the big job
    admit, _ = pol.select([big], n_active=0, used=0, free=10 ** 9, need=_need)
    assert admit == [big]
