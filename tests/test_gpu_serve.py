"""The request loop: continuous batching with admission by memory budget.

Requests arrive while others run; they are admitted into new micro-batches
as memory frees up, run beside the earlier ones, and each result (time
series and final membrane state) must equal the same request served alone,
bit for bit.
"""
import numpy as np
import pytest

import ensemble as en
import irbc_amd as ia
import serve


def _requests():
    reps = [
        en.replica(16, 12, 10, 10, 162, 2.6, 1.0, 162, kon=200.0, koff=20.0, shear=0.1, seed=41),
        en.replica(12, 12, 8, 6, 42, 1.8, 1.2, 30, kon=5.0, koff=1.0, shear=0.05, seed=42),
        en.replica(20, 16, 12, 8, 162, 3.0, 1.3, 100, kon=0.5, koff=0.05, shear=0.02, seed=43),
        en.replica(16, 16, 10, 5, 162, 2.8, 1.1, 60, kon=2.0, koff=0.5, shear=0.08, seed=44),
        en.replica(14, 12, 10, 7, 42, 2.0, 1.0, 42, kon=50.0, koff=5.0, shear=0.03, seed=45),
    ]
    arrivals = [0.0, 0.001, 0.002, 0.02, 0.02]
    return [serve.Request(i, d, 400, 50, t) for i, (d, t) in enumerate(zip(reps, arrivals))]


@pytest.mark.gpu
def test_served_results_equal_alone_and_admission_by_budget():
    reqs = _requests()
    # room for about two requests at a time: the rest wait for memory
    budget = int(2.2 * max(r.bytes for r in reqs))
    srv = serve.Server(budget, chunk=100, max_group=2)
    srv.run(reqs)
    assert all(r.finished is not None for r in reqs)
    assert srv.used == 0
    assert srv.stats["groups"] >= 3
    assert srv.stats["max_active_groups"] >= 2  # micro-batches ran together
    lat = [r.latency for r in reqs]
    assert min(lat) > 0
    for r in reqs:
        assert len(r.series) == r.steps // r.series_every
        alone = serve.Server(r.bytes * 2, chunk=400)
        a = serve.Request(r.id, r.replica, r.steps, r.series_every, 0.0)
        alone.run([a])
        np.testing.assert_array_equal(r.series, a.series, err_msg="request %d series" % r.id)
        for k in r.final:
            np.testing.assert_array_equal(r.final[k], a.final[k], err_msg="request %d %s" % (r.id, k))
    assert reqs[0].series[-1, 7] > 0  # the fast-binding request holds bonds


@pytest.mark.gpu
def test_batching_window_groups_arrivals_and_keeps_results():
    """With a batching window the loop waits for arrivals and forms fewer,
    larger micro-batches (eight or more replicas switch to four particles
    per lane); every result still equals the request served alone."""
    reps = [r.replica for r in _requests()] * 2
    reqs = [serve.Request(i, d, 200, 50, 0.002 * i) for i, d in enumerate(reps)]
    budget = int(4 * sum(r.bytes for r in reqs))
    srv = serve.Server(budget, chunk=100, max_active=4, window=0.05)
    srv.run(reqs)
    assert all(r.finished is not None for r in reqs)
    assert srv.used == 0
    assert srv.stats["groups"] == 1  # every request arrived inside the window
    assert srv.stats["max_resident_replicas"] == len(reqs)
    # the oldest request waited for the window before it was admitted
    assert reqs[0].admitted >= 0.05
    for r in reqs[:5]:
        alone = serve.Server(r.bytes * 2, chunk=200)
        a = serve.Request(r.id, r.replica, r.steps, r.series_every, 0.0)
        alone.run([a])
        np.testing.assert_array_equal(r.series, a.series, err_msg="request %d series" % r.id)
        for k in r.final:
            np.testing.assert_array_equal(r.final[k], a.final[k], err_msg="request %d %s" % (r.id, k))


@pytest.mark.gpu
def test_request_larger_than_budget_is_refused():
    r = _requests()[0]
    srv = serve.Server(r.bytes // 2)
    with pytest.raises(ValueError):
        srv.run([r])
