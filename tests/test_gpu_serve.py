"""The request loop: continuous batching with admission by memory budget.

Requests arrive while others run.  Re-packing (the default schedule) moves
the running requests and the newly admitted ones into one context at each
admission or once finished requests hold enough of it; micro-batches
(repack=False) run the requests admitted together as a context of their
own beside the earlier ones.  Either way each result (time series and
final membrane state) must equal the same request run alone, bit for bit.
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
    srv = serve.Server(budget, chunk=100, max_group=2, repack=False)
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


def _direct(r):
    """The request run as a plain one-replica context, no loop and no copy."""
    b = ia.Batch([r.spec])
    try:
        b.set_option(ia.OPT_ALIGN_LISTS, 1)
        en.load(b, 0, r.replica, device_solvent=True)
        en.start(b)
        b.set_series(0, r.series_every, max(1, r.steps // r.series_every))
        b.md_steps(r.steps)
        b.sync()
        v = b.vertices(0)
        return b.series(0), {k: v[k] for k in ("x", "y", "z", "vx", "vy", "vz", "ix", "iy", "bondi")}
    finally:
        b.close()


@pytest.mark.gpu
def test_repacked_requests_of_different_lengths_equal_direct_runs():
    """Re-packing: requests of 200-500 steps arriving within 5 ms (while
    the first context is still being built) share one running context,
    which is rebuilt as they arrive and as finished ones free a fifth of
    it; every running request is copied device to device at each switch.
    Each result equals the request run directly in a context of its own."""
    base = _requests()
    steps = [500, 300, 400, 200, 500, 300]
    arrivals = [0.0, 0.001, 0.002, 0.003, 0.004, 0.005]
    reps = [r.replica for r in base] + [dict(base[0].replica, seed=46)]
    reqs = [serve.Request(i, d, n, 50, t) for i, (d, n, t) in enumerate(zip(reps, steps, arrivals))]
    budget = int(8 * sum(r.bytes for r in reqs))
    srv = serve.Server(budget, chunk=100, window=0.0, repack_idle=0.2)
    srv.run(reqs)
    assert all(r.finished is not None for r in reqs)
    assert srv.used == 0
    assert srv.stats["groups"] >= 3  # several contexts, each taking over from the last
    assert srv.stats["max_resident_replicas"] >= 3
    for r in reqs:
        assert len(r.series) == r.steps // r.series_every
        series, final = _direct(r)
        np.testing.assert_array_equal(r.series, series, err_msg="request %d series" % r.id)
        for k in final:
            np.testing.assert_array_equal(r.final[k], final[k], err_msg="request %d %s" % (r.id, k))


@pytest.mark.gpu
def test_repacking_refuses_a_request_it_cannot_stage():
    r = _requests()[0]
    srv = serve.Server(int(1.5 * r.bytes))
    with pytest.raises(ValueError):
        srv.run([r])
