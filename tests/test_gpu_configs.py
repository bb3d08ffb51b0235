"""Every benchmark configuration through the HIP path, at its own size.

Free-running against the oracle (no state copied after the start, same
counter-based streams, the oracle storing the solvent in FP32 where the
device does; see freerun.py):
  - the README run: 60x40x40 at density 10 (960 000 particles), the
    reference's Troph600 mesh and 500 knobs, its initial state as the
    reference builds it (reproduced bit for bit by the oracle against the
    reference's own fixtures), 90 steps: 9 collisions and a list update;
    the same with fast bond kinetics (kon = koff = 40), in three calls;
    the time series of both runs (centroid, centroid velocity, bonded
    knobs) against the same quantities of the oracle's state;
  - the single large replica: 64x64x48 (1 966 080 particles), 642-vertex
    icosphere, 30 steps with a collision and a list update;
  - a 2562-vertex stretched membrane in a 48x48x40 box at density 8;
  - sampled replicas inside the 64-replica uniform batch and the 48-replica
    uneven batch, advanced as part of their batch.
Batch and serving invariance: replicas of those batches and requests of a
Poisson stream drawn like the uneven batch give the same bits alone.

Tolerances, from the FP32 storage of the solvent (the oracle keeps FP64
arithmetic, exact cell sums are compared with sequential FP64 sums, and the
GPU's libm can differ from glibc in the last bit):
  - exact: shifted-cell keys of the last collision, per-cell counts, inside
    flags, step counters, periodic images, bonds, ligand occupancy;
  - solvent: at most 1% of the particles differ, each by at most 8 FP32
    units at the box scale (position) or 1e-5 (velocity);
  - membrane: positions within 1e-6 cells, velocities and forces within
    1e-5 of their largest magnitude (the cell-momentum precision target);
    area and volume within 1e-8 relative (the area/volume target).
"""
import numpy as np
import pytest

import ensemble as en
import freerun as fr
import irbc_amd as ia
import pyoracle as po
import serve

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _threads():
    po.set_threads(16)


def _readme(seed, kon, koff, n, chunks, every=10):
    o, V, E, T = fr.readme_state(seed, kon, koff)
    o.set_fp32(True)
    o.use_philox(seed)
    b = fr.device_from_oracle(o, V, E, T, seed)
    want = []
    try:
        b.set_diagnostics(True)
        b.set_series(0, every, n // every)
        fr.run(b, o, n, chunks, each=lambda k: want.append(fr.series_record(o)) if k % every == 0 else None)
        m = fr.compare(b, 0, o, cells=True)
        got = b.series(0)
    finally:
        b.close()
    print("README seed %d kon %g koff %g, %d steps:" % (seed, kon, koff, n), m)
    fr.check(m, (o.p.Lx, o.p.Ly, o.p.Lz), "README")
    # the time series (centroid, centroid velocity, bonded knobs every
    # `every` steps) against the same quantities of the oracle's state
    want = np.array(want)
    assert got.shape == want.shape, (got.shape, want.shape)
    np.testing.assert_array_equal(got[:, 0], want[:, 0])
    np.testing.assert_array_equal(got[:, 7], want[:, 7])
    np.testing.assert_allclose(got[:, 1:4], want[:, 1:4], rtol=0, atol=1e-6)
    vscale = np.abs(o.vertices()["vx"]).max()
    np.testing.assert_allclose(got[:, 4:7], want[:, 4:7], rtol=0, atol=1e-5 * vscale)
    return m


def test_readme_run_free_against_oracle():
    m = _readme(12345, 0.005, 0.015, 90, (None,))
    assert m["n"] == 960000


def test_readme_run_fast_bonds_in_three_calls():
    m = _readme(777, 40.0, 40.0, 40, (7, 13, None))
    assert m["bonded"] > 0  # bonds formed, and equal on both sides


def _single(d, n, upd_before):
    p = en.params(d)
    b = ia.Batch([en.spec(d)])
    try:
        st = en.load(b, 0, d)
        en.start(b)
        b.set_counters(0, 0, -1, p.uplist - upd_before)  # a list update inside the run
        o = fr.oracle_from_device(b, 0, p, st["V"], st["E"], st["T"])
        o.set_fp32(True)
        o.use_philox(d["seed"])
        b.set_diagnostics(True)
        fr.run(b, o, n)
        m = fr.compare(b, 0, o, cells=True)
    finally:
        b.close()
    print(d, m)
    fr.check(m, (d["Lx"], d["Ly"], d["Lz"]), "replica")
    return m


def test_single_large_replica_free_against_oracle():
    m = _single(en.single(), 30, 21)
    assert m["n"] == 64 * 64 * 48 * 10


def test_2562_vertex_membrane_free_against_oracle():
    d = en.replica(48, 48, 40, 8, 2562, None, 1.2, 1500, 20.0, 5.0, 0.08, 4242)
    m = _single(d, 30, 12)
    assert m["bonded"] > 0


def _batch_case(ds, sample_oracle, sample_alone, n_oracle=20, n_total=120):
    """Sampled replicas inside the batch against the oracle over n_oracle
    steps, then other sampled replicas after n_total steps against the same
    replicas run alone."""
    b = ia.Batch([en.spec(d) for d in ds])
    try:
        sts = [en.load(b, r, d) for r, d in enumerate(ds)]
        en.start(b)
        os_ = {}
        for r in sample_oracle:
            p = en.params(ds[r])
            o = fr.oracle_from_device(b, r, p, sts[r]["V"], sts[r]["E"], sts[r]["T"])
            o.set_fp32(True)
            o.use_philox(ds[r]["seed"])
            os_[r] = o
        b.md_steps(n_oracle)
        b.sync()
        for r, o in os_.items():
            for _ in range(n_oracle):
                o.md_step()
            m = fr.compare(b, r, o)
            print("replica", r, ds[r], m)
            fr.check(m, (ds[r]["Lx"], ds[r]["Ly"], ds[r]["Lz"]), "replica %d in batch" % r)
        del os_
        b.md_steps(n_total - n_oracle)
        b.sync()
        got = {r: (b.solvent(r), b.vertices(r), b.ligands(r)) for r in sample_alone}
    finally:
        b.close()
    for r in sample_alone:
        a = en.batch([ds[r]])
        try:
            a.md_steps(n_total)
            a.sync()
            s, v, lg = a.solvent(0), a.vertices(0), a.ligands(0)
        finally:
            a.close()
        gs, gv, gl = got[r]
        for k in s:
            np.testing.assert_array_equal(gs[k], s[k], err_msg="replica %d solvent %s" % (r, k))
        for k in v:
            np.testing.assert_array_equal(gv[k], v[k], err_msg="replica %d vertices %s" % (r, k))
        np.testing.assert_array_equal(gl, lg)


def test_uniform_64_replica_batch():
    ds = en.uniform()
    assert len(ds) == 64 and all((d["Lx"], d["Ly"], d["Lz"]) == (32, 32, 24) for d in ds)
    _batch_case(ds, sample_oracle=(5, 50), sample_alone=(17, 63))


def test_uneven_48_replica_batch():
    ds = en.uneven()
    assert len(ds) == 48
    big = max(range(48), key=lambda r: en.n_solvent(ds[r]))
    m2562 = next(r for r in range(48) if ds[r]["nv"] == 2562)
    m162 = next(r for r in range(48) if ds[r]["nv"] == 162)
    _batch_case(ds, sample_oracle=(m2562, m162), sample_alone=(big, 47))


@pytest.mark.parametrize("options,repack", [(None, False), (12, False), (12, True)])
def test_poisson_stream_of_uneven_requests(options, repack):
    """Requests drawn like the uneven batch, Poisson arrivals faster than
    they run, served under a budget that makes some wait; sampled results
    equal the same requests served alone.  With options 12 contexts of
    several replicas run the membrane on a second stream as grid kernels,
    with XCD chunking and captured graphs (the shape the autotuner picks for
    the uneven batch), while contexts are built and closed on the builder
    thread.  Micro-batches (repack False) run side by side; re-packing moves
    the running requests into each new context device to device."""
    reqs = serve.poisson_requests(12, 200.0, 2000, seed=5, series_every=100)
    budget = int((2.0 if repack else 0.5) * sum(r.bytes for r in reqs))
    srv = serve.Server(budget, chunk=100, max_active=3, window=0.005, options=options, repack=repack)
    srv.run(reqs)
    assert all(r.finished is not None for r in reqs)
    assert srv.used == 0
    assert srv.stats["groups"] >= 2
    assert srv.stats["max_active_groups"] >= 2
    for r in (reqs[0], reqs[6], reqs[11]):
        alone = serve.Server(r.bytes * 2, chunk=2000)
        a = serve.Request(r.id, r.replica, r.steps, r.series_every, 0.0)
        alone.run([a])
        np.testing.assert_array_equal(r.series, a.series, err_msg="request %d series" % r.id)
        for k in r.final:
            np.testing.assert_array_equal(r.final[k], a.final[k], err_msg="request %d %s" % (r.id, k))
