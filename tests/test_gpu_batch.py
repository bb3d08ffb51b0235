"""Batching: a replica gives the same bits alone and inside an uneven batch.

Replicas share every kernel launch but no state: cell sums are exact
fixed point, membrane sums run in a fixed per-replica order and the
counter-based streams are keyed by (replica seed, step, cell).  So every
array of a replica run in a batch of three different boxes, meshes, rates
and shear rates must equal the same replica run alone, bit for bit, over
steps that include collisions, cell-list updates and bond events.
"""
import numpy as np
import pytest

import ensemble as en
import irbc_amd as ia


def _mix():
    return [
        en.replica(16, 12, 10, 10, 162, 2.6, 1.0, 162, kon=200.0, koff=20.0, shear=0.1, seed=11),
        en.replica(12, 12, 8, 6, 42, 1.8, 1.2, 30, kon=5.0, koff=1.0, shear=0.05, seed=12),
        en.replica(20, 16, 12, 8, 162, 3.0, 1.3, 100, kon=0.5, koff=0.05, shear=0.02, seed=13),
    ]


def _snapshot(b, r):
    out = dict(b.solvent(r))
    out.update({"v_" + k: a for k, a in b.vertices(r).items()})
    out["ligands"] = b.ligands(r)
    d, i = b.parasite(r)
    out["para_d"], out["para_i"] = d, i
    out["counters"] = np.array(b.counters(r))
    g = b.globals(r)
    out["globals"] = np.array([g[k] for k in ("Sarea", "Volume", "wallspeed", "step", "nbonds")])
    assert g["status"] == 0, "replica %d raised device status %d" % (r, g["status"])
    assert g["nbonds"] == out["v_bondi"].sum()
    out["series"] = b.series(r)
    return out


@pytest.mark.gpu
def test_replica_alone_equals_replica_in_batch():
    ds = _mix()
    steps = 60  # 6 collisions; list updates every uplist steps per replica
    b = en.batch(ds)
    for r in range(len(ds)):
        b.set_series(r, 10, 16)
    b.md_steps(steps)
    b.sync()
    mixed = [_snapshot(b, r) for r in range(len(ds))]
    b.close()
    assert len(mixed[0]["series"]) == steps // 10
    for r, d in enumerate(ds):
        a = en.batch([d])
        a.set_series(0, 10, 16)
        a.md_steps(steps)
        a.sync()
        alone = _snapshot(a, 0)
        a.close()
        for k in alone:
            np.testing.assert_array_equal(mixed[r][k], alone[k], err_msg="replica %d %s" % (r, k))
    assert mixed[0]["globals"][4] > 0  # the fast-rate replica formed bonds


@pytest.mark.gpu
def test_batch_split_calls_equal_one_call():
    """md_steps(a) then md_steps(b) equals md_steps(a+b): no state is lost
    between calls (device-advanced counters, parity buffers, pending shift)."""
    ds = _mix()
    b1 = en.batch(ds)
    b1.md_steps(37)
    b1.sync()
    b2 = en.batch(ds)
    for n in (5, 1, 11, 20):
        b2.md_steps(n)
    b2.sync()
    for r in range(len(ds)):
        s1, s2 = _snapshot(b1, r), _snapshot(b2, r)
        for k in s1:
            np.testing.assert_array_equal(s1[k], s2[k], err_msg="replica %d %s" % (r, k))


@pytest.mark.gpu
def test_graph_replay_equals_eager():
    """md_steps through captured HIP graphs (replayed several times, with
    host reads between the replays) gives the same bits as kernel-by-kernel
    launches."""
    ds = _mix()
    b1 = en.batch(ds)
    b1.md_steps(73)
    b1.sync()
    b2 = en.batch(ds)
    b2.set_option(ia.OPT_GRAPHS, 1)
    for n in (30, 20, 23):
        b2.md_steps(n)
        b2.sync()
        assert b2.globals(0)["status"] == 0
    for r in range(len(ds)):
        s1, s2 = _snapshot(b1, r), _snapshot(b2, r)
        for k in s1:
            np.testing.assert_array_equal(s1[k], s2[k], err_msg="replica %d %s" % (r, k))


def _fused_vs_grid(ds):
    out = []
    for fused in (0, 1):
        b = en.batch(ds)
        b.set_option(ia.OPT_MEMBRANE_CHAIN, 0)
        b.set_option(ia.OPT_MEMBRANE_FUSED, fused)
        b.md_steps(20)
        b.sync()
        out.append([(b.vertices(r), b.globals(r)) for r in range(len(ds))])
    for r in range(len(ds)):
        (v0, g0), (v1, g1) = out[0][r], out[1][r]
        for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
            scale = max(1.0, np.abs(v0[k]).max())
            np.testing.assert_allclose(v1[k], v0[k], rtol=1e-9, atol=1e-9 * scale, err_msg="%d %s" % (r, k))
        for k in ("ix", "iy", "bondi", "bondl"):
            np.testing.assert_array_equal(v1[k], v0[k])
        assert abs(g1["Sarea"] - g0["Sarea"]) <= 1e-12 * g0["Sarea"]
        assert abs(g1["Volume"] - g0["Volume"]) <= 1e-12 * g0["Volume"]


@pytest.mark.gpu
def test_fused_membrane_matches_grid_kernels():
    """The one-workgroup membrane step and the five grid-wide kernels differ
    only in the summation order of the global area and volume."""
    _fused_vs_grid(_mix())


@pytest.mark.gpu
def test_fused_membrane_with_lds_positions_far_into_the_batch():
    """Replicas whose vertices start far into the packed arrays (here 3 x 642
    before the last), so the one-workgroup step's LDS copy of the unwrapped
    positions is indexed relative to the replica, not by global index."""
    ds = [en.replica(16, 16, 16, 6, 642, None, 1.0, 100, kon=0.5, koff=0.05, shear=0.05, seed=40 + i)
          for i in range(4)]
    _fused_vs_grid(ds)


@pytest.mark.gpu
@pytest.mark.parametrize("graphs,dev,oneq,four", [(0, 0, 1, 0), (1, 0, 1, 0), (1, 0, 1, 1), (1, 0, 0, 0), (1, 1, 0, 0)])
def test_membrane_chain_equals_grid_kernels(graphs, dev, oneq, four):
    """The one-launch membrane step (workgroups ordered by ticket) -- inside
    the streaming and bounce-back launches of a one-queue step (either
    stream shape), or on its own stream with XCD-local or device-scope
    hand-offs -- runs the five grid kernels' arithmetic element by element
    (those on their own stream): same bits, in an
    uneven batch whose smaller meshes leave some of the launch's tickets
    idle, over steps with collisions, list updates and bond events."""
    ds = _mix() + [en.replica(16, 16, 16, 6, 642, None, 1.0, 300, kon=100.0, koff=10.0, shear=0.05, seed=21)]
    snaps = []
    for chain in (0, 1):
        b = en.batch(ds)
        b.set_option(ia.OPT_GRAPHS, graphs)
        b.set_option(ia.OPT_MEMBRANE_FUSED, 0)
        b.set_option(ia.OPT_MEMBRANE_CHAIN, chain)
        b.set_option(ia.OPT_CHAIN_DEVICE_SCOPE, dev)
        b.set_option(ia.OPT_ONE_QUEUE, oneq)
        b.set_option(ia.OPT_STREAM4, four)
        for n in (13, 40):
            b.md_steps(n)
            b.sync()
        snaps.append([_snapshot(b, r) for r in range(len(ds))])
        b.close()
    for r in range(len(ds)):
        for key in snaps[0][r]:
            np.testing.assert_array_equal(snaps[0][r][key], snaps[1][r][key], err_msg="replica %d %s" % (r, key))
    assert snaps[1][0]["globals"][4] > 0  # the fast-rate replica formed bonds


@pytest.mark.gpu
def test_near_masks_never_change_results():
    """The sub-cell near masks only decide which particles go to the
    bounce-back kernels: with them off (every particle of a listed cell
    queued) the trajectories are the same bits over 120 steps (12
    collisions, list updates), for fast walls and fast-moving membranes too,
    and with them on the queue is shorter."""
    ds = _mix() + [en.replica(20, 20, 16, 10, 642, None, 1.0, 300, kon=5.0, koff=1.0, shear=0.25, seed=31)]
    snaps, queues = [], []
    for masks in (0, 1):
        b = en.batch(ds)
        b.set_option(ia.OPT_NEAR_MASKS, masks)
        q = np.zeros(len(ds))
        for _ in range(12):
            b.md_steps(10)
            b.sync()
            q += [b.globals(r)["near_queue"] for r in range(len(ds))]
        snaps.append([_snapshot(b, r) for r in range(len(ds))])
        queues.append(q)
        b.close()
    for r in range(len(ds)):
        for key in snaps[0][r]:
            np.testing.assert_array_equal(snaps[0][r][key], snaps[1][r][key], err_msg="replica %d %s" % (r, key))
    assert np.all(queues[1] < queues[0]), queues


@pytest.mark.gpu
def test_sort_schedule_never_changes_results():
    """Collisions that leave the particles in their slots (k_bin's LDS-summed
    momentum, k_dvv, k_cell_the, the rotation in place) give the same bits
    as collisions that sort every time: 120 steps (12 collisions, list
    updates) of an uneven batch with the collision diagnostics on, sorting
    every collision, every third (in place in between, in both streaming
    shapes, on one queue with graphs and on two queues), and never after the
    first.  Solvent state, keys, vertices, bonds, counters and the last
    collision's cell parameters all equal."""
    ds = _mix() + [en.replica(20, 20, 16, 10, 642, None, 1.0, 300, kon=5.0, koff=1.0, shear=0.25, seed=31)]
    runs = []
    for every, opts in ((1, {}), (3, {}), (3, {ia.OPT_STREAM4: 1}),
                        (3, {ia.OPT_ONE_QUEUE: 0, ia.OPT_GRAPHS: 0}), (1000, {ia.OPT_STREAM4: 1}),
                        (3, {ia.OPT_BIN_SPAN: 4})):
        print("sort every", every, opts, flush=True)
        b = en.batch(ds)
        b.set_option(ia.OPT_SORT_EVERY, every)
        for k, v in opts.items():
            b.set_option(k, v)
        b.set_diagnostics(True)
        kinds = []
        for _ in range(12):
            b.md_steps(10)
            kinds.append(b.globals(0)["collision_kind"])
        # call k runs steps 10k .. 10k + 9, the collision (csteps 10) at
        # step 10k first; it sorts when k is a multiple of the schedule (none
        # before the first collision reads as sorted)
        assert kinds == [1 if k % every == 0 else 2 for k in range(12)], kinds
        snap = [_snapshot(b, r) for r in range(len(ds))]
        for r in range(len(ds)):
            snap[r].update({"cell_" + k: a for k, a in b.cells(r).items()})
        runs.append(snap)
        b.close()
    for run in runs[1:]:
        for r in range(len(ds)):
            for key in runs[0][r]:
                np.testing.assert_array_equal(run[r][key], runs[0][r][key], err_msg="replica %d %s" % (r, key))


@pytest.mark.gpu
def test_sort_every_rejects_zero():
    b = en.batch(_mix()[:1])
    try:
        with pytest.raises(ia.IrbcError, match="SORT_EVERY"):
            b.set_option(ia.OPT_SORT_EVERY, 0)
    finally:
        b.close()


@pytest.mark.gpu
def test_autotune_leaves_the_trajectory_alone():
    """The autotuner's trial steps are undone: a batch tuned and then run
    equals a batch run with the chosen options from the start."""
    ds = _mix()
    b1 = en.batch(ds)
    b1.md_steps(7)
    k, rep = b1.autotune(12)
    assert 0 <= k < 2 ** len(ia.OPTIONS) == len(rep)
    idx = np.arange(len(rep))
    # timed: no one-workgroup membrane, capped grid or late membrane start;
    # one queue only with the chain
    timed = ((idx & 0b1100001) == 0) & (((idx >> 7) & 1) | (1 - ((idx >> 8) & 1)) == 1)
    assert timed.sum() == 48
    assert np.all(np.isnan(rep[~timed])) and np.all(rep[timed] > 0) and timed[k]
    b1.md_steps(25)
    b1.sync()
    b2 = en.batch(ds)
    for bit, (_, opt) in enumerate(ia.OPTIONS):
        b2.set_option(opt, (k >> bit) & 1)
    b2.md_steps(32)
    b2.sync()
    for r in range(len(ds)):
        s1, s2 = _snapshot(b1, r), _snapshot(b2, r)
        for key in s1:
            np.testing.assert_array_equal(s1[key], s2[key], err_msg="replica %d %s" % (r, key))


@pytest.mark.gpu
def test_launch_options_do_not_change_results():
    """Near-membrane shape, XCD chunking, the stream shape and where the
    membrane part of a step starts are pure launch choices."""
    ds = _mix()
    snaps = []
    for near, xcd, s4, after in ((1, 1, 0, 0), (0, 1, 0, 0), (1, 0, 0, 0), (1, 1, 1, 0), (0, 1, 0, 1), (1, 1, 1, 1)):
        b = en.batch(ds)
        b.set_option(ia.OPT_NEAR_MERGED, near)
        b.set_option(ia.OPT_XCD_CHUNK, xcd)
        b.set_option(ia.OPT_STREAM4, s4)
        b.set_option(ia.OPT_MEM_AFTER_STREAM, after)
        b.md_steps(31)
        b.sync()
        snaps.append([_snapshot(b, r) for r in range(len(ds))])
    for other in snaps[1:]:
        for r in range(len(ds)):
            for key in snaps[0][r]:
                np.testing.assert_array_equal(snaps[0][r][key], other[r][key], err_msg="replica %d %s" % (r, key))


@pytest.mark.gpu
def test_capped_stream_grid_does_not_change_results():
    """The bulk streaming kernels on a capped grid, whose workgroups loop over
    more tiles than the grid has (two 32x32x24 replicas: 1920 tiles), give the
    same bits as one workgroup per tile, in both stream shapes."""
    ds = en.uniform(2)
    snaps = []
    for s4, capped in ((0, 0), (0, 1), (1, 0), (1, 1)):
        b = en.batch(ds)
        b.set_option(ia.OPT_STREAM4, s4)
        b.set_option(ia.OPT_STREAM_GRID, capped)
        b.md_steps(23)
        b.sync()
        assert b.globals(0)["status"] == 0
        snaps.append([_snapshot(b, r) for r in range(len(ds))])
        b.close()
    for other in snaps[1:]:
        for r in range(len(ds)):
            for key in snaps[0][r]:
                np.testing.assert_array_equal(snaps[0][r][key], other[r][key], err_msg="replica %d %s" % (r, key))


@pytest.mark.gpu
def test_mask_keep_and_queue_are_deterministic():
    """Two runs of the same batch give the same queue lengths and states,
    step after step, across collisions that keep or rebuild the near masks.
    (A rebuild decided inside the mask launch could be seen half-way: its
    first workgroup clears the drift that the others' decision reads, so
    some cells kept stale masks and the queues differed between runs.)"""
    ds = [en.replica(32, 32, 24, 10, nv, None, 1.0 + 0.1 * (r % 4), 100 + 40 * r, 20.0 * (1 + r % 3),
                     2.0 * (1 + r % 2), 0.05, 5000 + r) for r, nv in enumerate([162, 642, 2562, 642] * 2)]
    runs = []
    for _ in range(2):
        b = en.batch(ds)
        qs = []
        for _ in range(12):
            b.md_steps(25)
            b.sync()
            qs.append([b.globals(r)["near_queue"] for r in range(len(ds))])
            assert all(b.globals(r)["status"] == 0 for r in range(len(ds)))
        sol = [b.solvent(r)["x"] for r in range(len(ds))]
        runs.append((np.array(qs), sol))
        b.close()
    np.testing.assert_array_equal(runs[0][0], runs[1][0])
    for r in range(len(ds)):
        np.testing.assert_array_equal(runs[0][1][r], runs[1][1][r])


@pytest.mark.gpu
def test_membrane_chain_of_a_replica_stays_on_one_xcd():
    """The chain's XCD-local hand-offs need every workgroup of a replica's
    chain on one XCD; each workgroup checks its XCD from the hardware
    register.  On this device the dispatch keeps replica r's chain on XCD
    r % 8, on one queue and on two, and no replica raises the check."""
    ds = [en.replica(12, 12, 8, 6, 42, 1.8, 1.0, 20, seed=60 + r) for r in range(10)]
    for oneq in (1, 0):
        b = en.batch(ds)
        try:
            b.set_option(ia.OPT_ONE_QUEUE, oneq)
            b.md_steps(12)
            b.sync()  # raises on IRBC_ST_XCD
            for r in range(len(ds)):
                g = b.globals(r)
                assert g["status"] == 0
                assert g["chain_xcd"] == r % 8, (oneq, r, g["chain_xcd"])
        finally:
            b.close()


@pytest.mark.gpu
def test_lists_at_collisions_never_change_results():
    """Cell->triangle lists updated at the last collision before each
    replica's deadline (IRBC_OPT_ALIGN_LISTS) give the same bits as the
    reference's cadence, over 150 steps of a batch whose replicas have list
    cadences of 6 (shorter than a collision period), 12, 31 and 83 steps;
    with alignment, list updates happen on collision steps only."""
    ds = [en.replica(12, 12, 12, 10, 42, 1.8, 1.0, 42, kon=5.0, koff=1.0, shear=3.0, seed=71),
          en.replica(16, 12, 10, 10, 162, 2.6, 1.0, 162, kon=200.0, koff=20.0, shear=2.0, seed=72),
          en.replica(12, 12, 8, 8, 42, 1.8, 1.2, 30, kon=5.0, koff=1.0, shear=1.0, seed=73),
          en.replica(20, 16, 12, 8, 162, 3.0, 1.3, 100, kon=0.5, koff=0.05, shear=0.25, seed=74)]
    upl = [en.params(d).uplist for d in ds]
    assert sorted(upl) == [6, 12, 31, 83], upl
    snaps = []
    for align in (0, 1):
        b = en.batch(ds)
        b.set_option(ia.OPT_ALIGN_LISTS, align)
        for s in range(150):
            b.md_steps(1)
            for r in range(len(ds)):
                step, colsteps, upd = b.counters(r)
                if align and upd == 0 and upl[r] >= 10:
                    assert colsteps == 0, (r, step)  # a list update on a collision step
        b.sync()
        snaps.append([_snapshot(b, r) for r in range(len(ds))])
        b.close()
    for r in range(len(ds)):
        for key in snaps[0][r]:
            if key == "counters":
                continue  # the list counters differ by design
            np.testing.assert_array_equal(snaps[0][r][key], snaps[1][r][key], err_msg="replica %d %s" % (r, key))


@pytest.mark.gpu
def test_bin_span_does_not_change_results():
    """The bin-sort tile shape (1, 2 or 4 particle tiles per binning
    workgroup, spans crossing from one replica into the next in a batch of
    uneven sizes) gives the same bits over 41 steps (four collisions, list
    updates), with and without XCD chunking and graphs."""
    ds = _mix() + [en.replica(20, 20, 16, 10, 642, None, 1.0, 300, kon=5.0, koff=1.0, shear=0.25, seed=31)]
    snaps = []
    for span, xcd, graphs in ((1, 1, 1), (2, 1, 1), (4, 1, 1), (4, 0, 0), (2, 0, 1)):
        b = en.batch(ds)
        b.set_option(ia.OPT_BIN_SPAN, span)
        b.set_option(ia.OPT_XCD_CHUNK, xcd)
        b.set_option(ia.OPT_GRAPHS, graphs)
        assert b.option(ia.OPT_BIN_SPAN) == span
        b.md_steps(41)
        b.sync()
        snaps.append([_snapshot(b, r) for r in range(len(ds))])
        b.close()
    for other in snaps[1:]:
        for r in range(len(ds)):
            for key in snaps[0][r]:
                np.testing.assert_array_equal(snaps[0][r][key], other[r][key], err_msg="replica %d %s" % (r, key))
    b = en.batch(ds[:1])
    with pytest.raises(ia.IrbcError):
        b.set_option(ia.OPT_BIN_SPAN, 3)
    b.close()
