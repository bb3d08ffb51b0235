"""Checkpoint and resume: a replica saved between calls and loaded into
another context (another slot, other batch-mates) continues bit for bit
as if it had never stopped."""
import numpy as np
import pytest

import ensemble as en
import irbc_amd as ia
from test_gpu_batch import _mix, _snapshot


@pytest.mark.gpu
def test_resume_in_another_context_is_bit_exact():
    ds = _mix()
    b = en.batch(ds)
    b.md_steps(33)  # stops mid collision period and between list updates
    b.sync()
    ck = [b.checkpoint(r) for r in range(len(ds))]
    b.md_steps(40)
    b.sync()
    ref = [_snapshot(b, r) for r in range(len(ds))]
    b.close()
    order = [2, 0, 1]
    b2 = ia.Batch([en.spec(ds[i]) for i in order])
    for slot, i in enumerate(order):
        st = en.state(ds[i])
        b2.set_mesh(slot, st["V"], st["E"], st["T"])
        b2.restore(slot, ck[i])
    b2.md_steps(40)
    b2.sync()
    for slot, i in enumerate(order):
        got = _snapshot(b2, slot)
        for k in ref[i]:
            if k == "series":
                continue
            np.testing.assert_array_equal(got[k], ref[i][k], err_msg="replica %d %s" % (i, k))


@pytest.mark.gpu
def test_checkpoint_refuses_other_sizes():
    ds = _mix()
    b = en.batch(ds[:2])
    ck = b.checkpoint(0)
    with pytest.raises(ia.IrbcError):
        b.restore(1, ck)
    with pytest.raises(ia.IrbcError):
        b.restore(0, ck[:100])


@pytest.mark.gpu
def test_checkpoint_refuses_foreign_parameters_and_corrupt_indices():
    """A checkpoint of a replica with the same sizes but another shear rate
    is refused, and so is one whose cell->triangle entries or particle ids
    were corrupted: all are checked on the host before any copy, and the
    replica then runs on as it was."""
    d0 = _mix()[0]
    d1 = dict(d0, shear=d0["shear"] * 0.7)  # same sizes, other uplist and wall speed
    b = en.batch([d0, d1])
    try:
        b.md_steps(5)
        b.sync()
        ck = b.checkpoint(0)
        before = _snapshot(b, 1)
        with pytest.raises(ia.IrbcError, match="other parameters"):
            b.restore(1, ck)
        bad = bytearray(ck)
        bad[-4:] = (0x7FFFFFFF).to_bytes(4, "little")  # the last cell-list entry
        with pytest.raises(ia.IrbcError, match="entry out of range"):
            b.restore(0, bytes(bad))
        # the particle ids follow the six coordinate arrays
        spec = b.specs[0]
        ns, nv, nlig, ncell = spec.n_solvent, spec.n_vertices, b.nlig(0), b.ncells(0)
        llen = len(b.llist(0)[1])
        tail = ns * 28 + nv * (9 * 8 + 6 * 4) + nlig * 4 + 12 * 8 + 4 * 4 + (ncell + 1) * 4 + llen * 4
        ids = len(ck) - tail + 6 * ns * 4
        bad = bytearray(ck)
        bad[ids:ids + 4] = bad[ids + 4:ids + 8]  # a duplicated id
        with pytest.raises(ia.IrbcError, match="permutation"):
            b.restore(0, bytes(bad))
        b.restore(0, ck)  # the good one still loads
        b.md_steps(3)
        b.sync()
        after = _snapshot(b, 1)
        assert after["counters"][0] == before["counters"][0] + 3
    finally:
        b.close()


@pytest.mark.gpu
def test_counters_out_of_range_are_refused():
    b = en.batch(_mix()[:1])
    try:
        p = b.specs[0].params
        for bad in ((0, p.csteps, 0), (0, -2, 0), (0, -1, p.uplist), (-1, -1, 0)):
            with pytest.raises(ia.IrbcError):
                b.set_counters(0, *bad)
        b.set_counters(0, 0, p.csteps - 1, p.uplist - 1)
    finally:
        b.close()


@pytest.mark.gpu
def test_copy_replica_device_to_device_is_bit_exact():
    """irbc_copy_replica moves a running replica into a slot of another
    context without a host round trip of its state; the copy continues bit
    for bit, time series included, and the source is left as it was.  The
    first slot had its mesh and series set beforehand, the others take
    them from the source."""
    ds = _mix()
    b = en.batch(ds)
    for r in range(len(ds)):
        b.set_series(r, 10, 16)
    b.md_steps(33)  # mid collision period, between list updates
    b.sync()
    order = [1, 2, 0]
    b2 = ia.Batch([en.spec(ds[i]) for i in order])
    try:
        for slot, i in enumerate(order):
            if slot == 0:
                V, E, T = en.mesh(ds[i])
                b2.set_mesh(slot, V, E, T)
                b2.set_series(slot, 10, 16)
            b2.copy_from(slot, b, i)
        b.md_steps(40)
        b2.md_steps(40)
        b.sync()
        b2.sync()
        for slot, i in enumerate(order):
            want, got = _snapshot(b, i), _snapshot(b2, slot)
            assert len(got["series"]) == 7
            for k in want:
                np.testing.assert_array_equal(got[k], want[k], err_msg="replica %d %s" % (i, k))
    finally:
        b2.close()
        b.close()


@pytest.mark.gpu
def test_copy_replica_refuses_other_sizes_and_parameters():
    ds = _mix()
    d1 = dict(ds[0], shear=ds[0]["shear"] * 0.7)  # same sizes, other parameters
    b = en.batch([ds[0], ds[1], d1])
    b2 = ia.Batch([en.spec(ds[0])])
    try:
        b.set_series(0, 10, 4)
        b.md_steps(12)
        b.sync()
        with pytest.raises(ia.IrbcError, match="other sizes"):
            b2.copy_from(0, b, 1)
        with pytest.raises(ia.IrbcError, match="other parameters"):
            b2.copy_from(0, b, 2)
        with pytest.raises(ia.IrbcError, match="not set up"):
            b.copy_from(0, b2, 0)
        b2.copy_from(0, b, 0)
        assert len(b2.series(0)) == 1
        assert b2.counters(0) == b.counters(0)
    finally:
        b2.close()
        b.close()
