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
