"""The fused time loop (irbc_md_steps) against the oracle, one step at a time.

Each step starts the oracle from the device's state (solvent as stored in
FP32, membrane, bead, bonds, ligands, counters), runs one reference step
(maincode.cpp:461-744) drawing MT19937 streams, and hands the recorded tapes
to the device for the same step.  Cell keys, bond states and ligand
occupancy must match exactly at every step, positions and velocities within
FP32 storage rounding.  The schedule covers collision steps (with the
rotation fused into the next streaming kernel), a cell->triangle list
update, and binding/unbinding at high rates.
"""
import numpy as np
import pytest

from cases import Case, f32_ulps, load_oracle_from_gpu, periodic_diff

pytestmark = pytest.mark.gpu


def test_md_steps_match_oracle_per_step():
    import irbc_amd  # noqa: F401
    case = Case(seed=21, kon=20.0, koff=20.0, z_low=0.12, walls=1.0)
    o = case.o
    p = case.po_params
    b = case.make_gpu()
    o.llistupdate()
    b.llist_update()
    # collisions on steps 2, 12, 22; list update on step 4
    o.set_counters(100, 7, p.uplist - 5)
    b.set_counters(0, 100, 7, p.uplist - 5)
    b.use_tapes(0, True)
    b.set_diagnostics(True)
    o.seed_mt(4242, 4343)
    checked_keys = 0
    for step in range(26):
        if step:
            load_oracle_from_gpu(o, b)
        _, cs, upd = o.counters()
        collides = cs + 1 == p.csteps
        o.md_step()
        if collides:
            b.set_collision_tape(0, *o.get_collision_tape())
        b.set_bond_tape(0, o.get_bond_tape())
        b.md_steps(1)
        so, sg = o.solvent(), b.solvent()
        if collides:
            np.testing.assert_array_equal(sg["key"], so["pid"])
            np.testing.assert_array_equal(b.cells()["nc"], o.cells()["nc"])
            checked_keys += 1
        for k, L in (("x", p.Lx), ("y", p.Ly)):
            assert np.abs(periodic_diff(sg[k], so[k], L)).max() <= 4 * 2.0 ** -23 * L, (step, k)
        assert f32_ulps(so["z"], sg["z"]).max() <= 4, step
        for k in ("vx", "vy", "vz"):
            assert f32_ulps(so[k], sg[k]).max() <= 4, (step, k)
        vo, vg = o.vertices(), b.vertices()
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_allclose(vg[k], vo[k], rtol=1e-9, atol=1e-12, err_msg="%d %s" % (step, k))
        for k in ("ix", "iy", "bondi", "bondl", "pid"):
            np.testing.assert_array_equal(vg[k], vo[k], err_msg="%d %s" % (step, k))
        np.testing.assert_array_equal(b.ligands(), o.ligands())
        np.testing.assert_allclose(b.parasite()[0], o.parasite()[0], rtol=1e-9, atol=1e-12)
        assert b.counters() == o.counters()
    assert checked_keys == 3
    assert vg["bondi"].sum() > 0
