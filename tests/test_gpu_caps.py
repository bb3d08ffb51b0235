"""The device's fixed capacities: at the limit the results are the
oracle's, past it the step is refused loudly, never truncated silently.

- Membrane vertices per collision cell (IRBC_MAX_VPC = 32, the per-lane
  list k_cell keeps in LDS): a 2562-vertex sphere of radius 3.1 puts up to
  30 vertices in a cell, and the collision equals the oracle's (keys,
  counts, ghost masses, vertex velocities); at radius 2.5 some cells hold 50,
  and the collision raises the range status (irbc_sync reports it).
- Ligand candidates per knob (IRBC_MAX_CAND = 16 and a window of 6 grid
  lines per axis): with the default dthr = 0.5 on the 0.5 grid a knob has a
  handful; with dthr = 1.2 its window spans more lines than the device keeps,
  and the bond preparation raises the candidate-overflow status.
The reference has no such capacities (maincode.cpp:138-154,
Src/solfuncs.cpp:208-220 grow vectors); these inputs are far outside the
benchmark configurations, which the free-running tests cover.
"""
import numpy as np
import pytest

from cases import Case, f32_ulps

pytestmark = pytest.mark.gpu

IRBC_ST_CAND_OVERFLOW = 0x2
IRBC_ST_RANGE = 0x8


@pytest.fixture(scope="module")
def ia():
    import irbc_amd
    irbc_amd.load()
    return irbc_amd


def _collide(case, b):
    o = case.o
    o.set_wallspeed(0.1)
    b.set_wallspeed(0, 0.1)
    o.seed_mt(11, 22)
    o.collision()
    b.set_collision_tape(0, *o.get_collision_tape())
    b.collision_step()


def test_vertices_per_cell_at_the_cap_match_the_oracle(ia):
    case = Case(seed=5, nv=2562, radius=3.1, Lx=16, Ly=16, Lz=12)
    b = case.make_gpu()
    try:
        b.set_diagnostics(True)
        _collide(case, b)
        o = case.o
        per_cell = np.bincount(o.vertices()["pid"])
        assert 28 <= per_cell.max() <= 32, per_cell.max()
        so, sg = o.solvent(), b.solvent()
        np.testing.assert_array_equal(sg["key"], so["pid"])
        co, cg = o.cells(), b.cells()
        np.testing.assert_array_equal(cg["nc"], co["nc"])
        np.testing.assert_array_equal(cg["ncc"], co["ncc"])
        np.testing.assert_allclose(cg["vcm"], co["vcm"], rtol=1e-12, atol=1e-14)
        for k in ("vx", "vy", "vz"):
            assert f32_ulps(so[k], sg[k]).max() <= 2.0, k
        vo, vg = o.vertices(), b.vertices()
        np.testing.assert_array_equal(vg["pid"], vo["pid"])
        for k in ("vx", "vy", "vz"):
            np.testing.assert_allclose(vg[k], vo[k], rtol=1e-10, atol=1e-12)
    finally:
        b.close()


def test_vertices_per_cell_past_the_cap_are_refused(ia):
    case = Case(seed=5, nv=2562, radius=2.5, Lx=16, Ly=16, Lz=12)
    b = case.make_gpu()
    try:
        with pytest.raises(ia.IrbcError, match="status 0x%x" % IRBC_ST_RANGE):
            _collide(case, b)
        assert np.bincount(case.o.vertices()["pid"]).max() > 32
    finally:
        b.close()


def test_ligand_candidates_past_the_cap_are_refused(ia):
    case = Case(seed=8, kon=40.0, koff=40.0, z_low=0.1)
    gp = ia.Params.default(**case.kw)
    gp.dthr = 1.2
    spec = ia.make_spec(gp, case.ns, len(case.V), len(case.E), len(case.T), seed=case.seed)
    b = ia.Batch([spec])
    try:
        b.set_mesh(0, case.V, case.E, case.T)
        case.load_gpu(b, 0)
        _collide(case, b)  # the knobs' collision cells
        v = b.vertices()
        assert (v["knob"][v["z"] < gp.dthr] == 1).any()
        with pytest.raises(ia.IrbcError, match="status 0x%x" % IRBC_ST_CAND_OVERFLOW):
            b.bond_step()
    finally:
        b.close()
