"""HIP path against the FP64 oracle, stage by stage, through the C ABI.

Both sides start from the same state (FP32-representable solvent values) and
consume the same random streams: the oracle draws them from MT19937 in the
reference's order and records them as tapes, the device reads the tapes.
Integer results (cell keys, counts, cell lists, bond states, ligand
occupancy) must match exactly; floating-point results within the stated
tolerances:
  * per-cell momentum (vcm): relative 1e-12 (exact fixed-point sums on the
    device, sequential FP64 on the oracle);
  * membrane area/volume: relative 1e-12 (target 1e-8);
  * solvent positions/velocities: 2 FP32 ulps of the oracle's FP64 value
    (the device stores FP32, computes FP64);
  * membrane forces and velocities: relative 1e-10.
"""
import numpy as np
import pytest

from cases import Case, f32_ulps, periodic_diff

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ia():
    import irbc_amd
    irbc_amd.load()
    return irbc_amd


def _collide_both(case, b, wallspeed=0.1, seed=(11, 22)):
    o = case.o
    o.set_wallspeed(wallspeed)
    b.set_wallspeed(0, wallspeed)
    o.seed_mt(*seed)
    o.collision()
    tapes = o.get_collision_tape()
    b.set_collision_tape(0, *tapes)
    b.collision_step()
    return tapes


@pytest.mark.parametrize("seed", [1, 2])
def test_collision_keys_counts_and_vcm(ia, seed):
    case = Case(seed=seed)
    b = case.make_gpu()
    _collide_both(case, b)
    o = case.o
    so, sg = o.solvent(), b.solvent()
    # cell keys of every particle, bit-exact
    np.testing.assert_array_equal(sg["key"], so["pid"])
    co, cg = o.cells(), b.cells()
    np.testing.assert_array_equal(cg["nc"], co["nc"])
    np.testing.assert_array_equal(cg["ncc"], co["ncc"])
    np.testing.assert_array_equal(cg["rand1"], co["rand1"])
    np.testing.assert_array_equal(cg["rand2"], co["rand2"])
    np.testing.assert_allclose(cg["vcm"], co["vcm"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(cg["dvv"], co["dvv"], rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(cg["the"], co["the"], rtol=1e-10, atol=1e-12)
    # rotated velocities: FP32 storage of the FP64 result
    for k in ("vx", "vy", "vz"):
        assert f32_ulps(so[k], sg[k]).max() <= 2.0, k
    vo, vg = o.vertices(), b.vertices()
    np.testing.assert_array_equal(vg["pid"], vo["pid"])
    for k in ("vx", "vy", "vz"):
        np.testing.assert_allclose(vg[k], vo[k], rtol=1e-10, atol=1e-12)
    do, io = o.parasite()
    dg, ig = b.parasite()
    assert ig[2] == io[2]
    np.testing.assert_allclose(dg[3:6], do[3:6], rtol=1e-10)


def test_collision_conserves_momentum_and_sets_kT(ia):
    """Invariants the algebra fixes: a ghost-free bulk cell keeps its momentum
    and, after the thermostat, sum |v - vcm|^2 = 3 (nc - 1) kT."""
    case = Case(seed=5, nv=42, radius=1.2, z_low=4.0)
    b = case.make_gpu()
    before = b.solvent()
    b.set_wallspeed(0, 0.1)
    b.use_tapes(0, False)
    b.set_diagnostics(True)
    b.collision_step()
    after = b.solvent()
    c = b.cells()
    p = case.po_params
    key = after["key"]
    k = key % (p.Lz + 1)
    vkey = np.unique(b.vertices()["pid"])
    bulk = (k >= 3) & (k <= p.Lz - 2) & ~np.isin(key, vkey) & (key != b.parasite()[1][2])
    nc = np.bincount(key, minlength=p.ncells())
    for comp in ("vx", "vy", "vz"):
        s0 = np.bincount(key[bulk], weights=before[comp][bulk], minlength=p.ncells())
        s1 = np.bincount(key[bulk], weights=after[comp][bulk], minlength=p.ncells())
        assert np.max(np.abs(s1 - s0)) < 2e-5
    vc = c["vcm"].reshape(-1, 3)
    dv2 = np.zeros(p.ncells())
    for j, comp in enumerate(("vx", "vy", "vz")):
        dv2 += np.bincount(key[bulk], weights=(after[comp][bulk] - vc[key[bulk], j]) ** 2,
                           minlength=p.ncells())
    cells = np.unique(key[bulk])
    cells = cells[nc[cells] > 1]
    np.testing.assert_allclose(dv2[cells], 3.0 * (nc[cells] - 1), rtol=2e-6)


def test_llist_matches_oracle(ia):
    case = Case(seed=3)
    b = case.make_gpu()
    case.o.llistupdate()
    b.llist_update()
    off_o, ent_o = case.o.llist()
    off_g, ent_g = b.llist()
    np.testing.assert_array_equal(off_g, off_o)
    np.testing.assert_array_equal(ent_g, ent_o)
    assert len(ent_g) > 1000


def test_inside_flags_match_oracle(ia):
    case = Case(seed=4)
    b = case.make_gpu()
    b.classify()
    np.testing.assert_array_equal(b.solvent()["indio"], case.indio)
    assert (case.indio == 0).sum() > 100  # particles inside the cell exist


@pytest.mark.parametrize("wallspeed", [0.1, 2.4])
def test_stream_bounce_back_and_walls(ia, wallspeed):
    """Ten streaming steps; each step both sides start from the device state.
    Vertex forces are zero, so the first half-kick the device applies inside
    the stream kernel leaves the vertex velocities as the oracle has them."""
    case = Case(seed=6)
    o = case.o
    nv = len(case.V)
    o.set_vertices(fx=np.zeros(nv), fy=np.zeros(nv), fz=np.zeros(nv))
    b = case.make_gpu()
    o.llistupdate()
    b.llist_update()
    b.set_wallspeed(0, wallspeed)
    p = case.po_params
    moved = 0
    for it in range(10):
        if it:
            s = b.solvent()
            o.set_solvent(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
        before = o.solvent()
        o.stream(wallspeed)
        b.stream()
        so, sg = o.solvent(), b.solvent()
        for k, L in (("x", p.Lx), ("y", p.Ly)):
            d = periodic_diff(sg[k], so[k], L)
            assert np.abs(d).max() <= 4 * 2.0 ** -23 * L, (it, k)
        assert f32_ulps(so["z"], sg["z"]).max() <= 4, it
        for k in ("vx", "vy", "vz"):
            assert f32_ulps(so[k], sg[k]).max() <= 2, (it, k)
        # the same particles were reflected by the membrane or a wall
        moved_o = (so["vx"] != before["vx"]) | (so["vz"] != before["vz"])
        moved_g = (sg["vx"] != before["vx"]) | (sg["vz"] != before["vz"])
        np.testing.assert_array_equal(moved_g, moved_o)
        moved += int(moved_o.sum())
    assert moved > 0


def test_membrane_area_volume_and_forces(ia):
    case = Case(seed=7, nv=642, radius=3.0, Lx=20, Ly=16, Lz=12)
    b = case.make_gpu()
    o = case.o
    # perturb the shape so every penalty is active
    rng = np.random.default_rng(9)
    v = o.vertices()
    for k in ("x", "y", "z"):
        v[k] = v[k] + 0.02 * rng.standard_normal(len(v[k]))
    o.set_vertices(x=v["x"], y=v["y"], z=v["z"])
    b.set_vertices(0, x=v["x"], y=v["y"], z=v["z"])
    o.totalva()
    o.pfdt()
    o.forces()
    o.external()
    b.membrane_forces()
    go, gg = o.globals(), b.globals()
    assert abs(gg["Sarea"] - go["Sarea"]) <= 1e-12 * go["Sarea"]
    assert abs(gg["Volume"] - go["Volume"]) <= 1e-12 * go["Volume"]
    assert gg["Sarea0"] == go["Sarea0"] and gg["Volume0"] == go["Volume0"]
    vo, vg = o.vertices(), b.vertices()
    for k in ("fx", "fy", "fz"):
        scale = np.abs(vo[k]).max()
        np.testing.assert_allclose(vg[k], vo[k], rtol=1e-10, atol=1e-12 * scale)
    np.testing.assert_allclose(b.parasite()[0], o.parasite()[0], rtol=1e-12, atol=1e-12)


def test_bonds_bind_unbind_bit_exact(ia):
    case = Case(seed=8, kon=40.0, koff=40.0, z_low=0.1)
    b = case.make_gpu()
    o = case.o
    _collide_both(case, b)  # the knobs' collision cells select their ligand lists
    for step in range(12):
        o.bond_prep()
        o.seed_mt(100 + step, 200 + step)
        o.bond_dynamics()
        u = o.get_bond_tape()
        b.set_bond_tape(0, u)
        b.bond_step()
        vo, vg = o.vertices(), b.vertices()
        np.testing.assert_array_equal(vg["bondi"], vo["bondi"])
        np.testing.assert_array_equal(vg["bondl"], vo["bondl"])
        np.testing.assert_array_equal(b.ligands(), o.ligands())
        assert b.globals()["bond_draws"] == len(u)
        co, lo = o.pairs()
        cg, lg = b.pairs()
        np.testing.assert_array_equal(cg, co)
        np.testing.assert_array_equal(lg, lo)
    assert vo["bondi"].sum() > 0


@pytest.mark.gpu
def test_binned_classification_equals_every_triangle_at_full_size(ia):
    """The binned inside/outside flags equal the all-triangles loop on the
    64x64x48 replica and on a 2562-vertex stretched membrane (the oracle
    pins the loop on a small box)."""
    import ensemble as en
    for d in (en.single(), en.replica(48, 48, 40, 8, 2562, None, 1.2, 1500, seed=4242)):
        b = ia.Batch([en.spec(d)])
        try:
            en.load(b, 0, d)
            b.classify()
            binned = b.solvent(0)["indio"]
            b.set_option(ia.OPT_CLASSIFY_ALL, 1)
            b.classify()
            every = b.solvent(0)["indio"]
        finally:
            b.close()
        assert np.array_equal(binned, every)
        assert 0 < int((every == 0).sum()) < len(every)  # some inside, most outside
