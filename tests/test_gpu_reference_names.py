"""The reference's own call surface on the GPU, through libirbc_host.so.

1. The README run (maincode.cpp:56-770: 60x40x40 at density 10, the
   Troph600 mesh, 500 knobs, kon 0.005, koff 0.015, seed 12345) against the
   state the reference itself recorded (tests/golden/readme_runs.*).  The
   device starts from the README initial state (built by the oracle, which
   reproduces the reference's fixtures bit for bit) and runs the first 90
   steps of the time loop on the reference's MT19937 / StochasticLib1
   streams, replayed from the oracle as tapes; nothing is copied from the
   oracle after the start.  Step 90 is then driven the way
   maincode.cpp:472-604 calls it, through the reference-named functions:
   CollisionStep() when due, the first half kick, llistupdate() when due,
   streamstep(), the drift, then TotalVA(), pfdt() and force_calculate_dt().
   The state is compared where the reference recorded it: just before
   TotalVA(), and just after force_calculate_dt().
2. pfdt() and force_calculate_dt() as two calls (Src/memfuncs.cpp:443-484,
   Include/memfuncs.h:27-28): after pfdt() the vertex forces hold only the
   bead-membrane repulsion and the bead holds its reaction, as in the
   reference; the two calls give the same bits as the fused force pass.

The device stores the solvent in FP32 where the reference stores FP64, and
over 90 steps a few particles in 10^4 round into another cell, which moves
the membrane vertices sharing those cells.  So part 1 runs the oracle twice,
as the reference and in FP32 storage, requires the device to equal the FP32
run to the free-running tolerances, and reports both runs' departure from
the reference's record.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import freerun as fr
import irbc_amd as ia
import pyoracle as po
from cases import Case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
HOST = os.path.join(os.path.dirname(HERE), "irbc-code_amd", "libirbc_host.so")

# the reference's zero-argument functions (C++ names, Itanium mangling)
REF = {"CollisionStep": "_Z13CollisionStepv", "llistupdate": "_Z11llistupdatev",
       "streamstep": "_Z10streamstepv", "TotalVA": "_Z7TotalVAv", "pfdt": "_Z4pfdtv",
       "force_calculate_dt": "_Z18force_calculate_dtv", "externalforces": "_Z14externalforcesv",
       "bond_dynamics": "_Z13bond_dynamicsv", "sortsolvent_initial1": "_Z20sortsolvent_initial1v",
       "TotalVA0": "_Z8TotalVA0v", "centerofmass": "_Z12centerofmassv"}


class RefNames:
    """libirbc_host.so bound to one replica of a Batch."""

    def __init__(self, b, r=0):
        ia.load()
        self.L = C.CDLL(HOST)
        # one copy of libirbc_amd.so in the process: the batch's handle is
        # valid in the functions libirbc_host.so calls
        with open("/proc/self/maps") as f:
            paths = {ln.split()[-1] for ln in f if ln.rstrip().endswith("libirbc_amd.so")}
        assert len(paths) == 1, paths
        self.L.irbc_ref_bind.argtypes = [C.c_void_p, C.c_int]
        self.L.irbc_ref_bind(b.h, r)
        for k, sym in REF.items():
            setattr(self, k, getattr(self.L, sym))

    @property
    def Sarea(self):
        return C.c_double.in_dll(self.L, "Sarea").value

    @property
    def Volume(self):
        return C.c_double.in_dll(self.L, "Volume").value

    @property
    def Sarea0(self):
        return C.c_double.in_dll(self.L, "Sarea0").value

    @property
    def Volume0(self):
        return C.c_double.in_dll(self.L, "Volume0").value

    @property
    def center(self):
        return np.array((C.c_double * 3).in_dll(self.L, "center"))


def _partial_step(o, p, dev=None, h=None, b=None):
    """Step 90 of the time loop up to TotalVA() (maincode.cpp:472-598) on an
    oracle, and with dev (the oracle whose draws both replay) on the device
    through the reference names."""
    step, cs, upd = o.counters()
    cs, upd = cs + 1, upd + 1
    if cs == p.csteps:
        if dev:
            tape = dev.get_collision_tape()
            o.set_collision_tape(*tape)
            b.set_collision_tape(0, *tape)
            h.CollisionStep()
        o.collision()
        cs = 0
    o.kick1()
    if dev:
        b.membrane_phases(ia.PH_KICK1)
    if upd == p.uplist:
        o.llistupdate()
        if dev:
            h.llistupdate()
        upd = 0
    o.set_counters(step, cs, upd)
    w = p.ramp_speed if step < p.ramp_steps else p.walls
    o.stream(w)
    if dev:
        b.set_counters(0, step, cs, upd)
        b.set_wallspeed(0, w)
        h.streamstep()
    o.drift()
    if dev:
        b.membrane_phases(ia.PH_DRIFT)


def _from_reference(s, v, nc, gold, meta, p):
    """Distances of a state from the reference's recorded pre-TotalVA state
    of step 90 (membrane whole, solvent and cells as sampled)."""
    g = lambda k: gold["readme/pre_va/" + k]  # noqa: E731
    L = (p.Lx, p.Ly)
    m = {}
    xg = np.stack([v["x"] + v["ix"] * L[0], v["y"] + v["iy"] * L[1], v["z"]])
    xr = np.stack([g("mx") + g("ix") * L[0], g("my") + g("iy") * L[1], g("mz")])
    dx = np.abs(xg - xr).max(axis=0)
    m["mem_pos_max"] = float(dx.max())
    m["mem_pos_p90"] = float(np.quantile(dx, 0.9))
    vr = np.stack([g(k) for k in ("mvx", "mvy", "mvz")])
    m["mem_vel_rel"] = float(np.abs(np.stack([v[k] for k in ("vx", "vy", "vz")]) - vr).max() / np.abs(vr).max())
    st = meta["sample_stride_pre_va_sx"]
    dpos = []
    for k, Lk in (("x", p.Lx), ("y", p.Ly), ("z", None)):
        d = np.abs(s[k][::st] - g("s" + k + "@sample"))
        dpos.append(np.minimum(d, Lk - d) if Lk else d)
    dpos = np.max(dpos, axis=0)
    dvel = np.max([np.abs(s[k][::st] - g("s" + k + "@sample")) for k in ("vx", "vy", "vz")], axis=0)
    m["solvent_sampled"] = len(dpos)
    m["solvent_far"] = int(np.count_nonzero((dpos > 1e-4) | (dvel > 1e-4)))
    m["solvent_pos_med"] = float(np.median(dpos))
    m["key_diff"] = int(np.count_nonzero(s["key"][::st] != g("spid@sample")))
    m["indio_diff"] = int(np.count_nonzero(s["indio"][::st] != g("sindio@sample")))
    m["nc_diff"] = int(np.count_nonzero(nc[::meta["sample_stride_pre_va_nc"]] != g("nc@sample")))
    return m


def test_readme_run_through_reference_names_against_reference_fixtures():
    """Three runs from the README initial state on the reference's streams:
    A, the oracle as the reference (FP64 storage), whose state at step 90 is
    the reference's recorded one; B, the oracle storing the solvent in FP32
    as the device does; the device.  B and the device replay A's draws as
    tapes, addressed by cell (the stream itself skips wall cells that are
    full, so a run that drew on its own would lose step with A at the first
    cell it fills differently).  The device must equal B to the
    free-running tolerances; B, and so the
    device, departs from the reference only through FP32 storage: a few
    particles in 10^4 round into another cell, and the cells they leave and
    join move the membrane vertices in them."""
    po.set_threads(16)
    seed, n = 12345, 90
    gold = np.load(os.path.join(GOLD, "readme_runs.npz"))
    with open(os.path.join(GOLD, "readme_runs.json")) as f:
        meta = json.load(f)["runs"]["readme"]
    A, V, E, T = fr.readme_state(seed, 0.005, 0.015)
    B, _, _, _ = fr.readme_state(seed, 0.005, 0.015)
    B.set_fp32(True)
    B.use_tapes(True)
    b = fr.device_from_oracle(B, V, E, T, seed)
    try:
        h = RefNames(b)
        b.use_tapes(0, True)
        b.set_diagnostics(True)
        p = A.p
        for _ in range(n):
            _, cs, _ = A.counters()
            collides = cs + 1 == p.csteps
            A.md_step()
            if collides:
                tape = A.get_collision_tape()
                B.set_collision_tape(*tape)
                b.set_collision_tape(0, *tape)
            bt = A.get_bond_tape()
            B.set_bond_tape(bt)
            B.md_step()
            b.set_bond_tape(0, bt)
            b.md_steps(1)
        _partial_step(A, p)
        _partial_step(B, p, dev=A, h=h, b=b)
        # A is the reference at this point (pinned on the CPU bit for bit)
        np.testing.assert_array_equal(A.vertices()["x"], gold["readme/pre_va/mx"])
        mB = fr.compare(b, 0, B, cells=True)
        print("device against B (FP32 storage):", mB)
        fr.check(mB, (p.Lx, p.Ly, p.Lz), "README step 90, device against B")
        dev = _from_reference(b.solvent(), b.vertices(), b.cells()["nc"], gold, meta, p)
        sb = B.solvent()
        sb["key"] = sb["pid"]
        orc = _from_reference(sb, B.vertices(), B.cells()["nc"], gold, meta, p)
        print("device from the reference:", dev)
        print("B from the reference:     ", orc)
        # the cell->triangle lists as the oracle holds them: each cell's live
        # entries, in order; the reference's per-cell vectors keep a stale
        # tail beyond the last update's count (Src/solfuncs.cpp:353-360),
        # which the oracle reproduces and the device's compacted lists drop
        off, ent = b.llist()
        oo, oe = B.llist()
        live = np.diff(off)
        assert (live <= np.diff(oo)).all()
        pos = np.repeat(oo[:-1], live) + (np.arange(len(ent)) - np.repeat(off[:-1], live))
        np.testing.assert_array_equal(ent, oe[pos])

        # TotalVA(), pfdt(), force_calculate_dt() (maincode.cpp:601-604)
        h.TotalVA()
        for o in (A, B):
            o.totalva()
            o.pfdt()
            o.forces()
        gB = B.globals()
        assert abs(h.Sarea - gB["Sarea"]) <= 1e-8 * gB["Sarea"]
        assert abs(h.Volume - gB["Volume"]) <= 1e-8 * gB["Volume"]
        gl = gold["readme/post_forces/glob"]
        dev["area_rel"] = abs(h.Sarea - gl[0]) / gl[0]
        dev["volume_rel"] = abs(h.Volume - gl[1]) / gl[1]
        h.pfdt()
        h.force_calculate_dt()
        vg, vb = b.vertices(), B.vertices()
        fref = np.stack([gold["readme/post_forces/" + k] for k in ("fx", "fy", "fz")])
        np.testing.assert_array_equal(np.stack([A.vertices()[k] for k in ("fx", "fy", "fz")]), fref)
        fB = np.stack([vb[k] for k in ("fx", "fy", "fz")])
        fD = np.stack([vg[k] for k in ("fx", "fy", "fz")])
        assert np.abs(fD - fB).max() <= 1e-5 * np.abs(fB).max()
        dev["mem_force_rel"] = float(np.abs(fD - fref).max() / np.abs(fref).max())
        orc["mem_force_rel"] = float(np.abs(fB - fref).max() / np.abs(fref).max())
        np.testing.assert_allclose(b.parasite()[0], B.parasite()[0], rtol=1e-8, atol=1e-12)
    finally:
        b.close()
    print("device from the reference after the forces:", dev, "B:", orc)
    # the device's departure from the reference is B's (the triangle
    # inequality with the checks against B above)
    assert dev["key_diff"] == orc["key_diff"] and dev["nc_diff"] == orc["nc_diff"], (dev, orc)
    assert dev["indio_diff"] == 0 == orc["indio_diff"]
    assert dev["mem_pos_max"] <= orc["mem_pos_max"] + 1e-6, (dev, orc)
    assert dev["mem_vel_rel"] <= orc["mem_vel_rel"] + 1e-5, (dev, orc)
    assert dev["mem_force_rel"] <= orc["mem_force_rel"] + 1e-5, (dev, orc)
    # and that departure is small where it is not chaotic: most vertices,
    # the sampled solvent, area and volume
    assert dev["mem_pos_p90"] <= 1e-4 and dev["solvent_pos_med"] <= 1e-4, dev
    assert dev["key_diff"] <= 1e-3 * dev["solvent_sampled"], dev
    assert dev["area_rel"] <= 1e-6 and dev["volume_rel"] <= 1e-6, dev


def test_pfdt_and_force_calculate_dt_as_two_calls():
    case = Case(seed=11, nv=642, radius=3.0, Lx=20, Ly=16, Lz=12)
    o = case.o
    v = o.vertices()
    # the bead just inside the membrane under vertex 0, in the cutoff of a
    # few vertices; forces cleared as the first half kick leaves them
    c = np.array([v["x"].mean(), v["y"].mean(), v["z"].mean()])
    x0 = np.array([v["x"][0], v["y"][0], v["z"][0]])
    u = (c - x0) / np.linalg.norm(c - x0)
    d, i = o.parasite()
    d = d.copy()
    d[:3] = x0 + 0.8 * o.p.paraR * u
    d[6:] = 0.0
    z = np.zeros(len(v["x"]))
    o.set_parasite(d, i)
    o.set_vertices(fx=z, fy=z, fz=z)
    b = case.make_gpu()
    b2 = case.make_gpu()
    try:
        h = RefNames(b)
        o.totalva()
        h.TotalVA()
        go = o.globals()
        assert abs(h.Sarea - go["Sarea"]) <= 1e-12 * go["Sarea"]
        assert abs(h.Volume - go["Volume"]) <= 1e-12 * go["Volume"]
        # TotalVA0() of the same shape, and centerofmass()
        h.TotalVA0()
        assert (h.Sarea0, h.Volume0) == (h.Sarea, h.Volume)
        o.centerofmass()
        h.centerofmass()
        go = o.globals()
        np.testing.assert_allclose(h.center, [go["cx"], go["cy"], go["cz"]], rtol=1e-12, atol=1e-12)
        o.pfdt()
        h.pfdt()
        vo, vg = o.vertices(), b.vertices()
        hit = np.flatnonzero(vo["fx"] ** 2 + vo["fy"] ** 2 + vo["fz"] ** 2)
        assert 0 < len(hit) < 100, len(hit)
        for k in ("fx", "fy", "fz"):
            np.testing.assert_allclose(vg[k], vo[k], rtol=1e-12, atol=0)
        np.testing.assert_allclose(b.parasite()[0], o.parasite()[0], rtol=1e-12, atol=0)
        assert np.abs(b.parasite()[0][9:]).max() > 0  # the bead's reaction
        o.forces()
        h.force_calculate_dt()
        vo, vg = o.vertices(), b.vertices()
        for k in ("fx", "fy", "fz"):
            np.testing.assert_allclose(vg[k], vo[k], rtol=1e-10, atol=1e-12 * np.abs(vo[k]).max())
        # the fused pass (IRBC_PH_VA | IRBC_PH_FORCES) gives the same bits
        b2.membrane_phases(ia.PH_VA | ia.PH_FORCES)
        v2 = b2.vertices()
        for k in ("fx", "fy", "fz"):
            np.testing.assert_array_equal(v2[k], vg[k])
        np.testing.assert_array_equal(b2.parasite()[0], b.parasite()[0])
        # force_calculate_dt() alone leaves the bead out
        b2.set_vertices(0, fx=z, fy=z, fz=z)
        b2.membrane_phases(ia.PH_MEMBRANE)
        v3 = b2.vertices()
        rest = np.setdiff1d(np.arange(len(z)), hit)
        for k in ("fx", "fy", "fz"):
            np.testing.assert_array_equal(v3[k][rest], vg[k][rest])
        assert max(np.abs(v3[k][hit] - vg[k][hit]).max() for k in ("fx", "fy", "fz")) > 0
        with pytest.raises(ia.IrbcError):
            b2.membrane_phases(1 << 12)
    finally:
        b.close()
        b2.close()
