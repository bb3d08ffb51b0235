"""Free-running comparisons of the HIP path with the oracle.

Both sides start from one state and run N whole MD steps on the same
counter-based random streams (the oracle in its Philox mode, keyed by the
replica's seed as the device is), with the oracle storing the solvent in
FP32 exactly where the device stores it.  Nothing is copied between the two
after the start, so the comparison covers everything a trajectory carries
from step to step: the near masks kept or rebuilt at collisions, the drift
margins, the parity-buffered kinematics, list updates, the grid shifts and
the bond history.  What remains between the two is FP64 rounding (exact
fixed-point cell sums against sequential FP64 sums, the device's libm
against glibc in the transcendental functions), which rarely moves an FP32
value by one unit.
"""
import os

import numpy as np

import irbc_amd as ia
import pyoracle as po

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def oracle_params(p):
    """The oracle's parameter block with the device's values."""
    q = po.Params()
    for f, _ in ia.Params._fields_:
        setattr(q, f, getattr(p, f))
    return q


def device_params(q):
    p = ia.Params()
    for f, _ in po.Params._fields_:
        setattr(p, f, getattr(q, f))
    return p


def oracle_from_device(b, r, p, V, E, T):
    """An oracle holding replica r's state as the device holds it (the
    solvent as stored), its own cell->triangle lists of that state, and the
    same counters."""
    o = po.Oracle(oracle_params(p))
    o.set_mesh(V, E, T)
    s = b.solvent(r)
    o.set_solvent(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
    v = b.vertices(r)
    o.set_vertices(**{k: v[k] for k in po.Oracle.VD + po.Oracle.VI})
    d, i = b.parasite(r)
    o.set_parasite(d, i)
    o.set_ligands(b.ligands(r))
    o.llistupdate()
    o.set_counters(*b.counters(r))
    return o


def device_from_oracle(o, V, E, T, seed):
    """A one-replica batch holding the oracle's state (solvent rounded to
    FP32 on upload, as the oracle's FP32 mode has it), lists built on the
    device from that state."""
    p = device_params(o.p)
    s = o.solvent()
    spec = ia.make_spec(p, len(s["x"]), len(V), len(E), len(T), seed=seed)
    b = ia.Batch([spec])
    b.set_mesh(0, V, E, T)
    b.set_solvent(0, s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
    v = o.vertices()
    b.set_vertices(0, **{k: v[k] for k in b.VD + b.VI})
    d, i = o.parasite()
    b.set_parasite(0, d, i)
    b.set_ligands(0, o.ligands())
    b.set_counters(0, *o.counters())
    b.llist_update()
    return b


def readme_state(seed=12345, kon=0.005, koff=0.015, knobs=500, ns=960000):
    """The README run's initial state (maincode.cpp:56-446) from the oracle,
    which reproduces the reference's own fixtures bit for bit."""
    V, E, T = po.read_mesh(*[os.path.join(GOLD, f) for f in ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat")])
    o = po.Oracle(onrate=kon, offrate=koff)
    o.init_readme(seed, ns, V, E, T, knobs)
    # init_readme keeps the raw file coordinates; the mesh in the state is
    # the placed one
    v = o.vertices()
    Vp = np.stack([v["x"] + v["ix"] * o.p.Lx, v["y"] + v["iy"] * o.p.Ly, v["z"]], axis=1)
    return o, Vp, E, T


def run(b, o, n, chunks=(None,), each=None):
    """n steps on both sides; the device's in the given call sizes (None:
    one call for the rest); each(k) is called after the oracle's k-th step."""
    done = 0
    for c in chunks:
        k = n - done if c is None else min(c, n - done)
        if k > 0:
            b.md_steps(k)
            done += k
    if done < n:
        b.md_steps(n - done)
    b.sync()
    for k in range(n):
        o.md_step()
        if each:
            each(k + 1)


def compare(b, r, o, cells=False):
    """Differences between replica r of the device and the oracle; with
    cells (the device ran with diagnostics on), also the per-cell results of
    the last collision."""
    out = {}
    if cells:
        cg, co = b.cells(r), o.cells()
        out["nc_diff"] = int(np.count_nonzero(cg["nc"] != co["nc"]))
        scale = max(1e-300, np.abs(co["vcm"]).max())
        out["vcm_rel"] = float(np.abs(cg["vcm"] - co["vcm"]).max() / scale)
    sg, so = b.solvent(r), o.solvent()
    L = (o.p.Lx, o.p.Ly)
    out["n"] = len(so["x"])
    out["key_diff"] = int(np.count_nonzero(sg["key"] != so["pid"]))
    out["indio_diff"] = int(np.count_nonzero(sg["indio"] != so["indio"]))
    dx = np.abs(sg["x"] - so["x"])
    dx = np.minimum(dx, L[0] - dx)
    dy = np.abs(sg["y"] - so["y"])
    dy = np.minimum(dy, L[1] - dy)
    dz = np.abs(sg["z"] - so["z"])
    dv = np.max([np.abs(sg[k] - so[k]) for k in ("vx", "vy", "vz")], axis=0)
    moved = (dx > 0) | (dy > 0) | (dz > 0) | (dv > 0)
    out["solvent_diff"] = int(np.count_nonzero(moved))
    out["solvent_pos_max"] = float(max(dx.max(), dy.max(), dz.max()))
    out["solvent_vel_max"] = float(dv.max())
    vg, vo = b.vertices(r), o.vertices()
    ux = lambda v: v["x"] + v["ix"] * L[0]  # noqa: E731
    uy = lambda v: v["y"] + v["iy"] * L[1]  # noqa: E731
    out["mem_pos_max"] = float(max(np.abs(ux(vg) - ux(vo)).max(), np.abs(uy(vg) - uy(vo)).max(),
                                   np.abs(vg["z"] - vo["z"]).max()))
    vscale = max(1e-300, max(np.abs(vo[k]).max() for k in ("vx", "vy", "vz")))
    out["mem_vel_rel"] = float(max(np.abs(vg[k] - vo[k]).max() for k in ("vx", "vy", "vz")) / vscale)
    fscale = max(1e-300, max(np.abs(vo[k]).max() for k in ("fx", "fy", "fz")))
    out["mem_force_rel"] = float(max(np.abs(vg[k] - vo[k]).max() for k in ("fx", "fy", "fz")) / fscale)
    out["bonds_equal"] = bool(all(np.array_equal(vg[k], vo[k]) for k in ("bondi", "bondl", "knob")))
    out["images_equal"] = bool(np.array_equal(vg["ix"], vo["ix"]) and np.array_equal(vg["iy"], vo["iy"]))
    out["ligands_equal"] = bool(np.array_equal(b.ligands(r), o.ligands()))
    out["bonded"] = int(np.count_nonzero(vo["bondi"] == 1))
    gg, go = b.globals(r), o.globals()
    out["area_rel"] = abs(gg["Sarea"] - go["Sarea"]) / go["Sarea"]
    out["volume_rel"] = abs(gg["Volume"] - go["Volume"]) / go["Volume"]
    pg, pog = b.parasite(r), o.parasite()
    out["parasite_rel"] = float(np.abs(pg[0][:6] - pog[0][:6]).max() / max(1e-300, np.abs(pog[0][:6]).max()))
    out["counters_equal"] = tuple(b.counters(r)) == tuple(o.counters())
    return out


def series_record(o):
    """The time-series record of the oracle's state after a step: step,
    membrane centroid (unwrapped), centroid velocity, bonded knobs."""
    v = o.vertices()
    L = (o.p.Lx, o.p.Ly)
    return np.array([o.counters()[0], np.mean(v["x"] + v["ix"] * L[0]), np.mean(v["y"] + v["iy"] * L[1]),
                     np.mean(v["z"]), np.mean(v["vx"]), np.mean(v["vy"]), np.mean(v["vz"]),
                     np.count_nonzero(v["bondi"] == 1)])


def check(m, L, what):
    """The tolerances of a free-running comparison (FP32 solvent storage on
    both sides, FP64 arithmetic; the oracle's sequential sums against the
    device's exact fixed-point ones, glibc against the device's libm):
    exact keys, counts, flags, counters, images, bonds and ligands; at most
    1% of the particles off by at most 8 FP32 units at the box scale or
    1e-5 in velocity; the membrane within 1e-6 cells and 1e-5 of its
    velocity and force scales; area, volume and bead within 1e-8."""
    assert m["counters_equal"], what
    assert m["key_diff"] == 0, (what, m)
    assert m["indio_diff"] == 0, (what, m)
    if "nc_diff" in m:
        assert m["nc_diff"] == 0, (what, m)
        assert m["vcm_rel"] <= 1e-5, (what, m)
    assert m["solvent_diff"] <= 0.01 * m["n"], (what, m)
    assert m["solvent_pos_max"] <= 8 * np.spacing(np.float32(max(L))), (what, m)
    assert m["solvent_vel_max"] <= 1e-5, (what, m)
    assert m["mem_pos_max"] <= 1e-6, (what, m)
    assert m["mem_vel_rel"] <= 1e-5, (what, m)
    assert m["mem_force_rel"] <= 1e-5, (what, m)
    assert m["area_rel"] <= 1e-8 and m["volume_rel"] <= 1e-8, (what, m)
    assert m["parasite_rel"] <= 1e-8, (what, m)
    assert m["bonds_equal"] and m["images_equal"] and m["ligands_equal"], (what, m)
