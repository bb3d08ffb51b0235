"""The CPU oracle against golden fixtures written from the reference itself.

tests/golden/make_golden.py ran the unmodified reference main()
(maincode.cpp:56-770, one OpenMP thread, seed fixed through time()) and
recorded the full state at step N just before TotalVA() and the forces just
after force_calculate_dt().  The oracle must reproduce every array bit for
bit: solvent positions, velocities, cell keys and inside flags (960 000
particles), the 98 400 cells of the last collision (vcm, nc, ncc, dvv, ThE,
rand1, rand2), the cell->triangle lists with their stale tails, the
membrane, the parasite bead, bonds and ligand occupancy.
"""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def _digest(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _meta():
    with open(os.path.join(GOLD, "readme_runs.json")) as f:
        return json.load(f)


def _run_oracle(po, run):
    v, e, t = po.read_mesh(*[os.path.join(GOLD, f) for f in
                             ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat")])
    o = po.Oracle(onrate=run["kon"], offrate=run["koff"])
    stats = o.init_readme(run["seed"], 960000, v, e, t, run["knobs"])
    for _ in range(run["step"]):
        o.md_step()
    # step N up to TotalVA() (maincode.cpp:472-598)
    step, cs, upd = o.counters()
    cs, upd = cs + 1, upd + 1
    if cs == o.p.csteps:
        o.collision()
        cs = 0
    o.kick1()
    if upd == o.p.uplist:
        o.llistupdate()
        upd = 0
    o.set_counters(step, cs, upd)
    o.stream(o.p.ramp_speed)
    o.drift()
    return o, stats


def _state(o):
    s = o.solvent()
    vt = o.vertices()
    c = o.cells()
    pd, pi = o.parasite()
    off, ent = o.llist()
    g = o.globals()
    out = {"sx": s["x"], "sy": s["y"], "sz": s["z"], "svx": s["vx"], "svy": s["vy"],
           "svz": s["vz"], "spid": s["pid"], "sindio": s["indio"],
           "mx": vt["x"], "my": vt["y"], "mz": vt["z"], "mvx": vt["vx"], "mvy": vt["vy"],
           "mvz": vt["vz"], "fx": vt["fx"], "fy": vt["fy"], "fz": vt["fz"], "poff": vt["poff"],
           "pon": vt["pon"], "ix": vt["ix"], "iy": vt["iy"], "mpid": vt["pid"],
           "bondi": vt["bondi"], "bondl": vt["bondl"], "knob": vt["knob"],
           "para": pd, "parai": pi, "ligi": o.ligands(), "vcm": c["vcm"], "nc": c["nc"],
           "ncc": c["ncc"], "dvv": c["dvv"], "the": c["the"], "rand1": c["rand1"],
           "rand2": c["rand2"], "lloff": off.astype(np.int32), "llent": ent,
           "glob": np.array([g["Sarea"], g["Volume"], g["Sarea0"], g["Volume0"], g["theta0"],
                             g["wallspeed"], g["nsolb"]])}
    return out


@pytest.mark.parametrize("name", ["readme", "fastbonds"])
def test_oracle_matches_reference_bit_for_bit(oracle_lib, name):
    po = oracle_lib
    po.set_threads(os.cpu_count() or 1)
    run = _meta()["runs"][name]
    gold = np.load(os.path.join(GOLD, "readme_runs.npz"))
    o, stats = _run_oracle(po, run)
    # stdout known answers of the init (maincode.cpp:256-257,345-346,366-368,446)
    out = run["stdout"]
    assert "Mean edge is:\t%g" % stats[0] in out
    assert "Standard deviation is:\t%g" % stats[1] in out
    assert "Sarea:\t%g" % stats[2] in out and "Volume:\t%g" % stats[3] in out
    assert "Young's modulus::\t%g" % stats[5] in out
    assert "Poisson's ratio::\t%g" % stats[6] in out
    assert "no of vertices selected:\t%d" % int(stats[7]) in out

    pre = _state(o)
    bad = []
    for k, a in pre.items():
        want = run["hash"]["pre_va/" + k]
        if _digest(a) != want:
            key = "%s/pre_va/%s" % (name, k)
            detail = ""
            if key in gold:
                detail = "first diff at %s" % np.flatnonzero(gold[key] != a)[:3]
            elif key + "@sample" in gold:
                st = run["sample_stride_pre_va_%s" % k]
                detail = "sample diffs %d" % np.count_nonzero(gold[key + "@sample"] != a[::st])
            bad.append("%s %s" % (k, detail))
    assert not bad, "pre_va mismatches: %s" % bad
    # forces of step N (pfdt + force_calculate_dt on top of kick1's zeroing)
    o.totalva()
    o.pfdt()
    o.forces()
    vt = o.vertices()
    for k in ("fx", "fy", "fz"):
        np.testing.assert_array_equal(vt[k], gold["%s/post_forces/%s" % (name, k)])
    pd, _ = o.parasite()
    np.testing.assert_array_equal(pd, gold["%s/post_forces/para" % name])
    g = o.globals()
    want = gold["%s/post_forces/glob" % name]
    assert g["Sarea"] == want[0] and g["Volume"] == want[1]
    if name == "fastbonds":
        # the high-rate run must actually exercise binding and unbinding
        assert pre["bondi"].sum() > 0
