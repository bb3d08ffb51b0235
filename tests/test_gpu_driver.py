"""The command-line driver (irbc_main) on the README run, its output files
against the reference's formats and against the oracle.

irbc_main takes the reference's arguments (README.md usage,
maincode.cpp:56-770) and writes what the reference writes:
  - the set-up lines on stdout (maincode.cpp:256-257,345-346,366-368,446),
    each of which must appear verbatim in the reference's own stdout for the
    same seed (tests/golden/readme_runs.json);
  - knobs.dat, "%lf\\t%lf\\t%lf" per knob vertex (maincode.cpp:436-446),
    compared as text with the oracle's README set-up, which reproduces the
    reference's knob choice bit for bit;
  - frames "#k" then "x y z ix iy bondi" per vertex after steps F+1, 2F+1,
    ... (maincode.cpp:474-477,746-761), here F = 5 over 11 steps, compared
    with the oracle free-running on the device's streams (positions within
    the 6 printed decimals, images and bonds exact);
  - rbcK.stl with stl_flag 1 (Src/memfuncs.cpp:590-626): ASCII STL, one
    facet per triangle with the unnormalised normal of Normaldt
    (Src/memfuncs.cpp:110-122) and the unwrapped vertex positions, checked
    against the frame written at the same step.
"""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import freerun as fr
import pyoracle as po

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
MAIN = os.path.join(os.path.dirname(HERE), "irbc-code_amd", "irbc_main")
MESH = [os.path.join(GOLD, f) for f in ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat")]

FRAME_LINE = re.compile(r"^-?\d+\.\d{6}\t-?\d+\.\d{6}\t-?\d+\.\d{6}\t-?\d+\t-?\d+\t[01]$")


def _frames(path, nv):
    with open(path) as f:
        lines = f.read().splitlines()
    out = []
    i = 0
    while i < len(lines):
        assert lines[i] == "#%d" % (len(out) + 1), lines[i]
        rows = lines[i + 1:i + 1 + nv]
        assert len(rows) == nv
        for r in rows:
            assert FRAME_LINE.match(r), r
        a = np.array([r.split("\t") for r in rows])
        out.append((a[:, :3].astype(float), a[:, 3:].astype(int)))
        i += 1 + nv
    return out


def _stl(path):
    with open(path) as f:
        lines = f.read().splitlines()
    assert lines[0] == "solid ascii" and lines[-1] == "endsolid"
    body = lines[1:-1]
    assert len(body) % 7 == 0
    normals, verts = [], []
    for k in range(0, len(body), 7):
        n = body[k].split("\t")
        assert n[0] == "facet normal" and len(n) == 4, body[k]
        assert body[k + 1] == "outer loop" and body[k + 5] == "endloop" and body[k + 6] == "endfacet"
        tri = []
        for j in range(3):
            v = body[k + 2 + j].split("\t")
            assert v[0] == "vertex" and len(v) == 4, body[k + 2 + j]
            tri.append([float(x) for x in v[1:]])
        normals.append([float(x) for x in n[1:]])
        verts.append(tri)
    return np.array(normals), np.array(verts)


def test_irbc_main_short_readme_run(tmp_path):
    if not os.path.exists(MAIN):
        pytest.fail("irbc_main not built")
    seed, steps, every = 12345, 11, 5
    cmd = [MAIN, "0.005", "0.015"] + MESH + ["500", "RBC.dat", "1", "--steps", str(steps), "--seed", str(seed),
                                            "--frame-every", str(every)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr

    with open(os.path.join(GOLD, "readme_runs.json")) as f:
        ref_out = json.load(f)["runs"]["readme"]["stdout"].splitlines()
    ours = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(ours) == 8
    for ln in ours:
        assert ln in ref_out, ln

    po.set_threads(16)
    o, V, E, T = fr.readme_state(seed, 0.005, 0.015)
    v0 = o.vertices()
    knob = np.flatnonzero(v0["knob"] == 1)
    want = ["%f\t%f\t%f" % (v0["x"][i], v0["y"][i], v0["z"][i]) for i in knob]
    with open(tmp_path / "knobs.dat") as f:
        assert f.read().splitlines() == want

    nv = len(V)
    frames = _frames(tmp_path / "RBC.dat", nv)
    assert len(frames) == 2
    o.set_fp32(True)
    o.use_philox(seed)
    tris = np.asarray(T).reshape(-1, 3)
    for k in range(1, steps + 1):
        o.md_step()
        if k % every != 1 or k == 1:
            continue
        pos, ints = frames[k // every - 1]
        v = o.vertices()
        np.testing.assert_array_equal(ints, np.stack([v["ix"], v["iy"], v["bondi"]], 1))
        np.testing.assert_allclose(pos, np.stack([v["x"], v["y"], v["z"]], 1), rtol=0, atol=2e-6)
        # the STL of the same step: unwrapped positions of each triangle
        # and the cross product of its edges, to 6 significant digits
        normals, verts = _stl(tmp_path / ("rbc%d.stl" % (k // every)))
        assert len(normals) == len(tris)
        u = pos + ints[:, :2].astype(float) @ np.array([[o.p.Lx, 0, 0], [0, o.p.Ly, 0]])
        P = u[tris]
        np.testing.assert_allclose(verts, P, rtol=1e-5, atol=1e-5)
        N = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
        np.testing.assert_allclose(normals, N, rtol=0, atol=1e-4 * np.abs(N).max())
