"""Host side of the drop-in (no GPU needed): the README run's set-up drawn
by the C++ host library, the reference-named entry points it exports, and
the command-line driver's argument handling.

The set-up is checked against the oracle's restatement of maincode.cpp:
76-446, which the golden fixtures pin to the reference bit for bit, so the
host library must produce exactly the same knobs, placed vertices, vertex
and parasite velocities, rest area/volume and solvent (as a multiset; the
oracle reorders particles when it classifies them).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import irbc_amd as ia
import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
PKG = os.path.join(os.path.dirname(HERE), "irbc-code_amd")
HOST = os.path.join(PKG, "libirbc_host.so")
MAIN = os.path.join(PKG, "irbc_main")
MESH = [os.path.join(GOLD, f) for f in ("TrophV600.dat", "TrophE600.dat", "TrophT600.dat")]

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)


def _host():
    if not os.path.exists(HOST):
        pytest.fail("libirbc_host.so not built: run python irbc-code_amd/build.py")
    L = C.CDLL(HOST)
    L.irbc_host_readme_scenario.argtypes = [C.POINTER(ia.Params), C.c_char_p, C.c_char_p, C.c_char_p,
                                            C.c_int64, C.c_int, C.c_uint32, C.c_int, _d, _d, _d, _i, _d, _d]
    return L


def _scenario(seed, ns, nknobs, kon=0.005, koff=0.015):
    L = _host()
    p = ia.Params.default(onrate=kon, offrate=koff)
    nv = 600
    sol = np.zeros(6 * ns)
    xyz, vel, par, st = np.zeros(3 * nv), np.zeros(3 * nv), np.zeros(12), np.zeros(4)
    knob = np.zeros(nv, np.int32)
    r = L.irbc_host_readme_scenario(C.byref(p), *[m.encode() for m in MESH], ns, nknobs, seed, nv,
                                    sol.ctypes.data_as(_d), xyz.ctypes.data_as(_d), vel.ctypes.data_as(_d),
                                    knob.ctypes.data_as(_i), par.ctypes.data_as(_d), st.ctypes.data_as(_d))
    assert r == nv
    return dict(sol=sol.reshape(6, ns), xyz=xyz.reshape(nv, 3), vel=vel.reshape(3, nv), knob=knob, para=par,
                stats=st)


@pytest.mark.parametrize("seed,nknobs", [(12345, 500), (777, 40)])
def test_readme_setup_matches_oracle(seed, nknobs):
    ns = 20000
    h = _scenario(seed, ns, nknobs)
    v, e, t = po.read_mesh(*MESH)
    o = po.Oracle()
    st = o.init_readme(seed, ns, v, e, t, nknobs)
    vo = o.vertices()
    np.testing.assert_array_equal(h["knob"], vo["knob"])
    np.testing.assert_array_equal(h["xyz"], np.stack([vo["x"], vo["y"], vo["z"]], 1))
    np.testing.assert_array_equal(h["vel"], np.stack([vo["vx"], vo["vy"], vo["vz"]]))
    np.testing.assert_array_equal(h["para"][:6], o.parasite()[0][:6])
    np.testing.assert_array_equal(h["stats"], st[:4])
    so = o.solvent()
    ours = h["sol"]
    theirs = np.stack([so[k] for k in ("x", "y", "z", "vx", "vy", "vz")])
    a = ours[:, np.lexsort(ours[::-1])]
    b = theirs[:, np.lexsort(theirs[::-1])]
    np.testing.assert_array_equal(a, b)
    assert h["knob"].sum() == nknobs


def test_reference_names_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", HOST], capture_output=True, text=True, check=True).stdout
    for name in ("_Z13CollisionStepv", "_Z11llistupdatev", "_Z20sortsolvent_initial1v", "_Z7TotalVAv",
                 "_Z18force_calculate_dtv", "_Z4pfdtv", "_Z13bond_dynamicsv", "_Z8TotalVA0v",
                 "_Z12centerofmassv", "_Z6bbounddd", "irbc_ref_bind", "irbc_host_readme_scenario",
                 "Sarea0", "center"):
        assert name in out, name


def test_bbound_wraps_by_whole_box_lengths():
    """bbound (Src/solfuncs.cpp:603-622): [0, lx] by repeated whole-length
    steps, the ends kept as they are."""
    L = C.CDLL(HOST)
    f = L._Z6bbounddd
    f.argtypes = [C.c_double, C.c_double]
    f.restype = C.c_double
    lx = 60.0
    assert f(lx, lx) == lx and f(0.0, lx) == 0.0 and f(-0.0, lx) == 0.0
    for x in np.random.default_rng(3).uniform(-3 * lx, 4 * lx, 200):
        want = x
        while want > lx:
            want -= lx
        while want < 0:
            want += lx
        assert f(x, lx) == want
        assert 0.0 <= f(x, lx) <= lx


def test_driver_argument_errors():
    if not os.path.exists(MAIN):
        pytest.fail("irbc_main not built")
    r = subprocess.run([MAIN], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr
    r = subprocess.run([MAIN, "0.005", "0.015", "/nonexistent", MESH[1], MESH[2], "500", "/dev/null", "0"],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = subprocess.run([MAIN, "0.005", "0.015"] + MESH + ["601", "/dev/null", "0"], capture_output=True, text=True)
    assert r.returncode == 2 and "exceed" in r.stderr
