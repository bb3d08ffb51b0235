"""Matching test cases for the CPU oracle and the HIP path.

A case is a small box with seeded solvent (FP32-representable values, so
the oracle reads exactly what the device stores), an icosphere membrane
placed against the bottom wall, a parasite bead, knobs and forces.  The
oracle side is built without the device; make_gpu() mirrors it on the GPU.
"""
import numpy as np

import pyoracle as po
import scenario as sc


def params_kw(Lx=16, Ly=12, Lz=10, walls=1.0, ramp_steps=0, kon=0.005, koff=0.015, navg=10.0):
    return dict(Lx=Lx, Ly=Ly, Lz=Lz, walls=walls, ramp_steps=ramp_steps, onrate=kon, offrate=koff,
                navg=navg)


class Case:
    def __init__(self, seed=1, nv=162, radius=2.6, z_low=0.3, nknobs=None, density=10, **kw):
        self.kw = params_kw(**kw)
        self.po_params = po.Params.default(**self.kw)
        p = self.po_params
        V, E, T = sc.mesh_for(nv)
        V = sc.place(V * radius, p.Lx, p.Ly, z_lowest=z_low)
        self.V, self.E, self.T = V, E, T
        self.ns = int(density * p.Lx * p.Ly * p.Lz)
        self.sol = sc.solvent(self.ns, p.Lx, p.Ly, p.Lz, seed)
        self.vel = sc.vertex_velocities(nv, seed + 1)
        self.knob = sc.knobs_uniform(nv, nv if nknobs is None else nknobs, seed + 2)
        self.para_d, self.para_i = sc.parasite_state(V, seed + 3)
        o = po.Oracle(self.po_params)
        o.set_mesh(V, E, T)
        o.set_solvent(*self.sol)
        o.classify()
        self.indio = o.solvent()["indio"]
        o.set_vertices(vx=self.vel[:, 0], vy=self.vel[:, 1], vz=self.vel[:, 2], knob=self.knob)
        o.set_parasite(self.para_d, self.para_i)
        # start-of-loop forces, as maincode.cpp:385-387 leaves them
        o.totalva()
        o.pfdt()
        o.forces()
        self.o = o
        self.seed = seed

    def make_gpu(self):
        import irbc_amd as ia
        gp = ia.Params.default(**self.kw)
        spec = ia.make_spec(gp, self.ns, len(self.V), len(self.E), len(self.T), seed=self.seed)
        b = ia.Batch([spec])
        b.set_mesh(0, self.V, self.E, self.T)
        self.load_gpu(b, 0)
        return b

    def load_gpu(self, b, r):
        """Copy the oracle's whole state into replica r of a batch."""
        o = self.o
        s = o.solvent()
        b.set_solvent(r, s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
        v = o.vertices()
        b.set_vertices(r, **{k: v[k] for k in b.VD + b.VI})
        d, i = o.parasite()
        b.set_parasite(r, d, i)
        b.set_ligands(r, o.ligands())
        b.set_wallspeed(r, o.globals()["wallspeed"])


def load_oracle_from_gpu(o, b, r=0):
    """Copy a device replica's state into the oracle (keys are derived state)."""
    s = b.solvent(r)
    o.set_solvent(s["x"], s["y"], s["z"], s["vx"], s["vy"], s["vz"], s["indio"])
    v = b.vertices(r)
    o.set_vertices(**{k: v[k] for k in po.Oracle.VD + po.Oracle.VI})
    d, i = b.parasite(r)
    o.set_parasite(d, i)
    o.set_ligands(b.ligands(r))


def periodic_diff(a, b, L):
    d = a - b
    return d - L * np.round(d / L)


def f32_ulps(a, b, scale_floor=1.0):
    """|a-b| in units of the FP32 spacing at max(|a|, floor)."""
    m = np.maximum(np.abs(a), scale_floor)
    return np.abs(a - b) / (m * 2.0 ** -23)
