"""Replica ensembles: one large replica, uniform and uneven batches.

A replica is described by a plain dict (box, density, mesh size, radius,
aspect, knobs, rates, shear, seed); `spec()` turns it into an
irbc_replica_spec and `load()` puts its initial state into a slot of a
Batch.  The same description drives the benchmark, the tests and the
request loop, so a replica is the same object alone and in a mix.

Workloads:
  single()   64x64x48 at density 10, 642-vertex icosphere of radius 8,
             500 knobs, kon 0.005, koff 0.015, shear rate 0.05
  uniform()  64 replicas of 32x32x24 at density 10, 642-vertex icosphere of
             radius 6, kon/koff log-uniform in [1e-3, 1e-1], 200-800 knobs
  uneven()   48 replicas, boxes 24^3 .. 64x64x48 at density 5-10, meshes of
             162/642/2562 vertices stretched to aspect 1.0-1.4, 100-2000
             knobs (at most one per vertex), shear rate 0.01-0.1
"""
import numpy as np

import irbc_amd as ia
import scenario as sc

# radius at which an icosphere's mean edge is about one cell
NATURAL_RADIUS = {42: 2.0, 162: 4.0, 642: 8.0, 2562: 16.0}


def replica(Lx, Ly, Lz, density=10, nv=642, radius=None, aspect=1.0, nknobs=500, kon=0.005,
            koff=0.015, shear=0.05, seed=1, place="centre"):
    if radius is None:
        radius = NATURAL_RADIUS[nv]
    # keep the cell inside the box: the stretched axis along x, the others
    # clear of both walls
    ax = aspect ** (2.0 / 3.0)
    az = aspect ** (-1.0 / 3.0)
    radius = min(radius, 0.45 * Lx / ax, 0.45 * Ly / az, 0.4 * Lz / az)
    return dict(Lx=int(Lx), Ly=int(Ly), Lz=int(Lz), density=int(density), nv=int(nv),
                radius=float(radius), aspect=float(aspect), nknobs=int(min(nknobs, nv)),
                kon=float(kon), koff=float(koff), shear=float(shear), seed=int(seed), place=place)


def single(seed=12345):
    return replica(64, 64, 48, 10, 642, 8.0, 1.0, 500, 0.005, 0.015, 0.05, seed, place="reference")


def uniform(n=64, seed0=1000):
    out = []
    for r in range(n):
        g = np.random.default_rng(seed0 + r)
        kon, koff = 10.0 ** g.uniform(-3, -1, size=2)
        out.append(replica(32, 32, 24, 10, 642, 6.0, 1.0, int(g.integers(200, 801)), kon, koff,
                           0.05, seed0 + r))
    return out


def uneven(n=48, seed0=2000):
    out = []
    for r in range(n):
        g = np.random.default_rng(seed0 + r)
        Lx, Ly = (int(v) for v in g.integers(24, 65, size=2))
        Lz = int(g.integers(24, 49))
        nv = int(g.choice([162, 642, 2562]))
        kon, koff = 10.0 ** g.uniform(-3, -1, size=2)
        out.append(replica(Lx, Ly, Lz, int(g.integers(5, 11)), nv, None, float(g.uniform(1.0, 1.4)),
                           int(g.integers(100, 2001)), kon, koff, float(g.uniform(0.01, 0.1)),
                           seed0 + r))
    return out


def params(d):
    return ia.Params.default(Lx=d["Lx"], Ly=d["Ly"], Lz=d["Lz"], navg=float(d["density"]),
                             walls=d["shear"] * d["Lz"], ramp_steps=0, onrate=d["kon"],
                             offrate=d["koff"])


def n_solvent(d):
    return d["density"] * d["Lx"] * d["Ly"] * d["Lz"]


def mesh(d):
    V, E, T = sc.mesh_for(d["nv"])
    V = V * d["radius"]
    if d["aspect"] != 1.0:
        a = d["aspect"]
        V = V * np.array([a ** (2.0 / 3.0), a ** (-1.0 / 3.0), a ** (-1.0 / 3.0)])
    if d["place"] == "reference":
        V = sc.place_reference(V, d["Lx"], d["Ly"])
    else:
        V = sc.place(V, d["Lx"], d["Ly"])
    return V, E, T


def spec(d):
    V, E, T = sc.mesh_for(d["nv"])
    return ia.make_spec(params(d), n_solvent(d), len(V), len(E), len(T), seed=d["seed"])


def state(d, host_solvent=True):
    """The replica's initial state as host arrays (the solvent drawn on the
    host unless host_solvent is False)."""
    p = params(d)
    V, E, T = mesh(d)
    return dict(V=V, E=E, T=T,
                sol=sc.solvent(n_solvent(d), d["Lx"], d["Ly"], d["Lz"], d["seed"]) if host_solvent else None,
                vel=sc.vertex_velocities(d["nv"], d["seed"] + 1, pmass=p.pmass),
                knob=sc.knobs_uniform(d["nv"], d["nknobs"], d["seed"] + 2),
                para=sc.parasite_state(V, d["seed"] + 3))


def load(b, r, d, st=None, device_solvent=False):
    """Put replica d's initial state into slot r of batch b; with
    device_solvent the solvent is drawn on the device from the replica's
    seed instead of on the host (no host arrays, no upload)."""
    st = state(d, host_solvent=not device_solvent) if st is None else st
    b.set_mesh(r, st["V"], st["E"], st["T"])
    if device_solvent:
        b.init_solvent(r)
    else:
        b.set_solvent(r, *st["sol"])
    v = st["vel"]
    b.set_vertices(r, vx=v[:, 0], vy=v[:, 1], vz=v[:, 2], knob=st["knob"])
    b.set_parasite(r, *st["para"])
    b.set_counters(r, 0, -1, 0)
    return st


def start(b):
    """maincode.cpp:383-391 for every replica: forces at rest, inside flags,
    cell->triangle lists."""
    b.membrane_phases(ia.PH_VA | ia.PH_FORCES)
    b.classify()
    b.llist_update()
    b.sync()


def batch(ds, device=0):
    b = ia.Batch([spec(d) for d in ds], device=device)
    for r, d in enumerate(ds):
        load(b, r, d)
    start(b)
    return b
