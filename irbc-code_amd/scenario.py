"""Scenario setup for the iRBC step: meshes, placement, solvent, knobs.

The reference reads a triangulated cell from three files and places it
(maincode.cpp:165-272); the workloads here also need seeded icospheres and
ellipsoids of 162 / 642 / 2562 vertices.  Everything is numpy on the host;
the state is then handed to the device through the C ABI.
"""
import functools

import numpy as np


def icosphere(subdiv):
    """Unit icosphere: 10*4^s + 2 vertices, triangles oriented outward."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t),
         (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        mid = {}

        def midpoint(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                mid[k] = len(v) - 1
            return mid[k]

        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(v)
    T = np.array(f, dtype=np.int32)
    # orient outward: normal . centroid > 0
    n = np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]])
    flip = np.einsum("ij,ij->i", n, V[T].mean(axis=1)) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return V, T


def edges_of(tris):
    """Unique edges in order of first appearance over the triangle list."""
    seen = {}
    for a, b, c in tris:
        for p, q in ((a, b), (b, c), (c, a)):
            k = (min(p, q), max(p, q))
            if k not in seen:
                seen[k] = (int(p), int(q))
    return np.array(list(seen.values()), dtype=np.int32)


@functools.lru_cache(maxsize=None)
def mesh_for(nv):
    """The unit icosphere of nv vertices: positions, edges, triangles.  Built
    once per size (a request loop sets up many replicas a second) and
    returned read-only: callers scale and place copies."""
    sub = {12: 0, 42: 1, 162: 2, 642: 3, 2562: 4}
    if nv not in sub:
        raise ValueError("icosphere sizes are 12, 42, 162, 642, 2562 vertices")
    V, T = icosphere(sub[nv])
    E = edges_of(T)
    for a in (V, E, T):
        a.setflags(write=False)
    return V, E, T


def place_reference(V, Lx, Ly):
    """Centre as maincode.cpp:196-206: x = Lx/2 + 10, y = Ly/2, lowest vertex at z = 0.3."""
    zmin = np.sort(V[:, 2])[0]
    c = np.array([Lx / 2.0 + 10.0, Ly / 2.0, -zmin + 0.3])
    return V + c


def place(V, Lx, Ly, z_lowest=0.3, x=None):
    zmin = V[:, 2].min()
    cx = Lx / 2.0 if x is None else x
    return V + np.array([cx, Ly / 2.0, -zmin + z_lowest])


def solvent(n, Lx, Ly, Lz, seed, kT=1.0, mass=1.0, maxwell=True, fp32=True):
    """Uniform positions in the box; Maxwellian (or the reference's uniform
    +-sqrt(3kT/m)) velocities.  With fp32 the values are rounded to what the
    device stores, so a CPU check sees exactly the device's input."""
    rng = np.random.default_rng(seed)
    x = rng.random(n) * Lx
    y = rng.random(n) * Ly
    z = rng.random(n) * Lz
    if maxwell:
        v = rng.standard_normal((3, n)) * np.sqrt(kT / mass)
    else:
        vmax = np.sqrt(3.0 * kT / mass)
        v = -vmax + 2.0 * vmax * rng.random((3, n))
    out = [x, y, z, v[0], v[1], v[2]]
    if fp32:
        out = [a.astype(np.float32).astype(np.float64) for a in out]
    # a position that rounds up to the box edge stays inside
    out[0] = np.where(out[0] >= Lx, 0.0, out[0])
    out[1] = np.where(out[1] >= Ly, 0.0, out[1])
    out[2] = np.clip(out[2], 0.0, Lz)
    return out


def knobs_uniform(nv, n, seed):
    rng = np.random.default_rng(seed)
    k = np.zeros(nv, np.int32)
    k[rng.choice(nv, size=min(n, nv), replace=False)] = 1
    return k


def vertex_velocities(nv, seed, kT=1.0, pmass=10.0):
    rng = np.random.default_rng(seed)
    vmax = np.sqrt(3.0 * kT / pmass)
    return -vmax + 2.0 * vmax * rng.random((nv, 3))


def parasite_state(V_placed, seed, kT=1.0, PRmass=10.0, cutoff=1.7):
    """Bead behind the cell centre along -x (maincode.cpp:212-225): 2 units
    as in the reference, less in a small cell, so that the bead starts at
    least a WCA cutoff (plus a margin) from every vertex; a bead that starts
    on the membrane meets forces of order 1e20."""
    rng = np.random.default_rng(seed)
    c = np.array([V_placed[:, 0].mean(), V_placed[:, 1].mean(), V_placed[:, 2].mean()])
    r_in = np.sqrt(((V_placed - c) ** 2).sum(1)).min()
    off = min(2.0, max(0.0, r_in - (cutoff + 0.1)))
    vmax = np.sqrt(3.0 * kT / PRmass)
    d = np.zeros(12)
    d[0:3] = c - np.array([off, 0.0, 0.0])
    d[3:6] = -vmax + 2.0 * vmax * rng.random(3)
    return d, np.zeros(3, np.int32)
