"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_l = C.POINTER(C.c_int64)


class Params(C.Structure):
    """Mirror of irbc_params (include/irbc_params.h)."""

    _fields_ = [
        ("Lx", C.c_int32), ("Ly", C.c_int32), ("Lz", C.c_int32),
        ("csteps", C.c_int32), ("uplist", C.c_int32),
        ("navg", C.c_double), ("mass", C.c_double), ("Theta", C.c_double),
        ("dtcd", C.c_double), ("dtmd", C.c_double),
        ("cosalpha", C.c_double), ("sinalpha", C.c_double),
        ("walls", C.c_double), ("ramp_speed", C.c_double), ("ramp_steps", C.c_int64),
        ("kappasa", C.c_double), ("kappad", C.c_double), ("kappav", C.c_double),
        ("kappab", C.c_double), ("pmass", C.c_double),
        ("ks", C.c_double), ("dthr", C.c_double), ("FD", C.c_double), ("Fg", C.c_double),
        ("el0", C.c_double), ("elap", C.c_double), ("shear", C.c_double),
        ("PRmass", C.c_double), ("paraR", C.c_double),
        ("onrate", C.c_double), ("offrate", C.c_double), ("lig_spacing", C.c_double),
        ("Pi", C.c_double),
    ]

    @classmethod
    def default(cls, **kw):
        p = cls()
        p.Lx, p.Ly, p.Lz = 60, 40, 40
        p.navg, p.mass, p.Theta = 10.0, 1.0, 1.0
        p.dtcd, p.dtmd = 0.02, 0.002
        p.cosalpha, p.sinalpha = -0.70710678118, 0.70710678118
        p.walls, p.ramp_speed, p.ramp_steps = 3.0, 0.1, 250000
        p.kappasa, p.kappad, p.kappav, p.kappab = 60000.0, 1200.0, 6000.0, 100.0
        p.ks, p.dthr, p.FD, p.Fg, p.el0 = 6000.0, 0.50, 1200.0, 40.0, 0.1
        p.elap, p.shear = 0.0005, 1200.0
        p.PRmass, p.paraR = 10.0, 1.7
        p.onrate, p.offrate, p.lig_spacing = 0.005, 0.015, 0.5
        p.Pi = 3.141592653
        for k, v in kw.items():
            setattr(p, k, v)
        if "pmass" not in kw:
            p.pmass = p.navg * p.mass
        p.csteps = int(p.dtcd / p.dtmd)
        p.uplist = max(1, int(0.5 / (p.walls * p.dtmd)))
        return p

    def ncells(self):
        return self.Lx * self.Ly * (self.Lz + 1)


def _ptr(a, t):
    if a is None:
        return None
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params)]
        for name in ("orc_ncells", "orc_n_solvent", "orc_n_ligands", "orc_llist_total"):
            getattr(L, name).restype = C.c_int64
            getattr(L, name).argtypes = [C.c_void_p]
        L.orc_get_bond_tape.restype = C.c_int64
        L.orc_get_pairs.restype = C.c_int64
        L.orc_mt_uniform.restype = C.c_double
        L.orc_mt_bits.restype = C.c_uint32
        L.orc_sto_normal.restype = C.c_double
        L.orc_sto_normal.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_reflection.restype = C.c_double
        L.orc_reflection.argtypes = [C.c_void_p, C.c_int64]
        L.orc_solve.restype = C.c_double
        L.orc_solve.argtypes = [C.c_double] * 4
        L.orc_solve_p3.restype = C.c_int
        L.orc_solve_p3.argtypes = [_d, C.c_double, C.c_double, C.c_double]
        L.orc_stream.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_wallspeed.argtypes = [C.c_void_p, C.c_double]
        L.orc_set_counters.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int32]
        L.orc_set_bond_tape.argtypes = [C.c_void_p, _d, C.c_int64]
        L.orc_get_bond_tape.argtypes = [C.c_void_p, _d, C.c_int64]
        L.orc_get_pairs.argtypes = [C.c_void_p, _i, _i, C.c_int64]
        L.orc_set_solvent.argtypes = [C.c_void_p, C.c_int64] + [_d] * 6 + [_i]
        L.orc_seed_mt.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.orc_init_readme.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int, _d, C.c_int, _i,
                                      C.c_int, _i, C.c_int, _d]
        L.orc_set_mesh.argtypes = [C.c_void_p, C.c_int, _d, C.c_int, _i, C.c_int, _i]
        L.orc_use_philox.argtypes = [C.c_void_p, C.c_uint64]
        L.orc_philox_u2.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint32, _d]
        L.orc_philox_block.argtypes = [C.POINTER(C.c_uint32), C.c_uint32, C.c_uint32]
        L.orc_set_fp32.argtypes = [C.c_void_p, C.c_int]
        _lib = L
    return _lib


def set_threads(n):
    lib().orc_set_threads(int(n))


def threads():
    return lib().orc_get_threads()


def philox_block(ctr, key):
    """One Philox4x32-10 block: ctr = 4 words, key = 2 words; returns 4 words."""
    c = (C.c_uint32 * 4)(*[int(v) & 0xFFFFFFFF for v in ctr])
    lib().orc_philox_block(c, int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF)
    return [int(v) for v in c]


def philox_u2(seed, stream, index, sub):
    """The two uniforms of the counter-based stream (seed, stream, index, sub)."""
    u = np.zeros(2)
    lib().orc_philox_u2(int(seed), int(stream), int(index), int(sub), _ptr(u, _d))
    return float(u[0]), float(u[1])


def read_mesh(vfile, efile, tfile):
    """Mesh files of the reference format: x y z per line; 1-based index lists."""
    v = np.loadtxt(vfile, dtype=np.float64).reshape(-1, 3)
    e = np.loadtxt(efile, dtype=np.int64).reshape(-1, 2) - 1
    t = np.loadtxt(tfile, dtype=np.int64).reshape(-1, 3) - 1
    return v, e.astype(np.int32), t.astype(np.int32)


class Oracle:
    def __init__(self, params=None, **kw):
        self.p = params if params is not None else Params.default(**kw)
        self.h = C.c_void_p(lib().orc_create(C.byref(self.p)))
        self.L = lib()
        self.nc = self.L.orc_ncells(self.h)

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # ---- streams
    def seed_mt(self, ran, sto):
        self.L.orc_seed_mt(self.h, int(ran), int(sto))

    def use_tapes(self, on=True):
        self.L.orc_use_tapes(self.h, 1 if on else 0)

    def use_philox(self, seed):
        """Draw every random number from the device's counter-based streams."""
        self.L.orc_use_philox(self.h, int(seed))

    def set_fp32(self, on=True):
        """Store the solvent in FP32 where the device does (rounds it now)."""
        self.L.orc_set_fp32(self.h, 1 if on else 0)

    def set_collision_tape(self, shift_u, cell_u, wall_z, mem_z):
        self._tape = [_f64(shift_u), _f64(cell_u), _f64(wall_z), _f64(mem_z)]
        self.L.orc_set_collision_tape(self.h, *[_ptr(a, _d) for a in self._tape])

    def get_collision_tape(self):
        s = np.zeros(3)
        cu = np.zeros(2 * self.nc)
        wz = np.zeros(3 * self.nc)
        mz = np.zeros(3 * self.nc)
        self.L.orc_get_collision_tape(self.h, _ptr(s, _d), _ptr(cu, _d), _ptr(wz, _d), _ptr(mz, _d))
        return s, cu, wz, mz

    def set_bond_tape(self, u):
        self._bt = _f64(u)
        self.L.orc_set_bond_tape(self.h, _ptr(self._bt, _d), len(self._bt))

    def get_bond_tape(self):
        n = self.L.orc_get_bond_tape(self.h, None, 0)
        u = np.zeros(n)
        self.L.orc_get_bond_tape(self.h, _ptr(u, _d), n)
        return u

    def mt_uniform(self):
        return self.L.orc_mt_uniform(self.h)

    def mt_bits(self):
        return self.L.orc_mt_bits(self.h)

    def sto_normal(self, m=0.0, s=1.0):
        return self.L.orc_sto_normal(self.h, m, s)

    # ---- state
    def set_mesh(self, xyz, edges, tris):
        xyz, edges, tris = _f64(xyz), _i32(edges), _i32(tris)
        self.L.orc_set_mesh(self.h, len(xyz), _ptr(xyz, _d), len(edges), _ptr(edges, _i), len(tris),
                            _ptr(tris, _i))
        self.ne, self.nt = len(edges), len(tris)

    def set_solvent(self, x, y, z, vx, vy, vz, indio=None):
        arrs = [_f64(a) for a in (x, y, z, vx, vy, vz)]
        ind = None if indio is None else _i32(indio)
        self.L.orc_set_solvent(self.h, len(arrs[0]), *[_ptr(a, _d) for a in arrs], _ptr(ind, _i))

    def solvent(self):
        n = self.L.orc_n_solvent(self.h)
        d = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz")}
        d["pid"] = np.zeros(n, np.int32)
        d["indio"] = np.zeros(n, np.int32)
        self.L.orc_get_solvent(self.h, *[_ptr(d[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz")],
                               _ptr(d["pid"], _i), _ptr(d["indio"], _i))
        return d

    def nv(self):
        return self.L.orc_n_vertices(self.h)

    VD = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")
    VI = ("ix", "iy", "pid", "bondi", "bondl", "knob")

    def set_vertices(self, **kw):
        args = []
        keep = []
        for k in self.VD:
            a = None if kw.get(k) is None else _f64(kw[k])
            keep.append(a)
            args.append(_ptr(a, _d))
        for k in self.VI:
            a = None if kw.get(k) is None else _i32(kw[k])
            keep.append(a)
            args.append(_ptr(a, _i))
        self.L.orc_set_vertices(self.h, *args)

    def vertices(self):
        n = self.nv()
        d = {k: np.zeros(n) for k in self.VD + ("poff", "pon")}
        for k in self.VI:
            d[k] = np.zeros(n, np.int32)
        self.L.orc_get_vertices(self.h, *[_ptr(d[k], _d) for k in self.VD + ("poff", "pon")],
                                *[_ptr(d[k], _i) for k in self.VI])
        return d

    def set_parasite(self, d, i):
        d, i = _f64(d), _i32(i)
        self.L.orc_set_parasite(self.h, _ptr(d, _d), _ptr(i, _i))

    def parasite(self):
        d = np.zeros(12)
        i = np.zeros(3, np.int32)
        self.L.orc_get_parasite(self.h, _ptr(d, _d), _ptr(i, _i))
        return d, i

    def nlig(self):
        return self.L.orc_n_ligands(self.h)

    def ligands(self):
        a = np.zeros(self.nlig(), np.int32)
        self.L.orc_get_ligands(self.h, _ptr(a, _i))
        return a

    def set_ligands(self, a):
        a = _i32(a)
        self.L.orc_set_ligands(self.h, _ptr(a, _i))

    def ligand_xy(self):
        x = np.zeros(self.nlig())
        y = np.zeros(self.nlig())
        self.L.orc_get_ligand_xy(self.h, _ptr(x, _d), _ptr(y, _d))
        return x, y

    def cells(self):
        n = self.nc
        d = {"vcm": np.zeros(3 * n), "nc": np.zeros(n, np.int32)}
        for k in ("ncc", "rand1", "rand2", "dvv", "the"):
            d[k] = np.zeros(n)
        self.L.orc_get_cells(self.h, _ptr(d["vcm"], _d), _ptr(d["nc"], _i), _ptr(d["ncc"], _d),
                             _ptr(d["rand1"], _d), _ptr(d["rand2"], _d), _ptr(d["dvv"], _d),
                             _ptr(d["the"], _d))
        return d

    def globals(self):
        g = np.zeros(10)
        self.L.orc_get_globals(self.h, _ptr(g, _d))
        return dict(zip(("Sarea", "Volume", "Sarea0", "Volume0", "theta0", "wallspeed", "cx", "cy",
                         "cz", "nsolb"), g))

    def topology(self):
        l0 = np.zeros(self.ne)
        kp = np.zeros(self.ne)
        es = np.zeros(2 * self.ne, np.int32)
        ie = np.zeros(self.ne, np.int32)
        ta = np.zeros(self.nt)
        self.L.orc_get_topology(self.h, _ptr(l0, _d), _ptr(kp, _d), _ptr(es, _i), _ptr(ie, _i),
                                _ptr(ta, _d))
        return {"l0": l0, "kappap": kp, "edgest": es.reshape(-1, 2), "iedge": ie, "tarea": ta}

    def llist(self):
        tot = self.L.orc_llist_total(self.h)
        off = np.zeros(self.nc + 1, np.int64)
        ent = np.zeros(max(tot, 1), np.int32)
        self.L.orc_get_llist(self.h, _ptr(off, _l), _ptr(ent, _i))
        return off, ent[:tot]

    def pairs(self):
        n = self.nv()
        cnt = np.zeros(n, np.int32)
        tot = self.L.orc_get_pairs(self.h, _ptr(cnt, _i), None, 0)
        lig = np.zeros(max(tot, 1), np.int32)
        self.L.orc_get_pairs(self.h, _ptr(cnt, _i), _ptr(lig, _i), tot)
        return cnt, lig[:tot]

    def set_counters(self, step, colsteps, upd):
        self.L.orc_set_counters(self.h, int(step), int(colsteps), int(upd))

    def counters(self):
        s = C.c_int64()
        c = C.c_int32()
        u = C.c_int32()
        self.L.orc_get_counters(self.h, C.byref(s), C.byref(c), C.byref(u))
        return s.value, c.value, u.value

    # ---- stages
    def collision(self):
        self.L.orc_collision(self.h)

    def llistupdate(self):
        self.L.orc_llistupdate(self.h)

    def stream(self, wallspeed):
        self.L.orc_stream(self.h, float(wallspeed))

    def set_wallspeed(self, w):
        self.L.orc_set_wallspeed(self.h, float(w))

    def reflection(self, i):
        return self.L.orc_reflection(self.h, int(i))

    def __getattr__(self, name):
        # zero-argument stages: kick1, drift, totalva, totalva0, pfdt, forces,
        # external, bond_prep, bond_dynamics, kick2, centerofmass, makeiedge,
        # sortsolvent, tricm, md_step, classify
        if name in ("kick1", "drift", "totalva", "totalva0", "pfdt", "forces", "external",
                    "bond_prep", "bond_dynamics", "kick2", "centerofmass", "makeiedge",
                    "sortsolvent", "tricm", "md_step", "classify"):
            f = getattr(lib(), "orc_" + name)
            return lambda: f(self.h)
        raise AttributeError(name)

    def init_readme(self, seed, ns, xyz_raw, edges, tris, nknobs):
        xyz_raw, edges, tris = _f64(xyz_raw), _i32(edges), _i32(tris)
        st = np.zeros(8)
        r = self.L.orc_init_readme(self.h, int(seed), int(ns), len(xyz_raw), _ptr(xyz_raw, _d),
                                   len(edges), _ptr(edges, _i), len(tris), _ptr(tris, _i), int(nknobs),
                                   _ptr(st, _d))
        if r != 0:
            raise ValueError("orc_init_readme failed: %d" % r)
        self.ne, self.nt = len(edges), len(tris)
        return st


def solve_p3(a, b, c):
    x = np.zeros(3)
    n = lib().orc_solve_p3(_ptr(x, _d), a, b, c)
    return n, x


def solve(a, b, c, dtmd=0.002):
    return lib().orc_solve(a, b, c, dtmd)
