"""Python binding of the MI355X iRBC library (libirbc_amd.so) through its C ABI.

The library must be built (irbc-code_amd/build.py) and a HIP device present;
there is no CPU fallback: loading fails loudly otherwise.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libirbc_amd.so")

PH_KICK1, PH_DRIFT, PH_VA, PH_FORCES = 1, 2, 4, 8
PH_EXTERNAL, PH_BONDS, PH_KICK2, PH_COUNTERS = 16, 32, 64, 128
PH_ALL = 255
PH_PARASITE, PH_MEMBRANE = 512, 1024  # the two halves of PH_FORCES (pfdt, force_calculate_dt)

EXPORTS = (
    "irbc_create", "irbc_destroy", "irbc_last_error", "irbc_device_bytes", "irbc_estimate_bytes",
    "irbc_sync", "irbc_set_mesh", "irbc_set_vertices", "irbc_get_vertices", "irbc_set_solvent",
    "irbc_get_solvent", "irbc_set_parasite", "irbc_get_parasite", "irbc_n_ligands",
    "irbc_set_ligands", "irbc_get_ligands", "irbc_get_globals", "irbc_set_counters",
    "irbc_get_counters", "irbc_set_wallspeed", "irbc_set_seed", "irbc_use_tapes",
    "irbc_set_collision_tape", "irbc_set_bond_tape", "irbc_collision_step", "irbc_llist_update",
    "irbc_stream", "irbc_classify_solvent", "irbc_membrane_phases", "irbc_membrane_va",
    "irbc_membrane_forces", "irbc_bond_step", "irbc_get_cells", "irbc_get_llist",
    "irbc_get_pairs", "irbc_md_steps", "irbc_set_diagnostics", "irbc_profile",
    "irbc_profile_read", "irbc_set_series", "irbc_get_series", "irbc_query", "irbc_set_option",
    "irbc_checkpoint_bytes", "irbc_checkpoint_save", "irbc_checkpoint_load", "irbc_autotune",
    "irbc_membrane_va_static", "irbc_init_solvent", "irbc_get_spec", "irbc_copy_replica",
    "irbc_get_option", "irbc_autotune_spans",
)

OPT_MEMBRANE_FUSED = 1
OPT_GRAPHS = 2
OPT_NEAR_MERGED = 3
OPT_XCD_CHUNK = 4
OPT_STREAM4 = 5
OPT_STREAM_GRID = 6
OPT_MEM_AFTER_STREAM = 7
OPT_MEMBRANE_CHAIN = 8
OPT_CHAIN_DEVICE_SCOPE = 9
OPT_NEAR_MASKS = 10
OPT_ONE_QUEUE = 11
OPT_CLASSIFY_ALL = 12
OPT_ALIGN_LISTS = 13
OPT_SORT_EVERY = 14
OPT_BIN_SPAN = 15
OPTIONS = (("membrane_fused", 1), ("near_merged", 3), ("graphs", 2), ("xcd_chunk", 4), ("stream4", 5),
           ("stream_grid", 6), ("mem_after_stream", 7), ("membrane_chain", 8), ("one_queue", 11))

SERIES_FIELDS = ("step", "x", "y", "z", "vx", "vy", "vz", "bonds")

_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_l = C.POINTER(C.c_int64)


class Params(C.Structure):
    """irbc_params (include/irbc_params.h)."""

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
        """define.h / maincode.cpp values, with keyword overrides."""
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

    def copy_to(self, other):
        for f, _ in self._fields_:
            setattr(other, f, getattr(self, f))
        return other


class Spec(C.Structure):
    """irbc_replica_spec."""

    _fields_ = [("params", Params), ("n_solvent", C.c_int64), ("n_vertices", C.c_int32),
                ("n_edges", C.c_int32), ("n_triangles", C.c_int32), ("pad", C.c_int32),
                ("seed", C.c_uint64)]


class IrbcError(RuntimeError):
    pass


_lib = None


def load():
    """Load libirbc_amd.so; raises if it is missing (no silent fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IrbcError("libirbc_amd.so not built: run python irbc-code_amd/build.py")
    L = C.CDLL(LIB_PATH)
    L.irbc_last_error.restype = C.c_char_p
    L.irbc_estimate_bytes.restype = C.c_int64
    L.irbc_estimate_bytes.argtypes = [C.POINTER(Spec), C.c_int]
    L.irbc_create.argtypes = [C.POINTER(Spec), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
    L.irbc_n_ligands.restype = C.c_int64
    L.irbc_n_ligands.argtypes = [C.c_void_p, C.c_int]
    L.irbc_get_llist.restype = C.c_int64
    L.irbc_get_llist.argtypes = [C.c_void_p, C.c_int, _l, _i, C.c_int64]
    L.irbc_get_pairs.restype = C.c_int64
    L.irbc_get_pairs.argtypes = [C.c_void_p, C.c_int, _i, _i, C.c_int64]
    L.irbc_set_counters.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.c_int32, C.c_int32]
    L.irbc_set_wallspeed.argtypes = [C.c_void_p, C.c_int, C.c_double]
    L.irbc_set_seed.argtypes = [C.c_void_p, C.c_int, C.c_uint64]
    L.irbc_set_bond_tape.argtypes = [C.c_void_p, C.c_int, _d, C.c_int64]
    L.irbc_set_solvent.argtypes = [C.c_void_p, C.c_int, C.c_int64] + [_d] * 6 + [_i]
    L.irbc_md_steps.argtypes = [C.c_void_p, C.c_int64]
    L.irbc_membrane_va.argtypes = [C.c_void_p, C.c_int, _d, _d]
    L.irbc_set_series.argtypes = [C.c_void_p, C.c_int, C.c_int32, C.c_int32]
    L.irbc_get_series.restype = C.c_int64
    L.irbc_get_series.argtypes = [C.c_void_p, C.c_int, _d, C.c_int64]
    L.irbc_query.argtypes = [C.c_void_p]
    L.irbc_autotune.argtypes = [C.c_void_p, C.c_int32, _d]
    L.irbc_checkpoint_bytes.restype = C.c_int64
    L.irbc_checkpoint_bytes.argtypes = [C.c_void_p, C.c_int]
    L.irbc_checkpoint_save.restype = C.c_int64
    L.irbc_checkpoint_save.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
    L.irbc_checkpoint_load.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
    L.irbc_copy_replica.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    L.irbc_get_option.argtypes = [C.c_void_p, C.c_int, _i]
    L.irbc_autotune_spans.argtypes = [C.c_void_p, _d]
    L.irbc_membrane_va_static.argtypes = [C.c_void_p] * 5 + [C.c_int32, C.c_void_p, C.c_int32, C.c_double,
                                                             C.c_double, C.c_void_p, C.c_void_p]
    _lib = L
    return L


def last_error():
    return load().irbc_last_error().decode()


def _chk(rc):
    if rc < 0:
        raise IrbcError("irbc error %d: %s" % (rc, last_error()))
    return rc


def _ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def membrane_va_static(x, y, z, ix, iy, nv, tris, nt, Lx, Ly, out, stream=0):
    """TotalVA of a membrane held in caller-owned device memory, with no
    allocation: every argument but the sizes and box lengths is a device
    address (an int, e.g. a torch tensor's data_ptr()); area and volume go
    to out[0], out[1] on the device, ordered on `stream` (0: the null
    stream)."""
    _chk(load().irbc_membrane_va_static(x, y, z, ix, iy, int(nv), tris, int(nt), float(Lx), float(Ly), out,
                                        stream or None))


def make_spec(params, n_solvent, nv, ne, nt, seed=1):
    s = Spec()
    params.copy_to(s.params)
    s.n_solvent, s.n_vertices, s.n_edges, s.n_triangles = int(n_solvent), int(nv), int(ne), int(nt)
    s.seed = int(seed)
    return s


def estimate_bytes(specs):
    arr = (Spec * len(specs))(*specs)
    return load().irbc_estimate_bytes(arr, len(specs))


class Batch:
    """A context of one or more replicas on one GPU."""

    VD = ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz")
    VI = ("ix", "iy", "pid", "bondi", "bondl", "knob")

    def __init__(self, specs, device=0):
        L = load()
        self.L = L
        self.specs = list(specs)
        arr = (Spec * len(self.specs))(*self.specs)
        h = C.c_void_p()
        _chk(L.irbc_create(arr, len(self.specs), int(device), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.irbc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def ncells(self, r=0):
        return self.specs[r].params.ncells()

    def device_bytes(self):
        b = C.c_int64()
        _chk(self.L.irbc_device_bytes(self.h, C.byref(b)))
        return b.value

    def sync(self):
        _chk(self.L.irbc_sync(self.h))

    # ---- setup
    def set_mesh(self, r, xyz, edges, tris):
        xyz, edges, tris = _f64(xyz), _i32(edges), _i32(tris)
        _chk(self.L.irbc_set_mesh(self.h, r, _ptr(xyz, _d), _ptr(edges, _i), _ptr(tris, _i)))

    def set_vertices(self, r, **kw):
        args, keep = [], []
        for k in self.VD:
            a = _f64(kw.get(k))
            keep.append(a)
            args.append(_ptr(a, _d))
        for k in self.VI:
            a = _i32(kw.get(k))
            keep.append(a)
            args.append(_ptr(a, _i))
        _chk(self.L.irbc_set_vertices(self.h, r, *args))

    def vertices(self, r=0):
        n = self.specs[r].n_vertices
        d = {k: np.zeros(n) for k in self.VD + ("poff", "pon")}
        for k in self.VI:
            d[k] = np.zeros(n, np.int32)
        _chk(self.L.irbc_get_vertices(self.h, r, *[_ptr(d[k], _d) for k in self.VD + ("poff", "pon")],
                                      *[_ptr(d[k], _i) for k in self.VI]))
        return d

    def set_solvent(self, r, x, y, z, vx, vy, vz, indio=None):
        arrs = [_f64(a) for a in (x, y, z, vx, vy, vz)]
        ind = _i32(indio)
        _chk(self.L.irbc_set_solvent(self.h, r, len(arrs[0]), *[_ptr(a, _d) for a in arrs],
                                     _ptr(ind, _i)))

    def init_solvent(self, r, maxwell=True):
        """A fresh solvent drawn on the device from the replica's seed."""
        _chk(self.L.irbc_init_solvent(self.h, r, 1 if maxwell else 0))

    def solvent(self, r=0):
        n = self.specs[r].n_solvent
        d = {k: np.zeros(n) for k in ("x", "y", "z", "vx", "vy", "vz")}
        d["key"] = np.zeros(n, np.int32)
        d["indio"] = np.zeros(n, np.int32)
        _chk(self.L.irbc_get_solvent(self.h, r, *[_ptr(d[k], _d) for k in ("x", "y", "z", "vx", "vy", "vz")],
                                     _ptr(d["key"], _i), _ptr(d["indio"], _i)))
        return d

    def set_parasite(self, r, d, i):
        d, i = _f64(d), _i32(i)
        _chk(self.L.irbc_set_parasite(self.h, r, _ptr(d, _d), _ptr(i, _i)))

    def parasite(self, r=0):
        d = np.zeros(12)
        i = np.zeros(3, np.int32)
        _chk(self.L.irbc_get_parasite(self.h, r, _ptr(d, _d), _ptr(i, _i)))
        return d, i

    def nlig(self, r=0):
        return self.L.irbc_n_ligands(self.h, r)

    def ligands(self, r=0):
        a = np.zeros(self.nlig(r), np.int32)
        _chk(self.L.irbc_get_ligands(self.h, r, _ptr(a, _i)))
        return a

    def set_ligands(self, r, a):
        a = _i32(a)
        _chk(self.L.irbc_set_ligands(self.h, r, _ptr(a, _i)))

    def globals(self, r=0):
        g = np.zeros(13)
        _chk(self.L.irbc_get_globals(self.h, r, _ptr(g, _d)))
        return dict(zip(("Sarea", "Volume", "Sarea0", "Volume0", "theta0", "wallspeed", "status",
                         "bond_draws", "step", "nbonds", "near_queue", "chain_xcd", "collision_kind"), g))

    def set_counters(self, r, step, colsteps, upd):
        _chk(self.L.irbc_set_counters(self.h, r, int(step), int(colsteps), int(upd)))

    def counters(self, r=0):
        s, c, u = C.c_int64(), C.c_int32(), C.c_int32()
        _chk(self.L.irbc_get_counters(self.h, r, C.byref(s), C.byref(c), C.byref(u)))
        return s.value, c.value, u.value

    def set_wallspeed(self, r, w):
        _chk(self.L.irbc_set_wallspeed(self.h, r, float(w)))

    def set_seed(self, r, seed):
        _chk(self.L.irbc_set_seed(self.h, r, int(seed)))

    def use_tapes(self, r, on=True):
        _chk(self.L.irbc_use_tapes(self.h, r, 1 if on else 0))

    def set_collision_tape(self, r, shift_u, cell_u, wall_z, mem_z):
        t = [_f64(a) for a in (shift_u, cell_u, wall_z, mem_z)]
        _chk(self.L.irbc_set_collision_tape(self.h, r, *[_ptr(a, _d) for a in t]))

    def set_bond_tape(self, r, u):
        u = _f64(u)
        _chk(self.L.irbc_set_bond_tape(self.h, r, _ptr(u, _d), len(u)))

    # ---- stages
    def collision_step(self):
        _chk(self.L.irbc_collision_step(self.h))

    def llist_update(self):
        _chk(self.L.irbc_llist_update(self.h))

    def stream(self):
        _chk(self.L.irbc_stream(self.h))

    def classify(self):
        _chk(self.L.irbc_classify_solvent(self.h))

    def membrane_phases(self, phases):
        _chk(self.L.irbc_membrane_phases(self.h, int(phases)))

    def membrane_va(self, r=0):
        a, v = C.c_double(), C.c_double()
        _chk(self.L.irbc_membrane_va(self.h, r, C.byref(a), C.byref(v)))
        return a.value, v.value

    def membrane_forces(self):
        _chk(self.L.irbc_membrane_forces(self.h))

    def bond_step(self):
        _chk(self.L.irbc_bond_step(self.h))

    def cells(self, r=0):
        n = self.ncells(r)
        d = {"vcm": np.zeros(3 * n), "nc": np.zeros(n, np.int32)}
        for k in ("ncc", "rand1", "rand2", "dvv", "the"):
            d[k] = np.zeros(n)
        _chk(self.L.irbc_get_cells(self.h, r, _ptr(d["vcm"], _d), _ptr(d["nc"], _i), _ptr(d["ncc"], _d),
                                   _ptr(d["rand1"], _d), _ptr(d["rand2"], _d), _ptr(d["dvv"], _d),
                                   _ptr(d["the"], _d)))
        return d

    def llist(self, r=0):
        n = self.ncells(r)
        off = np.zeros(n + 1, np.int64)
        tot = _chk(self.L.irbc_get_llist(self.h, r, _ptr(off, _l), None, 0))
        ent = np.zeros(max(tot, 1), np.int32)
        _chk(self.L.irbc_get_llist(self.h, r, _ptr(off, _l), _ptr(ent, _i), tot))
        return off, ent[:tot]

    def pairs(self, r=0):
        n = self.specs[r].n_vertices
        cnt = np.zeros(n, np.int32)
        tot = _chk(self.L.irbc_get_pairs(self.h, r, _ptr(cnt, _i), None, 0))
        lig = np.zeros(max(tot, 1), np.int32)
        _chk(self.L.irbc_get_pairs(self.h, r, _ptr(cnt, _i), _ptr(lig, _i), tot))
        return cnt, lig[:tot]

    def md_steps(self, n):
        _chk(self.L.irbc_md_steps(self.h, int(n)))

    def set_series(self, r, every, cap):
        """Record (step, centroid xyz, centroid velocity, bonds) every `every` steps."""
        _chk(self.L.irbc_set_series(self.h, r, int(every), int(cap)))

    def series(self, r=0):
        n = _chk(self.L.irbc_get_series(self.h, r, None, 0))
        out = np.zeros((max(n, 1), len(SERIES_FIELDS)))
        _chk(self.L.irbc_get_series(self.h, r, _ptr(out, _d), n))
        return out[:n]

    def checkpoint(self, r=0):
        """The replica's resumable state as bytes."""
        n = _chk(self.L.irbc_checkpoint_bytes(self.h, r))
        buf = C.create_string_buffer(n)
        m = _chk(self.L.irbc_checkpoint_save(self.h, r, buf, n))
        return buf.raw[:m]

    def restore(self, r, data):
        """Load a checkpoint into replica r (its mesh must be set)."""
        buf = C.create_string_buffer(bytes(data), len(data))
        _chk(self.L.irbc_checkpoint_load(self.h, r, buf, len(data)))

    def copy_from(self, r, src, rs):
        """Replica rs of batch `src` into slot r of this batch, device to
        device: state, mesh and time series (same sizes and parameters)."""
        _chk(self.L.irbc_copy_replica(self.h, int(r), src.h, int(rs)))

    def autotune(self, steps=20):
        """Pick the fastest launch options for this batch; returns
        (chosen index, ms per step of each of the 512 indices, NaN where
        not timed); bit i of the index is option OPTIONS[i]."""
        rep = np.zeros(512)
        k = _chk(self.L.irbc_autotune(self.h, int(steps), _ptr(rep, _d)))
        return k, rep

    def done(self):
        """True when all launched work has finished (non-blocking)."""
        return _chk(self.L.irbc_query(self.h)) == 1

    def set_option(self, option, value):
        _chk(self.L.irbc_set_option(self.h, int(option), int(value)))

    def option(self, option):
        v = C.c_int32()
        _chk(self.L.irbc_get_option(self.h, int(option), C.byref(v)))
        return v.value

    def autotune_spans(self):
        """ms per step the autotuner measured with bin spans 1, 2, 4."""
        ms = np.zeros(3)
        _chk(self.L.irbc_autotune_spans(self.h, _ptr(ms, _d)))
        return ms

    def set_diagnostics(self, on=True):
        _chk(self.L.irbc_set_diagnostics(self.h, 1 if on else 0))

    def profile(self, every=1):
        """Time every `every`-th step's kernels with HIP events (0 = off)."""
        _chk(self.L.irbc_profile(self.h, int(every)))

    def profile_read(self):
        t = np.zeros(8)
        _chk(self.L.irbc_profile_read(self.h, _ptr(t, _d)))
        return {"stream_ms": t[0], "stream_n": int(t[1]), "collision_ms": t[2],
                "collision_n": int(t[3]), "membrane_ms": t[4], "membrane_n": int(t[5]),
                "near_ms": t[6], "near_n": int(t[7])}
