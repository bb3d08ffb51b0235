"""The allocation-free TotalVA entry point (irbc_membrane_va_static) against
the context's irbc_membrane_va: bit-identical area and volume on two calls in
a row, the second on a smaller membrane, through the C ABI on device memory
the caller allocated."""
import numpy as np
import pytest

import irbc_amd as ia
import scenario as sc

pytestmark = pytest.mark.gpu


def _membrane(nv, seed):
    """A perturbed icosphere in a 16x12x10 box, some vertices moved by whole
    box lengths and given the matching periodic images (ix, iy)."""
    p = ia.Params.default(Lx=16, Ly=12, Lz=10, walls=1.0, ramp_steps=0)
    V, E, T = sc.mesh_for(nv)
    V = sc.place(V * 2.6, p.Lx, p.Ly)
    rng = np.random.default_rng(seed)
    V = V + rng.normal(0.0, 0.02, V.shape)
    ix = np.zeros(nv, np.int32)
    iy = np.zeros(nv, np.int32)
    far = rng.choice(nv, nv // 5, replace=False)
    ix[far] = 1
    iy[far[: len(far) // 2]] = -1
    X = V.copy()
    X[:, 0] -= ix * p.Lx
    X[:, 1] -= iy * p.Ly
    return p, V, E, T, X, ix, iy


def _context_va(p, V, E, T, X, ix, iy):
    b = ia.Batch([ia.make_spec(p, 10, len(V), len(E), len(T), seed=3)])
    try:
        b.set_mesh(0, V, E, T)
        b.set_vertices(0, x=X[:, 0], y=X[:, 1], z=X[:, 2], ix=ix, iy=iy)
        return b.membrane_va(0)
    finally:
        b.close()


class _Hip:
    """Device buffers for the test through the HIP runtime the library
    itself links (torch bundles another runtime, which cannot share the
    process with it)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.h = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def put(self, a):
        C = self.C
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 8))) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p.value

    def get(self, ptr, n, dtype):
        C = self.C
        out = np.empty(n, dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(ptr), C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _static_va(p, T, X, ix, iy):
    hip = _Hip()
    try:
        d = [hip.put(np.ascontiguousarray(X[:, k], dtype=np.float64)) for k in range(3)]
        i = [hip.put(np.asarray(a, np.int32)) for a in (ix, iy)]
        t = hip.put(np.asarray(T, np.int32))
        out = hip.put(np.full(2, np.nan))
        ia.membrane_va_static(d[0], d[1], d[2], i[0], i[1], len(X), t, len(T), p.Lx, p.Ly, out, 0)
        return tuple(hip.get(out, 2, np.float64).tolist())
    finally:
        hip.free()


def test_static_va_equals_context_va_on_two_calls():
    for nv, seed in ((642, 1), (162, 2)):  # the second call is the smaller membrane
        p, V, E, T, X, ix, iy = _membrane(nv, seed)
        want = _context_va(p, V, E, T, X, ix, iy)
        got = _static_va(p, T, X, ix, iy)
        assert got == want, (nv, got, want)  # bit for bit
        assert np.isfinite(got).all() and got[0] > 0 and got[1] > 0
