"""CPU-side checks of the drop-in boundary: the library loads without a GPU
and exports every entry point include/irbc.h declares; the host-side scenario
code builds valid meshes."""
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)


def _declared():
    names = set()
    for hdr in ("irbc.h",):
        txt = open(os.path.join(REPO, "include", hdr)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(irbc_[a-z0-9_]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    import ctypes
    import irbc_amd
    lib = irbc_amd.load()
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(irbc_amd.EXPORTS) <= _declared()
    assert len(_declared()) >= 39
    assert isinstance(lib, ctypes.CDLL)


def test_create_without_device_fails_loudly():
    import irbc_amd
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        return  # a GPU may be present
    p = irbc_amd.Params.default(Lx=16, Ly=12, Lz=10)
    s = irbc_amd.make_spec(p, 100, 42, 120, 80)
    try:
        irbc_amd.Batch([s])
    except irbc_amd.IrbcError as e:
        assert "device" in str(e).lower()
    else:
        raise AssertionError("creating a context without a GPU must fail")


def test_params_struct_layout_matches_header():
    import irbc_amd
    import pyoracle
    assert [f for f, _ in irbc_amd.Params._fields_] == [f for f, _ in pyoracle.Params._fields_]
    import ctypes
    # 5 int32 + padding, 27 doubles, one int64
    assert ctypes.sizeof(irbc_amd.Params) == ctypes.sizeof(pyoracle.Params) == 24 + 27 * 8 + 8


def test_icospheres_are_closed_and_outward():
    import scenario as sc
    for nv, ne, nt in ((162, 480, 320), (642, 1920, 1280), (2562, 7680, 5120)):
        V, E, T = sc.mesh_for(nv)
        assert (len(V), len(E), len(T)) == (nv, ne, nt)
        assert nv - ne + nt == 2
        # every edge in exactly two triangles
        cnt = {}
        for a, b, c in T:
            for p, q in ((a, b), (b, c), (c, a)):
                k = (min(p, q), max(p, q))
                cnt[k] = cnt.get(k, 0) + 1
        assert set(cnt.values()) == {2}
        n = np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]])
        vol = np.einsum("ij,ij->i", n, V[T].sum(axis=1) / 3.0).sum() / 6.0
        assert 0.95 * 4.0 / 3.0 * np.pi < vol < 4.0 / 3.0 * np.pi
