"""Statistical checks of the production path (counter-based streams).

Couette flow between the walls: the bottom wall is at rest, the top wall
moves at `wallspeed` (maincode.cpp:523-531, Src/solfuncs.cpp:178-206).  Over
many collisions the mean x-velocity of the solvent must stay a linear
function of z with slope wallspeed/Lz, and the thermostat keeps the
temperature of the peculiar velocities at kT = 1.  Both are properties of
the physics, independent of how the draws are generated.
"""
import numpy as np
import pytest

import ensemble as en


def _couette(w, Lz=12, seed=31):
    d = en.replica(20, 10, Lz, 10, 42, 1.5, 1.0, 10, kon=0.005, koff=0.015, shear=w / Lz, seed=seed)
    b = en.batch([d])
    s = b.solvent(0)
    b.set_solvent(0, s["x"], s["y"], s["z"], s["vx"] + w * s["z"] / Lz, s["vy"], s["vz"], s["indio"])
    return d, b


@pytest.mark.gpu
def test_shear_profile_linear_and_kT_one():
    w, Lz = 1.2, 12
    d, b = _couette(w, Lz)
    b.md_steps(2000)
    nb = 2 * Lz
    sum_v, cnt, t_sum, t_n = np.zeros(nb), np.zeros(nb), 0.0, 0
    for _ in range(30):
        b.md_steps(100)
        s = b.solvent(0)
        out = s["indio"] != 0  # outside the cell
        z, vx = s["z"][out], s["vx"][out]
        k = np.clip((z / Lz * nb).astype(int), 0, nb - 1)
        sum_v += np.bincount(k, vx, nb)
        cnt += np.bincount(k, None, nb)
        u = w * z / Lz
        pec = (vx - u) ** 2 + s["vy"][out] ** 2 + s["vz"][out] ** 2
        t_sum += pec.sum()
        t_n += 3 * len(pec)
    g = b.globals(0)
    assert g["status"] == 0
    zc = (np.arange(nb) + 0.5) * Lz / nb
    prof = sum_v / cnt
    inner = (zc > 2.0) & (zc < Lz - 2.0)
    slope, icpt = np.polyfit(zc[inner], prof[inner], 1)
    fit = slope * zc[inner] + icpt
    r2 = 1.0 - np.sum((prof[inner] - fit) ** 2) / np.sum((prof[inner] - prof[inner].mean()) ** 2)
    assert r2 > 0.98, (r2, prof)
    assert abs(slope - w / Lz) < 0.15 * w / Lz, (slope, w / Lz)
    kT = t_sum / t_n
    assert abs(kT - 1.0) < 0.05, kT


@pytest.mark.gpu
def test_device_drawn_solvent():
    """irbc_init_solvent: the same seed draws the same solvent in any slot
    of any batch, another seed another one; positions lie in the box and
    velocities have kT = 1 (Maxwellian) or the reference's uniform spread
    (all read back in particle order, through the ids)."""
    import ensemble as en
    import irbc_amd as ia
    d = en.replica(20, 16, 12, 10, 162, 2.6, 1.0, 100, seed=91)
    e = en.replica(16, 12, 10, 10, 42, 1.8, 1.0, 20, seed=92)
    a = ia.Batch([en.spec(d)])
    b = ia.Batch([en.spec(e), en.spec(d)])
    try:
        for bb, r, dd in ((a, 0, d), (b, 0, e), (b, 1, d)):
            en.load(bb, r, dd, device_solvent=True)
        sa, sb, se = a.solvent(0), b.solvent(1), b.solvent(0)
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)
        assert not np.array_equal(sa["vx"][:len(se["vx"])], se["vx"])
        n = len(sa["x"])
        assert n == 20 * 16 * 12 * 10
        assert sa["x"].min() >= 0 and sa["x"].max() < 20 and sa["y"].max() < 16
        assert sa["z"].min() >= 0 and sa["z"].max() <= 12
        v = np.concatenate([sa["vx"], sa["vy"], sa["vz"]])
        assert abs(v.mean()) < 0.01 and abs(v.var() - 1.0) < 0.01
        assert abs(sa["x"].mean() - 10.0) < 0.1
        b.init_solvent(1, maxwell=False)
        u = b.solvent(1)
        assert np.abs(u["vx"]).max() <= np.sqrt(3.0) and abs(np.var(u["vx"]) - 1.0) < 0.02
        a.md_steps(20)
        a.sync()
    finally:
        a.close()
        b.close()
