"""The oracle's counter-based stream mode and FP32 storage mode, pinned on
the CPU.

Philox4x32-10 is checked against the published known-answer vectors of its
authors (Random123, Salmon et al. SC'11), the uniform mapping against the
block output itself, and the keying of every draw (grid shift, rotation
axes) against the stream ids and counters the device uses.  FP32 storage
must leave every solvent value representable in FP32 after whole steps, and
the mode must be deterministic: the same seed gives the same bits, another
seed does not.
"""
import numpy as np

import pyoracle as po
from cases import Case

# (counter, key) -> block, from the Random123 known-answer vectors
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answers(oracle_lib):
    for ctr, key, want in KAT:
        assert tuple(po.philox_block(ctr, key)) == want


def test_philox_uniforms_are_the_top_53_bits(oracle_lib):
    seed, stream, index, sub = 0x1234_5678_9abc_def0, 3, (7 << 32) | 11, 5
    c = po.philox_block((index & 0xFFFFFFFF, stream, index >> 32, sub), (seed & 0xFFFFFFFF, seed >> 32))
    a = ((c[0] << 32) | c[1]) >> 11
    b = ((c[2] << 32) | c[3]) >> 11
    assert po.philox_u2(seed, stream, index, sub) == (a / 2.0 ** 53, b / 2.0 ** 53)


def test_counter_mode_keys_shift_and_axes_by_step_and_cell(oracle_lib):
    k = Case(seed=3, nv=42, radius=1.8)
    o = k.o
    seed = 99
    o.use_philox(seed)
    o.set_counters(17, 9, 0)  # step 17 collides (colsteps 9 -> 10 = csteps)
    o.md_step()
    shift, cell_u, _, _ = o.get_collision_tape()
    s0, s1 = po.philox_u2(seed, 1, 17, 0)
    s2, _ = po.philox_u2(seed, 1, 17, 1)
    assert tuple(shift) == (s0, s1, s2)
    for c in (0, 1, 777, o.nc - 1):
        assert tuple(cell_u[2 * c:2 * c + 2]) == po.philox_u2(seed, 2, 17, c)


def _run(seed, fp32, steps=12):
    k = Case(seed=4, nv=42, radius=1.8, kon=40.0, koff=40.0)
    o = k.o
    if fp32:
        o.set_fp32(True)
    o.use_philox(seed)
    for _ in range(steps):
        o.md_step()
    return o


def test_fp32_storage_keeps_the_solvent_representable(oracle_lib):
    o = _run(5, True)
    s = o.solvent()
    for key in ("x", "y", "z", "vx", "vy", "vz"):
        a = s[key]
        np.testing.assert_array_equal(a, a.astype(np.float32).astype(np.float64), err_msg=key)


def test_counter_mode_is_deterministic_and_seeded(oracle_lib):
    a, b, c = _run(5, True), _run(5, True), _run(6, True)
    sa, sb, sc = a.solvent(), b.solvent(), c.solvent()
    for key in ("x", "vx", "pid"):
        np.testing.assert_array_equal(sa[key], sb[key])
    assert not np.array_equal(sa["vx"], sc["vx"])
    np.testing.assert_array_equal(a.vertices()["bondi"], b.vertices()["bondi"])
