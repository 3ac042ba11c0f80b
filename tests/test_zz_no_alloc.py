"""psx_statistics_noalloc (no allocation: its partials live in the library's fixed device scratch
array) against psx_statistics (partials in memory the call allocates), through the C ABI, on two
calls in a row, the second with a smaller input, and on a batch of scenes."""
import ctypes as C

import numpy as np
import pytest

from helpers import dense_box, small_cylinder
from psx import _lib, scenes
from psx._lib import psx_stats

pytestmark = pytest.mark.gpu


def _both(lib, handle, n_scenes):
    a = (psx_stats * n_scenes)()
    b = (psx_stats * n_scenes)()
    assert lib.psx_statistics(handle, C.cast(a, C.c_void_p)) == 0, lib.psx_last_error()
    assert lib.psx_statistics_noalloc(handle, C.cast(b, C.c_void_p)) == 0, lib.psx_last_error()
    return a, b


def _same(a, b, what):
    for k, (x, y) in enumerate(zip(a, b)):
        for f, _ in psx_stats._fields_:
            vx, vy = getattr(x, f), getattr(y, f)
            vx = list(vx) if not isinstance(vx, (int, float)) else [vx]
            vy = list(vy) if not isinstance(vy, (int, float)) else [vy]
            assert vx == vy, (what, k, f, vx, vy)


def test_noalloc_statistics_equal_allocating_entry(psx_mod):
    lib = _lib.load()
    # first call: 262,144 spheres in the cylinder, after a few steps (contacts, overlaps)
    big = scenes.cylinder_settle()
    with psx_mod.Simulation(big.pos, big.vel, big.radius, big.mass, big.params, big.mesh) as s:
        s.step(20)
        a, b = _both(lib, s._h, 1)
        _same(a, b, "262k cylinder")
        assert a[0].kinetic_energy > 0.0
    # second call, smaller input: 512 spheres of an overlapping box (collided pairs)
    small = dense_box(n_side=8, dtype="f64")
    with psx_mod.Simulation(small.pos, small.vel, small.radius, small.mass, small.params) as s:
        a, b = _both(lib, s._h, 1)
        _same(a, b, "512-sphere box")
        assert a[0].n_collided > 0 and a[0].max_overlap > 0.0
    # a batch: one record per scene, each combined from its own workgroups
    batch = [dense_box(n_side=k, seed=k, dtype="f32") for k in (3, 9, 5)]
    batch.append(small_cylinder(n=700, dtype="f32"))
    with psx_mod.Batch(batch) as bt:
        bt.step(3)
        a, b = _both(lib, bt._h, len(batch))
        _same(a, b, "batch")


def test_noalloc_statistics_reject_oversized_input(psx_mod):
    """An input whose partials exceed the fixed scratch array is refused with PSX_ERR_INVALID
    (checked before any work is enqueued); the allocating entry still serves it."""
    lib = _lib.load()
    side = (64, 64, 3840)  # 15.7 M spheres, one per cell: more than the 4 MiB of partials hold
    n = side[0] * side[1] * side[2]
    k = np.arange(n)
    p = np.stack([k % side[0], (k // side[0]) % side[1], k // (side[0] * side[1])], 1) + 0.5
    r = np.full(n, 0.5)
    prm = psx_mod.Params(gravity=(0, 0, 0), dt=1e-4, box_min=(0, 0, 0),
                         box_max=(float(side[0]), float(side[1]), float(side[2])), dtype="f32")
    with psx_mod.Simulation(p, np.zeros((n, 3)), r, None, prm) as s:
        out = (psx_stats * 1)()
        rc = lib.psx_statistics_noalloc(s._h, C.cast(out, C.c_void_p))
        assert rc == 1 and b"scratch" in lib.psx_last_error()
        assert lib.psx_statistics(s._h, C.cast(out, C.c_void_p)) == 0
        assert out[0].n_collided == 0
