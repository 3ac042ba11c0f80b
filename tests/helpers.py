"""Shared builders for the parity tests: small seeded scenes and oracle/HIP runners."""
from __future__ import annotations

import numpy as np

from psx import meshes
from psx.sim import Params, Scene, mass_of


def dense_box(n_side=10, spacing=0.95, seed=0, dtype="f64", spring=1e4, es=0.5, walls=None,
              vscale=1.0, box=(10.0, 10.0, 12.0)) -> Scene:
    """Overlapping lattice in a box: contacts and wall hits from the first step."""
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    k = np.arange(n)
    p = np.stack([k % n_side, (k // n_side) % n_side, k // (n_side * n_side)], 1) * spacing + 0.6
    p = p + rng.uniform(-0.05, 0.05, size=(n, 3))
    r = rng.uniform(0.45, 0.55, n)
    v = rng.normal(0.0, vscale, size=(n, 3))
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=walls or (0.5, 0.6, 0.7, 0.8, 0.9, 0.4),
                 spring=spring, restitution=es, dt=1e-4, box_min=(0.0, 0.0, 0.0), box_max=box,
                 dtype=dtype)
    return Scene(p, v, r, prm, mass_of(r))


def small_cylinder(n=800, seed=1, dtype="f64", radius=6.0, height=20.0, n_az=24, n_ax=4) -> Scene:
    """Particles packed near the floor and wall of a small open cylinder, moving outward/down."""
    rng = np.random.default_rng(seed)
    mesh = meshes.open_cylinder(radius, height, n_az, n_ax, n_az)
    pts = []
    while len(pts) < n:
        q = rng.uniform([-radius, -radius, 0.0], [radius, radius, 6.0])
        if q[0] ** 2 + q[1] ** 2 <= (radius - 0.3) ** 2:
            pts.append(q)
    p = np.asarray(pts)
    r = rng.uniform(0.3, 0.5, n)
    v = rng.normal(0.0, 1.0, size=(n, 3))
    v[:, :2] += 2.0 * p[:, :2] / radius  # outward
    v[:, 2] -= 2.0
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=100.0, restitution=0.5,
                 dt=1e-3, dtype=dtype)
    return Scene(p, v, r, prm, mass_of(r), mesh)


def oracle_run(oracle, sc: Scene, steps: int, dtype=None):
    dtype = dtype or sc.params.dtype
    return oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                      mesh=sc.mesh, dtype=dtype, semantics=sc.params.semantics)


def hip_run(sc: Scene, steps: int, graph=False, **tuning):
    from psx import Simulation
    with Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        if graph or tuning:
            s.set_tuning(use_graph=1 if graph else 0, **tuning)
        s.step(steps)
        s.synchronize()
        return s.state()


def as_dtype(x, dtype):
    return np.asarray(x, np.float64).astype(np.float32 if dtype == "f32" else np.float64)


def assert_bitwise(a, b, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    eq = (a == b) | (np.isnan(a) & np.isnan(b))
    if not eq.all():
        idx = np.argwhere(~eq)[:5]
        raise AssertionError(f"{what}: {int((~eq).sum())} mismatches, first {idx.tolist()}: "
                             f"{a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}")
