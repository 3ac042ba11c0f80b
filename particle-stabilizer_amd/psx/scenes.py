"""Seeded synthetic workloads (numpy PCG64 via default_rng(seed)).

    box_settle        4,096 spheres in a box, float64 (the CPU-runnable settle)
    cylinder_settle   262,144 spheres in an open-top cylinder mesh (2,176 triangles), float32
    hopper_settle     4,194,304 spheres in a cone-bottom hopper mesh (8,192 triangles), float32
    batched_boxes     256 independent box scenes of uneven size, float32
    calibration_scene 16,384 spheres in an 80-triangle open box mesh, float32

Each returns a `Scene` (or a list of them) whose `meta` records seed and generator.
"""
from __future__ import annotations

import math

import numpy as np

from . import meshes
from .sim import Params, Scene, mass_of

GENERATOR = f"numpy {np.__version__} PCG64 default_rng"


def _lattice_box(n: int, origin, spacing: float, nx: int, ny: int, rng, jitter: float):
    """n points on an nx x ny x ... lattice (x fastest, then y, then z layers), jittered."""
    k = np.arange(n)
    ix, iy, iz = k % nx, (k // nx) % ny, k // (nx * ny)
    p = np.stack([ix, iy, iz], axis=1).astype(np.float64) * spacing + np.asarray(origin, np.float64)
    return p + rng.uniform(-jitter, jitter, size=(n, 3))


def box_settle(n_side: int = 16, seed: int = 0, steps: int = 20000, dtype: str = "f64") -> Scene:
    """16^3 lattice from (0.5, 0.5, 20), spacing 1.2, jitter +-0.05; r ~ U[0.45, 0.55]; box
    (0,0,0)-(20,20,60); g (0,0,-9.8); spring 1e4; restitutions 0.5; dt 1e-4."""
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    pos = _lattice_box(n, (0.5, 0.5, 20.0), 1.2, n_side, n_side, rng, 0.05)
    r = rng.uniform(0.45, 0.55, size=n)
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=1e4, restitution=0.5,
                 dt=1e-4, box_min=(0.0, 0.0, 0.0), box_max=(20.0, 20.0, 60.0), dtype=dtype)
    return Scene(pos, np.zeros((n, 3)), r, prm, mass_of(r),
                 meta={"name": "box_settle", "seed": seed, "steps": steps, "generator": GENERATOR})


def _fill_axisymmetric(n: int, spacing: float, radius_at, z0: float, rng, jitter: float,
                       wall_clear: float):
    """Lattice points layer by layer from z0 upward inside r <= radius_at(z) - wall_clear."""
    out = []
    got = 0
    z = z0
    while got < n:
        rz = radius_at(z) - wall_clear
        if rz > 0:
            m = int(rz / spacing) + 1
            g = np.arange(-m, m + 1) * spacing
            X, Y = np.meshgrid(g, g, indexing="ij")
            keep = X * X + Y * Y <= rz * rz
            layer = np.stack([X[keep], Y[keep], np.full(keep.sum(), z)], axis=1)
            out.append(layer[: n - got])
            got += min(layer.shape[0], n - got)
        z += spacing
        if z > 1e6:
            raise RuntimeError("container too small for the requested particle count")
    p = np.concatenate(out)
    return p + rng.uniform(-jitter, jitter, size=p.shape)


def cylinder_settle(n: int = 262144, seed: int = 0, steps: int = 5000, dtype: str = "f32") -> Scene:
    """Open-top cylinder (R 40, H 300, 2,176 triangles); lattice spacing 1.7 inside R 38, jitter
    +-0.05; r = 0.5*exp(0.15 Z) clipped to [0.3, 0.8]; g (0,0,-9.8); spring 1e4; e 0.5; dt 1e-4."""
    rng = np.random.default_rng(seed)
    mesh = meshes.open_cylinder(40.0, 300.0, 128, 8, 128)
    pos = _fill_axisymmetric(n, 1.7, lambda z: 40.0, 1.7, rng, 0.05, 2.0)
    r = np.clip(0.5 * np.exp(0.15 * rng.standard_normal(n)), 0.3, 0.8)
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=1e4, restitution=0.5, dt=1e-4,
                 dtype=dtype)
    return Scene(pos, np.zeros((n, 3)), r, prm, mass_of(r), mesh,
                 meta={"name": "cylinder_settle", "seed": seed, "steps": steps,
                       "triangles": int(mesh.shape[0]), "generator": GENERATOR})


def cylinder_packed(n: int = 262144, seed: int = 0, steps: int = 1000, dtype: str = "f32",
                    factor: float = 0.6) -> Scene:
    """The cylinder settle with its lattice heights compressed by `factor` about z = 1 (0.6:
    vertical spacing 1.02, so every sphere's vertical neighbours are within the pair prefilter's
    reach and 0.88 pairs per sphere touch): a contact-rich state like a settled pile's, for timing
    the pair kernel's exact tests."""
    sc = cylinder_settle(n=n, seed=seed, steps=steps, dtype=dtype)
    sc.pos[:, 2] = 1.0 + (sc.pos[:, 2] - 1.0) * factor
    sc.meta.update(name="cylinder_packed", height_factor=factor)
    return sc


def hopper_settle(n: int = 4194304, seed: int = 0, steps: int = 2000, dtype: str = "f32") -> Scene:
    """Cone-bottom hopper (outlet R 10, cone half-angle 35 deg to R 150, cylinder height 400,
    8,192 triangles); r in {0.4 (70%), 0.8 (30%)}; lattice spacing 1.8, v0 (0,0,-1); spring 2e4;
    sphere e 0.4; wall e 0.3; dt 5e-5."""
    rng = np.random.default_rng(seed)
    mesh = meshes.hopper()
    t = math.tan(math.radians(35.0))
    cone_h = 140.0 / t

    def radius_at(z):
        return 150.0 if z >= cone_h else 10.0 + z * t

    pos = _fill_axisymmetric(n, 1.8, radius_at, 1.8, rng, 0.05, 2.0)
    r = np.where(rng.uniform(size=n) < 0.7, 0.4, 0.8)
    vel = np.zeros((n, 3))
    vel[:, 2] = -1.0
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.3,) * 6, spring=2e4, restitution=0.4,
                 dt=5e-5, dtype=dtype)
    return Scene(pos, vel, r, prm, mass_of(r), mesh,
                 meta={"name": "hopper_settle", "seed": seed, "steps": steps,
                       "triangles": int(mesh.shape[0]), "generator": GENERATOR})


def batched_boxes(n_scenes: int = 256, seed: int = 0, dtype: str = "f32",
                  n_min: int = 1024, n_max: int = 65536) -> list[Scene]:
    """Independent boxes: N_s = floor(exp(U[ln n_min, ln n_max])); edge 1.3*cbrt(N)*1.2 in x, y
    and 3x that in z; r ~ U[0.45, 0.55]; e_s ~ U[0.3, 0.9]; spring ~ U[5e3, 2e4]; g (0,0,-9.8);
    dt 1e-4; jittered lattice (spacing 1.2) seeded with seed + 1 + s."""
    rng = np.random.default_rng(seed)
    counts = np.floor(np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=n_scenes))).astype(int)
    es = rng.uniform(0.3, 0.9, size=n_scenes)
    ks = rng.uniform(5e3, 2e4, size=n_scenes)
    scenes = []
    for s in range(n_scenes):
        n = int(counts[s])
        srng = np.random.default_rng(seed + 1 + s)
        edge = 1.3 * n ** (1.0 / 3.0) * 1.2
        nx = max(1, int((edge - 1.2) / 1.2))
        pos = _lattice_box(n, (0.6 + 0.05, 0.6 + 0.05, 0.6 + 0.05), 1.2, nx, nx, srng, 0.05)
        r = srng.uniform(0.45, 0.55, size=n)
        prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=float(ks[s]),
                     restitution=float(es[s]), dt=1e-4, box_min=(0.0, 0.0, 0.0),
                     box_max=(edge, edge, 3.0 * edge), dtype=dtype)
        scenes.append(Scene(pos, np.zeros((n, 3)), r, prm, mass_of(r),
                            meta={"name": f"batched_boxes[{s}]", "seed": seed + 1 + s,
                                  "generator": GENERATOR}))
    return scenes


def calibration_scene(n_xy: int = 32, n_z: int = 16, seed: int = 0, dtype: str = "f32",
                      spring: float = 1e4, restitution: float = 0.5, v0z: float = -5.0) -> Scene:
    """n_xy x n_xy x n_z lattice (spacing 1.2, jitter +-0.05) in an 80-triangle open box mesh
    (0,0,0)-(40,40,60); r ~ U[0.45, 0.55]; g (0,0,-9.8); dt 1e-4; initial velocity (0, 0, v0z).
    The lattice starts 0.4-0.5 above the floor: from rest it would not reach it within a
    2,000-step window (0.2 s), and spring and restitution would have no contact to act on; at
    -5 it lands after about 900 steps and the layers above pile onto it for the rest."""
    rng = np.random.default_rng(seed)
    n = n_xy * n_xy * n_z
    pos = _lattice_box(n, (1.0, 1.0, 1.0), 1.2, n_xy, n_xy, rng, 0.05)
    r = rng.uniform(0.45, 0.55, size=n)
    edge = max(40.0, 1.2 * n_xy + 2.0)
    mesh = meshes.box_mesh((0.0, 0.0, 0.0), (edge, edge, 60.0))
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=spring,
                 restitution=restitution, dt=1e-4, dtype=dtype)
    vel = np.zeros((n, 3))
    vel[:, 2] = v0z
    return Scene(pos, vel, r, prm, mass_of(r), mesh,
                 meta={"name": "calibration_scene", "seed": seed, "generator": GENERATOR})
