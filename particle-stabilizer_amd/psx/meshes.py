"""Triangle-soup containers and OBJ I/O.

A container is an (T, 9) float64 array of de-indexed triangles (corner a, b, c), the layout
PolygonContainer builds from its mesh (PolygonContainer.cpp:26-42, 401-416). Normals follow the
winding, normalize(cross(b - a, c - a)) (PolygonContainer.cpp:67), and particles are kept on the
-normal side, so every generator here winds its triangles with OUTWARD normals.
"""
from __future__ import annotations

import math
import os

import numpy as np


def _quad(p00, p10, p11, p01):
    """Two triangles of a quad with corners in counter-clockwise order seen from outside."""
    return [np.concatenate([p00, p10, p11]), np.concatenate([p00, p11, p01])]


def _ring(radius: float, z: float, n_az: int) -> np.ndarray:
    th = 2.0 * np.pi * np.arange(n_az + 1) / n_az
    th[-1] = 0.0
    return np.stack([radius * np.cos(th), radius * np.sin(th), np.full(n_az + 1, z)], axis=1)


def _side(r0: float, z0: float, r1: float, z1: float, n_az: int) -> list:
    lo, hi = _ring(r0, z0, n_az), _ring(r1, z1, n_az)
    tris = []
    for k in range(n_az):
        tris += _quad(lo[k], lo[k + 1], hi[k + 1], hi[k])
    return tris


def _fan(radius: float, z: float, n: int) -> list:
    """Disc at height z facing down (normal -z), fanned from its centre."""
    ring = _ring(radius, z, n)
    c = np.array([0.0, 0.0, z])
    return [np.concatenate([c, ring[k + 1], ring[k]]) for k in range(n)]


def open_cylinder(radius: float = 40.0, height: float = 300.0, n_az: int = 128, n_ax: int = 8,
                  n_floor: int = 128) -> np.ndarray:
    """Open-top cylinder on z=0: n_az x n_ax side quads (2 triangles each) + an n_floor fan floor.

    Defaults: radius 40, height 300, 128 x 8 side quads = 2,048 triangles plus a 128-triangle
    floor fan, 2,176 triangles in all.
    """
    tris = []
    for k in range(n_ax):
        tris += _side(radius, height * k / n_ax, radius, height * (k + 1) / n_ax, n_az)
    tris += _fan(radius, 0.0, n_floor)
    return np.asarray(tris, dtype=np.float64)


def hopper(outlet_radius: float = 10.0, upper_radius: float = 150.0, half_angle_deg: float = 35.0,
           upper_height: float = 400.0, n_az: int = 128, n_cone: int = 15, n_cyl: int = 16,
           n_floor: int = 256) -> np.ndarray:
    """Cone-bottom hopper: a floor disc of the outlet radius at z=0, a cone (half-angle 35 deg)
    widening to the upper radius, then a cylinder of the upper height (open top).

    Defaults give 256 + 128*2*(15+16) = 8,192 triangles; the cone reaches z = 140/tan(35 deg).
    """
    cone_h = (upper_radius - outlet_radius) / math.tan(math.radians(half_angle_deg))
    tris = []
    for k in range(n_cone):
        r0 = outlet_radius + (upper_radius - outlet_radius) * k / n_cone
        r1 = outlet_radius + (upper_radius - outlet_radius) * (k + 1) / n_cone
        tris += _side(r0, cone_h * k / n_cone, r1, cone_h * (k + 1) / n_cone, n_az)
    for k in range(n_cyl):
        tris += _side(upper_radius, cone_h + upper_height * k / n_cyl, upper_radius,
                      cone_h + upper_height * (k + 1) / n_cyl, n_az)
    tris += _fan(outlet_radius, 0.0, n_floor)
    return np.asarray(tris, dtype=np.float64)


def box_mesh(lo=(0.0, 0.0, 0.0), hi=(40.0, 40.0, 60.0), nu: int = 2, nv: int = 4,
             open_top: bool = True) -> np.ndarray:
    """Axis-aligned box of nu x nv quads per face (2 triangles each), outward normals.

    Defaults: open top, 5 faces x 16 triangles = 80 triangles.
    """
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    tris = []

    def face(origin, eu, ev):
        # quads spanned by eu (nu steps) and ev (nv steps); outward normal = eu x ev
        for i in range(nu):
            for j in range(nv):
                p00 = origin + eu * (i / nu) + ev * (j / nv)
                p10 = origin + eu * ((i + 1) / nu) + ev * (j / nv)
                p11 = origin + eu * ((i + 1) / nu) + ev * ((j + 1) / nv)
                p01 = origin + eu * (i / nu) + ev * ((j + 1) / nv)
                tris.extend(_quad(p00, p10, p11, p01))

    d = hi - lo
    ex, ey, ez = np.array([d[0], 0, 0]), np.array([0, d[1], 0]), np.array([0, 0, d[2]])
    face(lo, ey, ex)                                   # z min, normal -z
    face(lo, ex, ez)                                   # y min, normal -y
    face(lo + ey, ez, ex)                              # y max, normal +y
    face(lo, ez, ey)                                   # x min, normal -x
    face(lo + ex, ey, ez)                              # x max, normal +x
    if not open_top:
        face(lo + ez, ex, ey)                          # z max, normal +z
    return np.asarray(tris, dtype=np.float64)


def normals(tris: np.ndarray) -> np.ndarray:
    t = np.asarray(tris, np.float64).reshape(-1, 3, 3)
    c = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    return c / np.linalg.norm(c, axis=1, keepdims=True)


def write_obj(path: str | os.PathLike, tris: np.ndarray, name: str = "container") -> None:
    """Triangle-only, single-object OBJ; the vertex list is the de-indexed soup in face order."""
    t = np.asarray(tris, np.float64).reshape(-1, 9)
    with open(path, "w") as f:
        f.write(f"o {name}\n")
        for row in t:
            for k in range(3):
                f.write("v %.17g %.17g %.17g\n" % (row[3 * k], row[3 * k + 1], row[3 * k + 2]))
        for i in range(t.shape[0]):
            f.write(f"f {3 * i + 1} {3 * i + 2} {3 * i + 3}\n")
