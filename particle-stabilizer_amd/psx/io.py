"""CSV particles and OBJ containers through the library's host-side readers and writers
(psx_read_csv / psx_write_csv / psx_read_obj, include/psx.h), which follow ParticleIO.cpp:8-144
and PolygonContainer::readFromFile (PolygonContainer.cpp:381-417), and the post-run outputs
(psx_resolve_overlaps, psx_write_obj_spheres: ParticleModel.cpp:154-252, ObjData.cpp:8-140)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check

_DP = C.POINTER(C.c_double)


def _take(lib, ptr, count, shape):
    arr = np.ctypeslib.as_array(ptr, shape=(count,)).copy().reshape(shape)
    lib.psx_free(C.cast(ptr, C.c_void_p))
    return arr


def read_csv(path: str | os.PathLike, velocity=(0.0, 0.0, -10.0), density: float = 1.0) -> dict:
    """ParticleData::readFromCSV: returns position, velocity, radius, mass and skipped lines."""
    lib = _lib.load()
    v = (C.c_double * 3)(*[float(x) for x in velocity])
    n = C.c_int64()
    skipped = C.c_int64()
    ptrs = [_DP() for _ in range(4)]
    check(lib.psx_read_csv(os.fsencode(path), float(density), v, C.byref(n),
                           *[C.byref(p) for p in ptrs], C.byref(skipped)))
    k = n.value
    return {"position": _take(lib, ptrs[0], 3 * k, (k, 3)),
            "velocity": _take(lib, ptrs[1], 3 * k, (k, 3)),
            "radius": _take(lib, ptrs[2], k, (k,)),
            "mass": _take(lib, ptrs[3], k, (k,)),
            "skipped": skipped.value}


def write_csv(path: str | os.PathLike, position, radius) -> None:
    """ParticleData::writeAsCSV: header x,y,z,radius and 10-character fixed-point values."""
    lib = _lib.load()
    pos = np.ascontiguousarray(position, np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(radius, np.float64).reshape(-1)
    check(lib.psx_write_csv(os.fsencode(path), pos.shape[0], pos.ctypes.data_as(_DP),
                            r.ctypes.data_as(_DP)))


def read_obj(path: str | os.PathLike) -> np.ndarray:
    """De-indexed (T, 9) triangle soup of an OBJ container, face order kept."""
    lib = _lib.load()
    nt = C.c_int64()
    ptr = _DP()
    check(lib.psx_read_obj(os.fsencode(path), C.byref(nt), C.byref(ptr)))
    return _take(lib, ptr, 9 * nt.value, (nt.value, 9))


def resolve_overlaps(position, radius, axis: int):
    """ParticleModel::resolveOverlaps: returns (lifted positions, collided count, max overlap),
    the count and maximum being the overlap check the reference logs afterwards."""
    lib = _lib.load()
    pos = np.array(position, np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(radius, np.float64).reshape(-1)
    if r.shape[0] != pos.shape[0]:
        raise ValueError("position and radius differ in length")
    nc = C.c_int64()
    mo = C.c_double()
    check(lib.psx_resolve_overlaps(pos.shape[0], pos.ctypes.data_as(_DP), r.ctypes.data_as(_DP),
                                   int(axis), C.byref(nc), C.byref(mo)))
    return pos, nc.value, mo.value


def write_obj_spheres(path: str | os.PathLike, position, radius) -> None:
    """ObjData::write: every particle as a triangulated UV sphere in one OBJ file."""
    lib = _lib.load()
    pos = np.ascontiguousarray(position, np.float64).reshape(-1, 3)
    r = np.ascontiguousarray(radius, np.float64).reshape(-1)
    if r.shape[0] != pos.shape[0]:
        raise ValueError("position and radius differ in length")
    check(lib.psx_write_obj_spheres(os.fsencode(path), pos.shape[0], pos.ctypes.data_as(_DP),
                                    r.ctypes.data_as(_DP)))
