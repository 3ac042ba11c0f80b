"""ctypes wrapper of oracle/_build/libpsx_oracle.so (test infrastructure only).

Every function mirrors an `extern "C"` entry of psx_oracle.cpp, which cites the reference file and
line it restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libpsx_oracle.so"

_lib = None
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        lib.oracle_grid.argtypes = [C.c_double, _DP, _DP, _DP, C.POINTER(C.c_int64)]
        lib.oracle_grid.restype = None
        lib.oracle_axis_index.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64]
        lib.oracle_axis_index.restype = C.c_int64
        lib.oracle_run_f64.argtypes = [_DP, C.c_int64, _DP, _DP, _DP, _DP, C.c_int64, _DP,
                                       C.c_int64, C.c_int]
        lib.oracle_run_f32.argtypes = [_DP, C.c_int64, _FP, _FP, _FP, _FP, C.c_int64, _DP,
                                       C.c_int64, C.c_int]
        lib.oracle_bucketing_f64.argtypes = [C.c_double, _DP, _DP, C.c_int64, _DP, _I32P, _U32P,
                                             _U32P, _I32P]
        lib.oracle_bucketing_f32.argtypes = [C.c_double, _DP, _DP, C.c_int64, _FP, _I32P, _U32P,
                                             _U32P, _I32P]
        lib.oracle_closest_point_f64.argtypes = [_DP, _DP, _DP, _DP, _DP]
        lib.oracle_closest_point_f64.restype = None
        lib.oracle_mesh_normals_f64.argtypes = [C.c_int64, _DP, _DP]
        lib.oracle_mesh_normals_f64.restype = None
        lib.oracle_mass.argtypes = [C.c_double, C.c_double]
        lib.oracle_mass.restype = C.c_double
        lib.oracle_max_threads.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def grid(rmax: float, mn, mx):
    lib = load()
    mn = np.ascontiguousarray(mn, np.float64)
    mx = np.ascontiguousarray(mx, np.float64)
    h = C.c_double()
    n3 = np.zeros(3, np.int64)
    lib.oracle_grid(rmax, _p(mn, C.c_double), _p(mx, C.c_double), C.byref(h), _p(n3, C.c_int64))
    return h.value, n3


def axis_index(x: float, mn: float, h: float, n: int) -> int:
    return int(load().oracle_axis_index(x, mn, h, n))


def run(params: np.ndarray, pos, vel, radius, mass, nsteps: int, mesh=None, dtype="f64",
        nthreads: int | None = None):
    """Runs nsteps reference steps; returns (pos, vel) in `dtype`."""
    lib = load()
    t = np.float64 if dtype in ("f64", np.float64) else np.float32
    ct = C.c_double if t is np.float64 else C.c_float
    params = np.ascontiguousarray(params, np.float64)
    assert params.shape == (18,)
    pos = np.array(pos, dtype=t, order="C", copy=True).reshape(-1, 3)
    vel = np.array(vel, dtype=t, order="C", copy=True).reshape(-1, 3)
    r = np.ascontiguousarray(radius, dtype=t)
    m = np.ascontiguousarray(mass, dtype=t)
    n = r.shape[0]
    ntri = 0
    tri = None
    if mesh is not None:
        tri = np.ascontiguousarray(mesh, np.float64).reshape(-1, 9)
        ntri = tri.shape[0]
    fn = lib.oracle_run_f64 if t is np.float64 else lib.oracle_run_f32
    nthreads = nthreads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or lib.oracle_max_threads()
    rc = fn(_p(params, C.c_double), n, _p(pos, ct), _p(vel, ct), _p(r, ct), _p(m, ct), ntri,
            _p(tri, C.c_double) if tri is not None else None, int(nsteps), int(nthreads))
    if rc != 0:
        raise RuntimeError(f"oracle run failed ({rc})")
    return pos, vel


def bucketing(rmax: float, mn, mx, pos, dtype="f64"):
    lib = load()
    t = np.float64 if dtype in ("f64", np.float64) else np.float32
    ct = C.c_double if t is np.float64 else C.c_float
    h, n3 = grid(rmax, mn, mx)
    ncell = int(n3.prod())
    pos = np.ascontiguousarray(pos, dtype=t).reshape(-1, 3)
    n = pos.shape[0]
    cell = np.empty(n, np.int32)
    count = np.empty(ncell, np.uint32)
    offset = np.empty(ncell + 1, np.uint32)
    perm = np.empty(n, np.int32)
    mn = np.ascontiguousarray(mn, np.float64)
    mx = np.ascontiguousarray(mx, np.float64)
    fn = lib.oracle_bucketing_f64 if t is np.float64 else lib.oracle_bucketing_f32
    fn(rmax, _p(mn, C.c_double), _p(mx, C.c_double), n, _p(pos, ct), _p(cell, C.c_int32),
       _p(count, C.c_uint32), _p(offset, C.c_uint32), _p(perm, C.c_int32))
    return {"cell": cell, "count": count, "offset": offset, "perm": perm, "h": h, "n": n3}


def closest_point(a, b, c, p):
    lib = load()
    arrs = [np.ascontiguousarray(x, np.float64) for x in (a, b, c, p)]
    q = np.zeros(3)
    lib.oracle_closest_point_f64(*[_p(x, C.c_double) for x in arrs], _p(q, C.c_double))
    return q


def mesh_normals(tri):
    lib = load()
    tri = np.ascontiguousarray(tri, np.float64).reshape(-1, 9)
    out = np.zeros((tri.shape[0], 3))
    lib.oracle_mesh_normals_f64(tri.shape[0], _p(tri, C.c_double), _p(out, C.c_double))
    return out


def mass(density: float, radius: float) -> float:
    return float(load().oracle_mass(density, radius))


def max_threads() -> int:
    return int(load().oracle_max_threads())
