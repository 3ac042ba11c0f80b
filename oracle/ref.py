"""ctypes wrapper of oracle/_build/libpsx_oracle.so (test infrastructure only).

Every function mirrors an `extern "C"` entry of psx_oracle.cpp, which cites the reference file and
line it restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# $PSX_ORACLE_LIBRARY selects another build of the same source (the sanitizer build)
LIB = Path(os.environ.get("PSX_ORACLE_LIBRARY") or HERE / "_build" / "libpsx_oracle.so")

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
        lib.oracle_grid.argtypes = [C.c_double, _DP, _DP, C.c_int, _DP, C.POINTER(C.c_int64)]
        lib.oracle_grid.restype = None
        lib.oracle_axis_index.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64]
        lib.oracle_axis_index.restype = C.c_int64
        lib.oracle_run_f64.argtypes = [_DP, C.c_int64, _DP, _DP, _DP, _DP, C.c_int64, _DP,
                                       C.c_int64, C.c_int, C.c_int]
        lib.oracle_run_f32.argtypes = [_DP, C.c_int64, _FP, _FP, _FP, _FP, C.c_int64, _DP,
                                       C.c_int64, C.c_int, C.c_int]
        lib.oracle_bucketing_f64.argtypes = [C.c_double, _DP, _DP, C.c_int64, _DP, _I32P, _U32P,
                                             _U32P, _I32P]
        lib.oracle_bucketing_f32.argtypes = [C.c_double, _DP, _DP, C.c_int64, _FP, _I32P, _U32P,
                                             _U32P, _I32P]
        lib.oracle_closest_point_f64.argtypes = [_DP, _DP, _DP, _DP, _DP]
        lib.oracle_closest_point_f64.restype = None
        lib.oracle_closest_point_cuda_f64.argtypes = [_DP, _DP, _DP, _DP, _DP]
        lib.oracle_closest_point_cuda_f64.restype = None
        lib.oracle_mesh_normals_f64.argtypes = [C.c_int64, _DP, _DP]
        lib.oracle_mesh_normals_f64.restype = None
        lib.oracle_mass.argtypes = [C.c_double, C.c_double]
        lib.oracle_mass.restype = C.c_double
        lib.oracle_max_threads.restype = C.c_int
        _lib = lib
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def grid(rmax: float, mn, mx, semantics: str = "cpu"):
    lib = load()
    mn = np.ascontiguousarray(mn, np.float64)
    mx = np.ascontiguousarray(mx, np.float64)
    h = C.c_double()
    n3 = np.zeros(3, np.int64)
    lib.oracle_grid(rmax, _p(mn, C.c_double), _p(mx, C.c_double), _cuda(semantics), C.byref(h),
                    _p(n3, C.c_int64))
    return h.value, n3


def axis_index(x: float, mn: float, h: float, n: int) -> int:
    return int(load().oracle_axis_index(x, mn, h, n))


def _cuda(semantics: str) -> int:
    if semantics not in ("cpu", "cuda"):
        raise ValueError(f"semantics must be 'cpu' or 'cuda', not {semantics!r}")
    return 1 if semantics == "cuda" else 0


def run(params: np.ndarray, pos, vel, radius, mass, nsteps: int, mesh=None, dtype="f64",
        nthreads: int | None = None, fast: bool = False, semantics: str = "cpu"):
    """Runs nsteps reference steps (semantics "cpu": the reference's CPU backend; "cuda": its CUDA
    backend's step, the kicks and mesh terms added in ascending order); returns (pos, vel) in
    `dtype`. fast: the mesh loop visits only the triangles a triangle-grid index offers and the
    pair loop is split among threads (the results of the reference's loops, much sooner); by
    default every triangle and a serial pair loop, as the reference runs them."""
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
            _p(tri, C.c_double) if tri is not None else None, int(nsteps), int(nthreads),
            (1 if fast else 0) | (2 * _cuda(semantics)))
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


def closest_point(a, b, c, p, semantics: str = "cpu"):
    lib = load()
    arrs = [np.ascontiguousarray(x, np.float64) for x in (a, b, c, p)]
    q = np.zeros(3)
    fn = lib.oracle_closest_point_cuda_f64 if _cuda(semantics) else lib.oracle_closest_point_f64
    fn(*[_p(x, C.c_double) for x in arrs], _p(q, C.c_double))
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


def statistics(pos, vel, radius, mass) -> dict:
    """ParticleStatistics::calculateStatistics (ParticleStatistics.cpp:3-43), O(N^2)."""
    lib = load()
    lib.oracle_statistics.argtypes = [C.c_int64, _DP, _DP, _DP, _DP, _DP]
    lib.oracle_statistics.restype = None
    pos = np.ascontiguousarray(pos, np.float64).reshape(-1, 3)
    vel = np.ascontiguousarray(vel, np.float64).reshape(-1, 3)
    radius = np.ascontiguousarray(radius, np.float64).reshape(-1)
    mass = np.ascontiguousarray(mass, np.float64).reshape(-1)
    out = np.zeros(9)
    lib.oracle_statistics(pos.shape[0], _p(pos, C.c_double), _p(vel, C.c_double),
                          _p(radius, C.c_double), _p(mass, C.c_double), _p(out, C.c_double))
    return {"kinetic_energy": out[0], "min": out[1:4].copy(), "max": out[4:7].copy(),
            "n_collided": int(out[7]), "max_overlap": out[8]}


def resolve_overlaps(pos, radius, axis: int):
    """ParticleModel::resolveOverlaps (ParticleModel.cpp:154-252), literal O(N^2) restatement:
    returns (lifted positions, collided count, max overlap)."""
    lib = load()
    lib.oracle_resolve_overlaps.argtypes = [C.c_int64, _DP, _DP, C.c_int32, _DP]
    lib.oracle_resolve_overlaps.restype = None
    pos = np.array(pos, np.float64).reshape(-1, 3)
    radius = np.ascontiguousarray(radius, np.float64).reshape(-1)
    out = np.zeros(2)
    lib.oracle_resolve_overlaps(pos.shape[0], _p(pos, C.c_double), _p(radius, C.c_double),
                                int(axis), _p(out, C.c_double))
    return pos, int(out[0]), float(out[1])


def obj_sphere_template(n: int):
    """ObjData::write (ObjData.cpp:20-73): the UV-sphere template for n particles, returned as
    (ndivs, unit vertices (3 * faces, 3) float32). Angles and products in binary32, sin / cos by the
    C library's sinf / cosf; theta + dtheta and phi + dphi are rounded sums as in the reference."""
    libm = C.CDLL("libm.so.6")
    for fn in ("sinf", "cosf"):
        getattr(libm, fn).argtypes = [C.c_float]
        getattr(libm, fn).restype = C.c_float
    f32 = np.float32
    ndivs = max(int(math.sqrt(1000000.0 / (6.0 + n))) + 1, 6)
    nth, nph = ndivs, 2 * ndivs
    dth = f32(math.pi / float(f32(nth - 1)))
    dph = f32(2 * math.pi / float(f32(nph - 1)))
    th = np.arange(nth - 1, dtype=f32) * dth
    th1 = th + dth
    ph = np.arange(nph - 1, dtype=f32) * dph
    ph1 = ph + dph
    trig = lambda fn, a: np.array([getattr(libm, fn)(float(x)) for x in a], f32)  # noqa: E731
    sin_t, cos_t, sin_t1, cos_t1 = trig("sinf", th), trig("cosf", th), trig("sinf", th1), trig("cosf", th1)
    sin_p, cos_p, sin_p1, cos_p1 = trig("sinf", ph), trig("cosf", ph), trig("sinf", ph1), trig("cosf", ph1)

    def unit(st, ct, sp, cp):  # (nth-1, nph-1, 3)
        a, b = np.meshgrid(np.arange(nth - 1), np.arange(nph - 1), indexing="ij")
        return np.stack([st[a] * cp[b], st[a] * sp[b], ct[a]], axis=-1).astype(f32)

    q0 = unit(sin_t, cos_t, sin_p, cos_p)
    q1 = unit(sin_t1, cos_t1, sin_p, cos_p)
    q2 = unit(sin_t1, cos_t1, sin_p1, cos_p1)
    q3 = unit(sin_t, cos_t, sin_p1, cos_p1)
    tpl = np.stack([q0, q1, q2, q2, q3, q0], axis=2)  # (nth-1, nph-1, 6, 3)
    return ndivs, tpl.reshape(-1, 3)


def obj_vertex(center, radius, q) -> tuple:
    """One "v" entry of ObjData::write: radius * q + centre in binary32 (ObjData.cpp:84-100)."""
    f32 = np.float32
    c = [f32(x) for x in center]
    r = f32(radius)
    return tuple(float(f32(f32(r * f32(q[a])) + c[a])) for a in range(3))


def stream_float(x: float) -> str:
    """A float or double as an output stream with default flags prints it (%g, precision 6)."""
    return "%g" % x
