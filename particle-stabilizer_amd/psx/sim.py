"""Host-side mirror of particle-stabilizer's simulator interface over the HIP backend.

`run_simulate` keeps the argument list and meaning of the reference's backend entry
(`runSimulate` / `runSimulateCUDA`, include/ParticleStabilizer/PhysicsEngine.hpp:31-42,
include/ParticleStabilizerCUDA/PhysicsEngineCUDA.hpp:43-54). `Simulation` and `Batch` expose the
stepping, state and bucketing entry points of include/psx.h.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _lib
from ._lib import (PSX_F32, PSX_F64, PSX_FLAG_SEMANTICS_CUDA, check, psx_mesh, psx_params,
                   psx_scene, psx_tuning)

_DTYPES = {"f32": PSX_F32, "float32": PSX_F32, "f64": PSX_F64, "float64": PSX_F64}


def _dtype_code(dtype) -> int:
    if isinstance(dtype, int):
        return dtype
    if dtype in (np.float32,):
        return PSX_F32
    if dtype in (np.float64,):
        return PSX_F64
    return _DTYPES[str(dtype)]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _c3(a: np.ndarray, n: int, name: str) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n, 3):
        raise ValueError(f"{name} must have shape ({n}, 3), got {a.shape}")
    return a


@dataclass
class Params:
    """Simulation constants (CLI defaults of ParticleStabilizerMain.cpp:41-57)."""

    gravity: Sequence[float] = (0.0, 0.0, -10.0)
    # CoefficientOfRestitutionWall order: zMin, zMax, xMin, xMax, yMin, yMax
    walls: Sequence[float] = (0.5, 0.5, 0.5, 0.5, 0.5, 0.5)
    spring: float = 1.0
    restitution: float = 0.1
    dt: float = 1e-5
    box_min: Sequence[float] = (0.0, 0.0, 0.0)
    box_max: Sequence[float] = (10.0, 20.0, 30.0)
    dtype: str = "f64"
    # "cpu": the reference CPU backend's step (PhysicsEngine.cpp:38-48); "cuda": its CUDA backend's
    # (PhysicsEngineCUDA.cpp:36-43), PSX_FLAG_SEMANTICS_CUDA
    semantics: str = "cpu"

    def to_c(self) -> psx_params:
        p = psx_params()
        p.gravity[:] = [float(x) for x in self.gravity]
        p.walls[:] = [float(x) for x in self.walls]
        p.spring = float(self.spring)
        p.restitution = float(self.restitution)
        p.dt = float(self.dt)
        p.box_min[:] = [float(x) for x in self.box_min]
        p.box_max[:] = [float(x) for x in self.box_max]
        p.dtype = _dtype_code(self.dtype)
        if self.semantics not in ("cpu", "cuda"):
            raise ValueError(f"semantics must be 'cpu' or 'cuda', not {self.semantics!r}")
        p.flags = PSX_FLAG_SEMANTICS_CUDA if self.semantics == "cuda" else 0
        return p

    def as_array(self) -> np.ndarray:
        """The 18 doubles gravity[3], walls[6], spring, restitution, dt, box_min[3], box_max[3]."""
        return np.array(list(self.gravity) + list(self.walls) +
                        [self.spring, self.restitution, self.dt] +
                        list(self.box_min) + list(self.box_max), dtype=np.float64)


def mass_of(radius, density: float = 1.0) -> np.ndarray:
    """Mass of ParticleData::readFromCSV (ParticleIO.cpp:131): density*4/3*pi*r^3, same order."""
    r = np.asarray(radius, dtype=np.float64)
    return density * 4.0 / 3.0 * np.pi * np.power(r, 3.0)


class Simulation:
    """One handle of the HIP backend (a single scene, or a batch built by `Batch`)."""

    def __init__(self, pos, vel, radius, mass=None, params: Params | None = None,
                 mesh: np.ndarray | None = None, *, _handle=None, _n=None):
        self._lib = _lib.load()
        self._h = C.c_void_p()
        if _handle is not None:
            self._h = _handle
            self.n = int(_n)
            self._keep = None
            self.params = params
            return
        params = params or Params()
        radius = np.ascontiguousarray(radius, dtype=np.float64).reshape(-1)
        n = radius.shape[0]
        pos = _c3(pos, n, "pos")
        vel = _c3(vel, n, "vel")
        mass = mass_of(radius) if mass is None else np.ascontiguousarray(mass, dtype=np.float64)
        cp = params.to_c()
        mptr = None
        if mesh is not None:
            mesh = np.ascontiguousarray(mesh, dtype=np.float64).reshape(-1, 9)
            m = psx_mesh()
            m.n_triangles = mesh.shape[0]
            m.vertices = _dptr(mesh)
            mptr = C.byref(m)
        check(self._lib.psx_create(C.byref(cp), n, _dptr(pos), _dptr(vel), _dptr(radius),
                                   _dptr(mass), mptr, C.byref(self._h)))
        self.n = n
        self.params = params
        self._keep = (mesh,)

    # -- lifetime --------------------------------------------------------------------------------
    def close(self) -> None:
        if self._h and self._h.value:
            check(self._lib.psx_destroy(self._h))
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- stepping --------------------------------------------------------------------------------
    def step(self, nsteps: int = 1) -> None:
        check(self._lib.psx_step(self._h, int(nsteps)))

    def prepare(self, nsteps: int) -> None:
        """Captures the graphs a later step(nsteps) replays from the current state (no step)."""
        check(self._lib.psx_prepare(self._h, int(nsteps)))

    def synchronize(self) -> None:
        check(self._lib.psx_synchronize(self._h))

    def set_stream(self, stream_ptr: int | None) -> None:
        check(self._lib.psx_set_stream(self._h, C.c_void_p(stream_ptr or 0)))

    @property
    def tuning(self) -> dict:
        t = psx_tuning()
        check(self._lib.psx_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in psx_tuning._fields_}

    def set_tuning(self, **kw) -> None:
        cur = self.tuning
        cur.update(kw)
        t = psx_tuning(**cur)
        check(self._lib.psx_set_tuning(self._h, C.byref(t)))

    # -- state -----------------------------------------------------------------------------------
    def state(self) -> tuple[np.ndarray, np.ndarray]:
        pos = np.empty((self.n, 3), np.float64)
        vel = np.empty((self.n, 3), np.float64)
        check(self._lib.psx_get_state(self._h, _dptr(pos), _dptr(vel)))
        return pos, vel

    def set_state(self, pos, vel) -> None:
        pos = _c3(pos, self.n, "pos")
        vel = _c3(vel, self.n, "vel")
        check(self._lib.psx_set_state(self._h, _dptr(pos), _dptr(vel)))

    def get_state_device(self, pos_ptr: int, vel_ptr: int) -> None:
        check(self._lib.psx_get_state_device(self._h, C.c_void_p(pos_ptr), C.c_void_p(vel_ptr)))

    def set_state_device(self, pos_ptr: int, vel_ptr: int) -> None:
        check(self._lib.psx_set_state_device(self._h, C.c_void_p(pos_ptr), C.c_void_p(vel_ptr)))

    @property
    def dtype(self):
        return np.float32 if self._lib.psx_dtype_of(self._h) == PSX_F32 else np.float64

    @property
    def num_cells(self) -> int:
        return int(self._lib.psx_num_cells(self._h))

    @property
    def num_scenes(self) -> int:
        return int(self._lib.psx_num_scenes(self._h))

    def grid(self, scene: int = 0) -> dict:
        mn = np.zeros(3)
        h = C.c_double()
        n3 = np.zeros(3, np.int64)
        base = C.c_int64()
        check(self._lib.psx_get_grid(self._h, scene, _dptr(mn), C.byref(h),
                                     n3.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(base)))
        return {"min": mn, "h": h.value, "n": n3, "cell_base": base.value}

    def bucketing(self) -> dict:
        """Buckets the current positions (no step) and returns cell, count, offset, perm."""
        nc = self.num_cells
        cell = np.empty(self.n, np.int32)
        count = np.empty(nc, np.uint32)
        offset = np.empty(nc + 1, np.uint32)
        perm = np.empty(self.n, np.int32)
        check(self._lib.psx_debug_bucketing(
            self._h, cell.ctypes.data_as(C.POINTER(C.c_int32)),
            count.ctypes.data_as(C.POINTER(C.c_uint32)),
            offset.ctypes.data_as(C.POINTER(C.c_uint32)),
            perm.ctypes.data_as(C.POINTER(C.c_int32))))
        return {"cell": cell, "count": count, "offset": offset, "perm": perm}

    def set_params(self, params: "Params", scene: int = 0) -> None:
        """Replaces a scene's coefficients (gravity, walls, spring, restitution, dt)."""
        cp = params.to_c()
        check(self._lib.psx_set_params(self._h, scene, C.byref(cp)))
        if scene == 0:
            self.params = params

    def window_forward(self, nsteps: int, checkpoint_every: int = 100) -> int:
        """Runs a window of nsteps steps with checkpoints; returns the window's id."""
        wid = C.c_int64()
        check(self._lib.psx_window_forward(self._h, int(nsteps), int(checkpoint_every),
                                           C.byref(wid)))
        return wid.value

    def window_backward(self, window_id: int, dl_dpos_ptr: int, dl_dvel_ptr: int,
                        out_pos_ptr: int, out_vel_ptr: int) -> np.ndarray:
        """Adjoint of window `window_id` (the latest one). Device pointers (n x 3, the handle's
        dtype); returns dL/d(spring, restitution, g_z) per scene as an (n_scenes, 3) array."""
        g = np.zeros(3 * self.num_scenes)
        check(self._lib.psx_window_backward(self._h, int(window_id), C.c_void_p(dl_dpos_ptr),
                                            C.c_void_p(dl_dvel_ptr), C.c_void_p(out_pos_ptr),
                                            C.c_void_p(out_vel_ptr), _dptr(g)))
        return g.reshape(-1, 3)

    def window_stats(self) -> dict:
        """Device time (ms) of the last backward: segment recompute and adjoint steps."""
        a, b = C.c_double(), C.c_double()
        check(self._lib.psx_window_stats(self._h, C.byref(a), C.byref(b)))
        return {"replay_ms": a.value, "adjoint_ms": b.value}

    def statistics(self) -> list[dict]:
        """ParticleStatistics::calculateStatistics of every scene (ParticleStatistics.cpp:3-43)."""
        from ._lib import psx_stats
        out = (psx_stats * self.num_scenes)()
        check(self._lib.psx_statistics(self._h, C.cast(out, C.c_void_p)))
        return [{"kinetic_energy": o.kinetic_energy, "min": np.array(o.min[:]),
                 "max": np.array(o.max[:]), "n_collided": int(o.n_collided),
                 "max_overlap": o.max_overlap} for o in out]

    def statistics_noalloc(self) -> list[dict]:
        """statistics() through psx_statistics_noalloc (no device allocation; fixed scratch)."""
        from ._lib import psx_stats
        out = (psx_stats * self.num_scenes)()
        check(self._lib.psx_statistics_noalloc(self._h, C.cast(out, C.c_void_p)))
        return [{"kinetic_energy": o.kinetic_energy, "min": np.array(o.min[:]),
                 "max": np.array(o.max[:]), "n_collided": int(o.n_collided),
                 "max_overlap": o.max_overlap} for o in out]

    def profile(self, nsteps: int) -> dict:
        """Mean device time (ms) per launch of each kernel over nsteps steps (HIP events)."""
        ms = np.zeros(32)
        per = np.zeros(32)
        nk = C.c_int32()
        check(self._lib.psx_profile(self._h, int(nsteps), _dptr(ms), _dptr(per), 32, C.byref(nk)))
        out = {}
        for k in range(nk.value):
            out[self._lib.psx_kernel_name(k).decode()] = {"ms": ms[k], "per_step": per[k]}
        return out


@dataclass
class Scene:
    """One scene of a batch."""

    pos: np.ndarray
    vel: np.ndarray
    radius: np.ndarray
    params: Params
    mass: np.ndarray | None = None
    mesh: np.ndarray | None = None
    meta: dict = field(default_factory=dict)


class Batch(Simulation):
    """Independent scenes advanced together with one launch per kernel per step."""

    def __init__(self, scenes: Sequence[Scene]):
        lib = _lib.load()
        ns = len(scenes)
        arr = (psx_scene * ns)()
        keep = []
        pos, vel, rad, mas = [], [], [], []
        for s, sc in enumerate(scenes):
            r = np.ascontiguousarray(sc.radius, np.float64).reshape(-1)
            n = r.shape[0]
            arr[s].params = sc.params.to_c()
            arr[s].n = n
            if sc.mesh is not None:
                m = np.ascontiguousarray(sc.mesh, np.float64).reshape(-1, 9)
                keep.append(m)
                arr[s].mesh.n_triangles = m.shape[0]
                arr[s].mesh.vertices = _dptr(m)
            pos.append(_c3(sc.pos, n, "pos"))
            vel.append(_c3(sc.vel, n, "vel"))
            rad.append(r)
            mas.append(mass_of(r) if sc.mass is None else np.asarray(sc.mass, np.float64))
        pos = np.ascontiguousarray(np.concatenate(pos))
        vel = np.ascontiguousarray(np.concatenate(vel))
        rad = np.ascontiguousarray(np.concatenate(rad))
        mas = np.ascontiguousarray(np.concatenate(mas))
        h = C.c_void_p()
        check(lib.psx_batch_create(ns, arr, _dptr(pos), _dptr(vel), _dptr(rad), _dptr(mas),
                                   C.byref(h)))
        super().__init__(None, None, None, params=scenes[0].params, _handle=h, _n=rad.shape[0])
        self.scene_sizes = np.array([len(np.asarray(s.radius).reshape(-1)) for s in scenes])
        self._keep = keep


def run_simulate(particles, box_min, box_max, container, gravity, walls, spring,
                 restitution_sphere, dt, max_steps, to_visualize=False, *, dtype="f64"):
    """The reference's backend entry (runSimulate / runSimulateCUDA, PhysicsEngine.hpp:31-42).

    particles: dict or object with `position` (n,3), `velocity` (n,3), `radius` (n,), `mass` (n,)
    container: None / "" for the box, an (T,9) triangle array, or a path to an OBJ file.
    Returns the resulting particles as a dict of arrays (resultParticles).
    """
    if to_visualize:
        raise ValueError("visualization is not part of this backend")
    get = (lambda k: particles[k]) if isinstance(particles, dict) else (lambda k: getattr(particles, k))
    mesh = None
    if container is not None and not (isinstance(container, str) and container == ""):
        if isinstance(container, (str, bytes)) or hasattr(container, "__fspath__"):
            from .io import read_obj
            mesh = read_obj(container)
        else:
            mesh = np.asarray(container, np.float64)
    prm = Params(gravity=gravity, walls=walls, spring=spring, restitution=restitution_sphere,
                 dt=dt, box_min=box_min, box_max=box_max, dtype=dtype)
    with Simulation(get("position"), get("velocity"), get("radius"), get("mass"), prm, mesh) as s:
        s.step(int(max_steps))
        s.synchronize()
        pos, vel = s.state()
    return {"position": pos, "velocity": vel, "radius": np.asarray(get("radius"), np.float64),
            "mass": np.asarray(get("mass"), np.float64)}
