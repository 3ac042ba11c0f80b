"""ctypes binding of libparticle_stabilizer_hip.so (C ABI in include/psx.h).

The library is the only compute path: if it is missing this module raises, there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent.parent          # particle-stabilizer_amd/
LIB_PATH = PKG_ROOT / "lib" / "libparticle_stabilizer_hip.so"
BIN_PATH = PKG_ROOT / "bin" / "particle_stabilizer_main"

PSX_OK = 0
PSX_F32 = 0
PSX_F64 = 1
PSX_FLAG_SEMANTICS_CUDA = 1  # psx_params.flags: the reference CUDA backend's step

STATUS_NAMES = {0: "PSX_OK", 1: "PSX_ERR_INVALID", 2: "PSX_ERR_HIP", 3: "PSX_ERR_ALLOC",
                4: "PSX_ERR_STATE", 5: "PSX_ERR_NUMERIC", 6: "PSX_ERR_IO"}


class PsxError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class psx_params(C.Structure):
    _fields_ = [("gravity", C.c_double * 3), ("walls", C.c_double * 6), ("spring", C.c_double),
                ("restitution", C.c_double), ("dt", C.c_double), ("box_min", C.c_double * 3),
                ("box_max", C.c_double * 3), ("dtype", C.c_int32), ("flags", C.c_int32)]


class psx_mesh(C.Structure):
    _fields_ = [("n_triangles", C.c_int64), ("vertices", C.POINTER(C.c_double))]


class psx_scene(C.Structure):
    _fields_ = [("params", psx_params), ("mesh", psx_mesh), ("n", C.c_int64)]


class psx_stats(C.Structure):
    _fields_ = [("kinetic_energy", C.c_double), ("min", C.c_double * 3), ("max", C.c_double * 3),
                ("n_collided", C.c_int64), ("max_overlap", C.c_double)]


class psx_tuning(C.Structure):
    _fields_ = [("block_integrate", C.c_int32), ("block_sort", C.c_int32),
                ("block_pairs", C.c_int32), ("use_graph", C.c_int32),
                ("pair_variant", C.c_int32), ("sort_variant", C.c_int32),
                ("pair_chunk", C.c_int32)]


_DP = C.POINTER(C.c_double)
_I32P = C.POINTER(C.c_int32)
_U32P = C.POINTER(C.c_uint32)
_I64P = C.POINTER(C.c_int64)

# name -> (restype, argtypes)
_SIGNATURES = {
    "psx_create": (C.c_int, [C.POINTER(psx_params), C.c_int64, _DP, _DP, _DP, _DP,
                             C.POINTER(psx_mesh), C.POINTER(C.c_void_p)]),
    "psx_batch_create": (C.c_int, [C.c_int32, C.POINTER(psx_scene), _DP, _DP, _DP, _DP,
                                   C.POINTER(C.c_void_p)]),
    "psx_destroy": (C.c_int, [C.c_void_p]),
    "psx_step": (C.c_int, [C.c_void_p, C.c_int64]),
    "psx_prepare": (C.c_int, [C.c_void_p, C.c_int64]),
    "psx_synchronize": (C.c_int, [C.c_void_p]),
    "psx_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "psx_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(psx_tuning)]),
    "psx_resolve_overlaps": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_int32,
                                       C.POINTER(C.c_int64), C.POINTER(C.c_double)]),
    "psx_write_obj_spheres": (C.c_int, [C.c_char_p, C.c_int64, C.c_void_p, C.c_void_p]),
    "psx_statistics": (C.c_int, [C.c_void_p, C.c_void_p]),
    "psx_statistics_noalloc": (C.c_int, [C.c_void_p, C.c_void_p]),
    "psx_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(psx_tuning)]),
    "psx_get_state": (C.c_int, [C.c_void_p, _DP, _DP]),
    "psx_set_state": (C.c_int, [C.c_void_p, _DP, _DP]),
    "psx_get_state_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "psx_set_state_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "psx_num_particles": (C.c_int64, [C.c_void_p]),
    "psx_num_scenes": (C.c_int32, [C.c_void_p]),
    "psx_num_cells": (C.c_int64, [C.c_void_p]),
    "psx_dtype_of": (C.c_int32, [C.c_void_p]),
    "psx_get_grid": (C.c_int, [C.c_void_p, C.c_int32, _DP, _DP, _I64P, _I64P]),
    "psx_debug_bucketing": (C.c_int, [C.c_void_p, _I32P, _U32P, _U32P, _I32P]),
    "psx_profile": (C.c_int, [C.c_void_p, C.c_int64, _DP, _DP, C.c_int32, _I32P]),
    "psx_kernel_name": (C.c_char_p, [C.c_int32]),
    "psx_set_params": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(psx_params)]),
    "psx_window_forward": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, _I64P]),
    "psx_window_stats": (C.c_int, [C.c_void_p, _DP, _DP]),
    "psx_window_backward": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p, _DP]),
    "psx_read_csv": (C.c_int, [C.c_char_p, C.c_double, _DP, _I64P, C.POINTER(_DP), C.POINTER(_DP),
                               C.POINTER(_DP), C.POINTER(_DP), _I64P]),
    "psx_write_csv": (C.c_int, [C.c_char_p, C.c_int64, _DP, _DP]),
    "psx_read_obj": (C.c_int, [C.c_char_p, _I64P, C.POINTER(_DP)]),
    "psx_free": (None, [C.c_void_p]),
    "psx_last_error": (C.c_char_p, []),
    "psx_abi_version": (C.c_int, []),
}

_lib = None


def exported_symbols() -> list[str]:
    """Entry points this binding expects the library to export."""
    return sorted(_SIGNATURES)


def load(path: str | os.PathLike | None = None) -> C.CDLL:
    """Loads the HIP backend. Raises if it has not been built: there is no CPU fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    # an explicit path, else $PSX_LIBRARY (as the CLI honours it), else the in-tree build
    p = Path(path) if path else Path(os.environ.get("PSX_LIBRARY") or LIB_PATH)
    if not p.exists():
        raise RuntimeError(
            f"HIP backend not built: {p} is missing (run `make -C particle-stabilizer_amd` "
            f"or __graft_entry__.build())")
    lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(code: int) -> None:
    if code != PSX_OK:
        msg = load().psx_last_error()
        raise PsxError(code, msg.decode() if msg else "")
