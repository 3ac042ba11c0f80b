"""The C ABI of libparticle_stabilizer_hip.so (include/psx.h), checked without a GPU: the library
loads, exports every declared entry point, its structs have the C layout, and invalid arguments are
rejected with a status code (never exit(), never a crash)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from psx import _lib
from psx.sim import Params

REPO = Path(__file__).resolve().parent.parent
HEADER = REPO / "include" / "psx.h"


def declared_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int32_t|int64_t|void|const char\*)\s+(psx_\w+)\(",
                                 text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    assert sorted(set(names)) == sorted(set(_lib.exported_symbols()))
    version = int(re.search(r"#define PSX_ABI_VERSION (\d+)", HEADER.read_text()).group(1))
    assert lib.psx_abi_version() == version


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "psx.h"\n'
                   "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\","
                   "sizeof(psx_params), sizeof(psx_mesh), sizeof(psx_scene), sizeof(psx_tuning),"
                   "offsetof(psx_params, dt), offsetof(psx_params, dtype),"
                   "offsetof(psx_scene, n));return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(REPO / "include"), str(src), "-o", str(exe)],
                   check=True)
    got = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                       check=True).stdout.split()))
    exp = [C.sizeof(_lib.psx_params), C.sizeof(_lib.psx_mesh), C.sizeof(_lib.psx_scene),
           C.sizeof(_lib.psx_tuning), _lib.psx_params.dt.offset, _lib.psx_params.dtype.offset,
           _lib.psx_scene.n.offset]
    assert got == exp


def _create(params: Params, n=4, radius=0.5, mass=1.0, mesh=None):
    lib = _lib.load()
    pos = np.zeros((n, 3)) + 1.0
    vel = np.zeros((n, 3))
    r = np.full(n, radius)
    m = np.full(n, mass)
    h = C.c_void_p()
    cp = params.to_c()
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    rc = lib.psx_create(C.byref(cp), n, dp(pos), dp(vel), dp(r), dp(m), mesh, C.byref(h))
    return rc, h, lib.psx_last_error().decode()


@pytest.mark.parametrize("kw,frag", [
    ({"dt": 0.0}, "dt"),
    ({"dt": -1e-3}, "dt"),
    ({"dt": float("nan")}, "dt"),
    ({"dtype": 7}, "dtype"),
    ({"box_min": (0, 0, 5), "box_max": (1, 1, 5)}, "box"),
    ({"gravity": (0, float("inf"), 0)}, "gravity"),
])
def test_invalid_params_rejected(kw, frag):
    rc, h, err = _create(Params(**kw))
    assert rc == 1 and frag in err and not h.value


def test_invalid_particles_rejected():
    rc, h, err = _create(Params(), n=0)
    assert rc == 1 and not h.value
    rc, h, err = _create(Params(), radius=-0.5)
    assert rc == 1 and "radii" in err
    rc, h, err = _create(Params(), mass=0.0)
    assert rc == 1 and "mass" in err
    m = _lib.psx_mesh()
    m.n_triangles = 3
    m.vertices = None
    rc, h, err = _create(Params(), mesh=C.byref(m))
    assert rc == 1 and "mesh" in err
    # triangle indices are int32 on the device: a count of 2^31 is refused before any vertex is read
    one = np.zeros(9)
    m.n_triangles = 1 << 31
    m.vertices = one.ctypes.data_as(C.POINTER(C.c_double))
    rc, h, err = _create(Params(), mesh=C.byref(m))
    assert rc == 1 and "triangles" in err and not h.value


def test_null_handles_rejected():
    lib = _lib.load()
    assert lib.psx_step(None, 1) == 1
    assert lib.psx_get_state(None, None, None) == 1
    assert lib.psx_destroy(None) == 0
    assert lib.psx_num_particles(None) == -1
    assert lib.psx_kernel_name(0) == b"k_integrate"
    assert lib.psx_kernel_name(3) == b"k_pairs"
    assert lib.psx_kernel_name(4) == b"k_scatter_ids"
    assert lib.psx_kernel_name(5) is None
    assert lib.psx_kernel_name(99) is None


def test_semantics_flags_checked():
    """psx_params.flags takes PSX_FLAG_SEMANTICS_CUDA and nothing else, and a batch's scenes must
    agree on it (both refused before any device work)."""
    assert _lib.PSX_FLAG_SEMANTICS_CUDA == 1
    assert Params(semantics="cuda").to_c().flags == 1 and Params().to_c().flags == 0
    with pytest.raises(ValueError):
        Params(semantics="gpu").to_c()
    lib = _lib.load()
    n = 4
    pos = np.zeros((2 * n, 3)) + 1.0
    vel = np.zeros((2 * n, 3))
    r = np.full(2 * n, 0.5)
    m = np.full(2 * n, 1.0)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    bad = Params().to_c()
    bad.flags = 2
    h = C.c_void_p()
    rc = lib.psx_create(C.byref(bad), n, dp(pos), dp(vel), dp(r), dp(m), None, C.byref(h))
    assert rc == 1 and "flags" in lib.psx_last_error().decode() and not h.value
    scenes = (_lib.psx_scene * 2)()
    scenes[0].params = Params(semantics="cuda").to_c()
    scenes[1].params = Params().to_c()
    scenes[0].n = scenes[1].n = n
    rc = lib.psx_batch_create(2, scenes, dp(pos), dp(vel), dp(r), dp(m), C.byref(h))
    assert rc == 1 and "semantics" in lib.psx_last_error().decode() and not h.value
