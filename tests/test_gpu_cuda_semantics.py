"""The PSX_FLAG_SEMANTICS_CUDA mode (the reference CUDA backend's step, PhysicsEngineCUDA.cpp:36-43)
against the oracle's restatement of it (pinned in test_oracle_cuda_pins.py), bit for bit: box and
mesh containers in binary64 and binary32, several launch shapes, graph replay and a batch; the
grid of the mode; the differentiable window refused."""
import dataclasses

import numpy as np
import pytest

from helpers import as_dtype, assert_bitwise, dense_box, hip_run, oracle_run, small_cylinder
from psx import PsxError, scenes

pytestmark = pytest.mark.gpu

PSX_ERR_STATE = 4


def _cuda(sc):
    sc.params = dataclasses.replace(sc.params, semantics="cuda")
    return sc


SHAPES = [dict(),
          dict(block_pairs=128, block_integrate=64, pair_variant=2, sort_variant=2),
          dict(block_pairs=256, block_sort=128, pair_chunk=2, sort_variant=1),
          dict(block_pairs=256, pair_chunk=2, sort_variant=3),  # the CUDA step takes id slots
          dict(block_pairs=128, block_integrate=512, pair_variant=32, sort_variant=1)]


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_cuda_box_bitexact(psx_mod, oracle, dtype, shape):
    """An overlapping lattice in a box with six different wall coefficients: wall hits and
    contacts from the first step."""
    sc = _cuda(dense_box(dtype=dtype))
    steps = 30
    rp, rv = oracle_run(oracle, sc, steps)
    p, v = hip_run(sc, steps, **SHAPES[shape])
    assert_bitwise(as_dtype(p, dtype), rp, "pos")
    assert_bitwise(as_dtype(v, dtype), rv, "vel")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_cuda_mesh_bitexact_eager_and_graph(psx_mod, oracle, dtype):
    """800 spheres thrown outward and down in a 96-triangle open cylinder: wall and floor hits
    (summed corrections, the dot-picked edge points) and pair contacts, 40 steps, launched step by
    step and replayed from graphs."""
    sc = _cuda(small_cylinder(dtype=dtype))
    steps = 40
    rp, rv = oracle_run(oracle, sc, steps)
    for graph in (False, True):
        p, v = hip_run(sc, steps, graph=graph)
        assert_bitwise(as_dtype(p, dtype), rp, f"pos graph={graph}")
        assert_bitwise(as_dtype(v, dtype), rv, f"vel graph={graph}")


def test_cuda_calibration_box_bitexact(psx_mod, oracle):
    """The 80-triangle open box of the calibration workload (8 x 8 x 4 lattice, binary32): the
    lattice lands on the floor mesh after about 900 steps; 1,100 steps."""
    sc = _cuda(scenes.calibration_scene(n_xy=8, n_z=4))
    steps = 1100
    rp, rv = oracle_run(oracle, sc, steps)
    p, v = hip_run(sc, steps, graph=True)
    assert_bitwise(as_dtype(p, "f32"), rp, "pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "vel")
    assert np.asarray(p)[:, 2].min() < sc.pos[:, 2].min() - 0.3  # the bottom layer landed


def test_cuda_differs_from_cpu_and_grid(psx_mod, oracle):
    """The flag reaches the device: the same scene under the two semantics ends apart, and the
    grid width is the CUDA backend's."""
    sc = dense_box(dtype="f64")
    pc, _ = hip_run(sc, 10)
    sc = _cuda(sc)
    pg, _ = hip_run(sc, 10)
    assert not np.array_equal(pc, pg)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        g = s.grid()
    h, n3 = oracle.grid(float(sc.radius.max()), sc.params.box_min, sc.params.box_max,
                        semantics="cuda")
    assert g["h"] == h == float(sc.radius.max()) * 2.0 * 1.05
    assert np.array_equal(g["n"], n3)


def test_cuda_batch_equals_plain_runs(psx_mod, oracle):
    """Three box scenes of different sizes in one batch: each equals its own plain run, and the
    first equals the oracle."""
    scs = [_cuda(dense_box(n_side=k, seed=k, dtype="f32")) for k in (6, 9, 7)]
    steps = 20
    with psx_mod.Batch(scs) as b:
        b.set_tuning(use_graph=1)
        b.step(steps)
        bp, bv = b.state()
    off = np.concatenate([[0], np.cumsum([sc.radius.shape[0] for sc in scs])])
    for k, sc in enumerate(scs):
        p, v = hip_run(sc, steps)
        assert_bitwise(bp[off[k]:off[k + 1]], p, f"scene {k} pos")
        assert_bitwise(bv[off[k]:off[k + 1]], v, f"scene {k} vel")
    rp, rv = oracle_run(oracle, scs[0], steps)
    assert_bitwise(as_dtype(bp[:off[1]], "f32"), rp, "scene 0 pos vs oracle")
    assert_bitwise(as_dtype(bv[:off[1]], "f32"), rv, "scene 0 vel vs oracle")


def test_cuda_window_refused(psx_mod):
    sc = _cuda(dense_box(n_side=4, dtype="f32"))
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        with pytest.raises(PsxError) as e:
            s.window_forward(10, 5)
        assert e.value.code == PSX_ERR_STATE
        s.step(3)  # the handle still steps
        p, _ = s.state()
        assert np.isfinite(p).all()


@pytest.mark.parametrize("workload", ["cylinder_settle", "hopper_settle"])
def test_cuda_full_size_workloads_bitexact(psx_mod, oracle, workload):
    """The 262,144-sphere cylinder (2,176 triangles) and the 4,194,304-sphere hopper (8,192
    triangles) in the CUDA engine's step, binary32, with their tuned launch shapes: 3 steps equal
    the oracle's (fast mode: the results of its reference loops)."""
    import json
    sc = _cuda(getattr(scenes, workload)())
    n = sc.radius.shape[0]
    cache = json.loads(open(scenes.__file__.replace("scenes.py", "tuning_mi355x.json")).read())
    tuning = next(e["tuning"] for k, e in cache.items() if f"|n={n}|" in k)
    steps = 3
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        s.set_tuning(**dict(tuning, use_graph=0))
        s.step(steps)
        p, v = s.state()
    rp, rv = oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                        mesh=sc.mesh, dtype="f32", fast=True, semantics="cuda")
    assert_bitwise(as_dtype(p, "f32"), rp, f"{workload} pos")
    assert_bitwise(as_dtype(v, "f32"), rv, f"{workload} vel")
