"""The incremental bucketing (sort_variant 3): counts kept from one step to the next, only the spheres
that change cell update them, and the rank kernel keeps a sphere's place in a cell that nothing
entered or left. Its results must be those of the full bucketing bit for bit, so every case here
compares against the oracle (or against the full form where the full form is already pinned to the
oracle at that size):

- every pair-kernel shape, eager and graph-replayed, on the dense box (binary64 and binary32);
- a fast gas (most spheres change cell within a few steps, some by more than one cell per step);
- the mesh scene, whose container pushes move spheres out of their bucket cells;
- cells of 300 spheres and more than the eight id slots (arrivals on the overflow chains);
- runs interleaved with statistics passes, state uploads and retunes between the forms, retunes
  that drop a pending hash pass with no scan after it, and the autotuner's sweep;
- the full-size hopper, the contact-rich hopper pressed into its wall, the 256-scene batch and a
  crowded 1M-sphere box;
- the differentiable window's gradients (the window's recompute uses the full form)."""
import numpy as np
import pytest

from helpers import as_dtype, assert_bitwise, dense_box, hip_run, oracle_run, small_cylinder
from psx import scenes
from psx.sim import Params, Scene, mass_of

pytestmark = pytest.mark.gpu

TUNE_CACHE = scenes.__file__.replace("scenes.py", "tuning_mi355x.json")

SHAPES = [dict(block_pairs=256, block_integrate=256),
          dict(block_pairs=128, pair_variant=1, pair_chunk=2, block_sort=64),
          dict(block_pairs=64, block_integrate=1024, pair_variant=4),
          dict(block_pairs=256, block_sort=128, pair_variant=32),
          dict(block_pairs=256, pair_variant=64)]


def fast_gas(dtype="f64", n_side=8, seed=3):
    """Spheres at up to a few thousand times their radius per second in a closed box: with dt 1e-4
    a typical sphere crosses a cell face every few steps and the fastest cross one per step or
    more; the walls push the spheres that run into them back by up to a cell."""
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    k = np.arange(n)
    p = np.stack([k % n_side, (k // n_side) % n_side, k // (n_side * n_side)], 1) * 1.4 + 1.0
    r = rng.uniform(0.3, 0.5, n)
    v = rng.normal(0.0, 3000.0, size=(n, 3))
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.9,) * 6, spring=1e3, restitution=0.8,
                 dt=1e-4, box_min=(0.0, 0.0, 0.0), box_max=(13.0, 13.0, 13.0), dtype=dtype)
    return Scene(p, v, r, prm, mass_of(r))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("shape", range(len(SHAPES)))
def test_box_steps_bitexact(psx_mod, oracle, dtype, shape):
    sc = dense_box(dtype=dtype)
    ref_p, ref_v = oracle_run(oracle, sc, 60)
    kw = dict(SHAPES[shape])
    if dtype == "f64":  # the binary64 rank/gather kernel takes workgroups of at most 512
        kw["block_sort"] = min(kw.get("block_sort", 256), 512)
    for graph in (False, True):
        p, v = hip_run(sc, 60, graph=graph, sort_variant=3, **kw)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos graph={graph}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel graph={graph}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fast_gas_bitexact(psx_mod, oracle, dtype):
    sc = fast_gas(dtype)
    ref_p, ref_v = oracle_run(oracle, sc, 100)
    # most spheres have changed cell over the run
    h = 2.0 * float(as_dtype(sc.radius, dtype).max()) * float(np.float32(1.05))
    moved = (np.floor(ref_p / h) != np.floor(as_dtype(sc.pos, dtype) / h)).any(axis=1).mean()
    assert moved > 0.5, moved
    for graph in (False, True):
        p, v = hip_run(sc, 100, graph=graph, sort_variant=3)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos graph={graph}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel graph={graph}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_mesh_steps_bitexact(psx_mod, oracle, dtype):
    sc = small_cylinder(dtype=dtype)
    ref_p, ref_v = oracle_run(oracle, sc, 150)
    for graph in (False, True):
        for kw in (dict(pair_variant=1, pair_chunk=2), dict(pair_variant=8)):
            p, v = hip_run(sc, 150, graph=graph, sort_variant=3, block_pairs=256, **kw)
            assert_bitwise(as_dtype(p, dtype), ref_p, f"pos graph={graph} {kw}")
            assert_bitwise(as_dtype(v, dtype), ref_v, f"vel graph={graph} {kw}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_crowded_cells_bitexact(psx_mod, oracle, dtype):
    """One cell of 300 spheres and forty of 9-12 (more than the eight id slots): the first pass
    files every sphere as an arrival, overflow chains included; later steps rank the stayers of
    such cells from their old ranges."""
    rng = np.random.default_rng(11)
    h_r = 0.5
    prm = Params(box_min=(0.0, 0.0, 0.0), box_max=(12.0, 12.0, 12.0), dtype=dtype, dt=1e-4,
                 spring=5.0, restitution=0.3, gravity=(0.0, 0.0, -1.0))
    h = h_r * 2.0 * float(np.float32(1.05))
    parts = [rng.uniform(2 * h + 0.01, 3 * h - 0.01, size=(300, 3))]
    for _ in range(40):
        corner = rng.integers(4, 10, size=3) * h
        parts.append(corner + rng.uniform(0.01, h - 0.01, size=(int(rng.integers(9, 13)), 3)))
    parts.append(rng.uniform(0.5, 11.5, size=(500, 3)))
    pos = np.concatenate(parts)
    rng.shuffle(pos)
    n = pos.shape[0]
    r = rng.uniform(0.3, h_r, n)
    r[0] = h_r
    vel = rng.normal(0.0, 30.0, size=(n, 3))  # spheres leave and enter the crowded cells
    sc = Scene(pos, vel, r, prm, mass_of(r))
    ref_p, ref_v = oracle_run(oracle, sc, 12)
    for kw in (dict(pair_variant=1, pair_chunk=2), dict(pair_variant=1, pair_chunk=4),
               dict(pair_variant=32, block_pairs=64)):
        p, v = hip_run(sc, 12, sort_variant=3, **kw)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos {kw}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel {kw}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_large_grid_bitexact(psx_mod, oracle, dtype):
    """A box of 104^3 cells (more than the direct scan's 2^20), where the look-back scan is also the
    full form's."""
    sc = dense_box(dtype=dtype, box=(120.0, 120.0, 120.0))
    ref_p, ref_v = oracle_run(oracle, sc, 20)
    for graph in (False, True):
        p, v = hip_run(sc, 20, graph=graph, sort_variant=3)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos graph={graph}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel graph={graph}")


RUNS = ((16, 3), (16, 1), (64, 3), (83, 3), (5, 1), (16, 3), (3, 2), (64, 3))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_runs_with_retunes_statistics_and_state_uploads(psx_mod, oracle, dtype):
    """Runs of graph-replayed and single steps, switching the cell-list form between runs
    (incremental, id slots, dense lists), with a statistics pass (a full bucketing on the same
    arrays) and a state read-back and upload between runs: the trajectory equals the oracle's, and
    every statistics pass equals that of a handle run with the full form alongside."""
    for sc in (dense_box(dtype=dtype), small_cylinder(n=400, dtype=dtype), fast_gas(dtype)):
        with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s, \
                psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as f:
            f.set_tuning(use_graph=1, sort_variant=1)
            for r, (k, form) in enumerate(RUNS):
                s.set_tuning(use_graph=1, sort_variant=form)
                s.step(k)
                f.step(k)
                s.synchronize()
                ss, fs = s.statistics()[0], f.statistics()[0]
                for key in ss:
                    assert np.array_equal(ss[key], fs[key]), (r, key, ss[key], fs[key])
                if r % 3 == 2:
                    p, v = s.state()
                    s.set_state(p, v)
            p, v = s.state()
            pf, vf = f.state()
        assert_bitwise(p, pf, "pos vs full form")
        assert_bitwise(v, vf, "vel vs full form")
        ref_p, ref_v = oracle_run(oracle, sc, sum(k for k, _ in RUNS))
        assert_bitwise(as_dtype(p, dtype), ref_p, "pos vs oracle")
        assert_bitwise(as_dtype(v, dtype), ref_v, "vel vs oracle")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_retunes_between_runs_without_a_scan(psx_mod, oracle, dtype):
    """Retunes between short runs of the fast gas, each dropping the last step's pending hash pass
    (many movers) with no scan after it, as the autotuner's rounds do: every first pass after a
    drop takes a generation of its own, so the arrivals of the dropped pass are not counted."""
    sc = fast_gas(dtype)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        for r in range(8):
            kw = dict(SHAPES[r % len(SHAPES)])
            if dtype == "f64":
                kw["block_sort"] = min(kw.get("block_sort", 256), 512)
            s.set_tuning(use_graph=r % 2, sort_variant=3, **kw)
            s.step(5)
        s.synchronize()
        p, v = s.state()
    ref_p, ref_v = oracle_run(oracle, sc, 40)
    assert_bitwise(as_dtype(p, dtype), ref_p, "pos")
    assert_bitwise(as_dtype(v, dtype), ref_v, "vel")


def test_autotune_rounds_then_run_bitexact(psx_mod, oracle):
    """The autotuner's sweep (every launch shape of the three cell-list forms, a retune per round)
    on the dense box, then a run from the restored state with the chosen shape."""
    from psx.tune import autotune
    sc = dense_box(dtype="f32")
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        out = autotune(s, steps=2)
        assert out["kernel_ms"]["step_incremental"] is not None
        s.step(30)
        s.synchronize()
        p, v = s.state()
    ref_p, ref_v = oracle_run(oracle, sc, 30)
    assert_bitwise(as_dtype(p, "f32"), ref_p, "pos")
    assert_bitwise(as_dtype(v, "f32"), ref_v, "vel")


def test_batch_equals_full_form(psx_mod, oracle):
    """A batch of three dense boxes and a fast gas: the incremental form's batch equals the full
    form's batch and the oracle, eager and replayed."""
    scs = [dense_box(n_side=6, seed=s, dtype="f32") for s in range(3)] + [fast_gas("f32", 7)]
    steps = 40
    got = {}
    for form in (1, 3):
        for graph in (0, 1):
            with psx_mod.Batch(scs) as b:
                b.set_tuning(use_graph=graph, sort_variant=form)
                b.step(steps)
                b.synchronize()
                got[form, graph] = b.state()
    for key, (p, v) in got.items():
        assert_bitwise(p, got[1, 0][0], f"pos {key}")
        assert_bitwise(v, got[1, 0][1], f"vel {key}")
    sizes = np.array([sc.radius.shape[0] for sc in scs])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    for k, sc in enumerate(scs):
        rp, rv = oracle_run(oracle, sc, steps)
        assert_bitwise(as_dtype(got[3, 1][0][offs[k]:offs[k + 1]], "f32"), rp, f"scene {k} pos")
        assert_bitwise(as_dtype(got[3, 1][1][offs[k]:offs[k + 1]], "f32"), rv, f"scene {k} vel")


def test_window_gradients_equal_full_form(psx_mod):
    """A differentiable window stepped with the incremental form (the mesh scene of the adjoint
    tests) gives the full form's final state, loss and gradients bit for bit: the backward's
    recompute buckets with the full form."""
    from test_gpu_adjoint import _hip, _weights
    from test_torch_step import adjoint_scene
    sc = adjoint_scene(True)
    wx, wv = _weights(sc.radius.shape[0])
    full = _hip(psx_mod, sc, wx, wv, sort_variant=1)
    inc = _hip(psx_mod, sc, wx, wv, sort_variant=3)
    inc_graph = _hip(psx_mod, sc, wx, wv, graph=True, sort_variant=3)
    for key in ("x0", "v0", "k", "es", "gz", "loss"):
        assert np.array_equal(np.asarray(inc[key]), np.asarray(full[key])), key
        assert np.array_equal(np.asarray(inc_graph[key]), np.asarray(full[key])), key


def _tuned(sim):
    from psx.tune import autotune
    return autotune(sim, cache=TUNE_CACHE)["tuning"]


def _oracle(oracle, sc, steps):
    return oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                      mesh=sc.mesh, dtype=sc.params.dtype, fast=True)


def test_hopper_lattice_and_wall_bitexact(psx_mod, oracle):
    """The 4,194,304-sphere hopper with its tuned pair shape and the incremental form: 3 eager steps
    from the lattice; then the contact-rich state pressed into the cone wall, one eager step and
    16 replayed from a prepared graph."""
    from test_gpu_configs import _hopper_against_the_wall
    sc = scenes.hopper_settle()
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        _tuned(s)
        s.set_tuning(use_graph=0, sort_variant=3)
        s.step(3)
        p, v = s.state()
    rp, rv = _oracle(oracle, sc, 3)
    assert_bitwise(as_dtype(p, "f32"), rp, "hopper lattice pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "hopper lattice vel")
    sc, _ = _hopper_against_the_wall()
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        _tuned(s)
        s.set_tuning(use_graph=1, sort_variant=3)
        s.step(1)
        s.prepare(16)
        s.step(16)
        p, v = s.state()
    rp, rv = _oracle(oracle, sc, 17)
    assert_bitwise(as_dtype(p, "f32"), rp, "hopper wall pos, 17 steps")
    assert_bitwise(as_dtype(v, "f32"), rv, "hopper wall vel, 17 steps")


def test_full_batch_equals_full_form(psx_mod):
    """All 256 box scenes (4.6M spheres, 35.5M cells), 16 steps from a prepared graph: the
    incremental form's state equals the full form's bit for bit (the full form's batch is pinned to
    the plain runs and the oracle at this size)."""
    scs = scenes.batched_boxes()
    got = []
    for form in (1, 3):
        with psx_mod.Batch(scs) as b:
            _tuned(b)
            b.set_tuning(use_graph=1, sort_variant=form)
            b.prepare(16)
            b.step(16)
            b.synchronize()
            got.append(b.state())
    assert_bitwise(got[1][0], got[0][0], "batch pos")
    assert_bitwise(got[1][1], got[0][1], "batch vel")


def test_crowded_box_bitexact(psx_mod, oracle):
    from test_gpu_configs import crowded_box
    sc = crowded_box()
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        s.set_tuning(sort_variant=3)
        s.step(3)
        p, v = s.state()
    rp, rv = _oracle(oracle, sc, 3)
    assert_bitwise(as_dtype(p, "f32"), rp, "crowded pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "crowded vel")


def test_full_batch_2000_steps_forms_equal_and_oracle(psx_mod, oracle):
    """All 256 box scenes for 2,000 steps, graph-replayed: the incremental and the full form stay
    bit-identical over the long run, and scene 80 (23,927 spheres) equals the oracle's 2,000 steps.
    (Past ~2,250 steps the reference's explicit kick blows up in several of these scenes, in the oracle
    as here, so the run stops before.)"""
    scs = scenes.batched_boxes()
    steps = 2000
    got = []
    for form in (1, 3):
        with psx_mod.Batch(scs) as b:
            _tuned(b)
            b.set_tuning(use_graph=1, sort_variant=form)
            b.prepare(steps)
            b.step(steps)
            b.synchronize()
            got.append(b.state())
    assert_bitwise(got[1][0], got[0][0], "batch pos, incremental vs full")
    assert_bitwise(got[1][1], got[0][1], "batch vel, incremental vs full")
    sizes = np.array([sc.radius.shape[0] for sc in scs])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    k = 80
    rp, rv = _oracle(oracle, scs[k], steps)
    assert_bitwise(as_dtype(got[1][0][offs[k]:offs[k + 1]], "f32"), rp, "scene 80 pos vs oracle")
    assert_bitwise(as_dtype(got[1][1][offs[k]:offs[k + 1]], "f32"), rv, "scene 80 vel vs oracle")


def test_hopper_2000_steps_forms_equal(psx_mod):
    """The 4,194,304-sphere hopper for 2,000 steps with its tuned shape, replayed from prepared graphs:
    the incremental form's state equals the full form's bit for bit (the full form's hopper steps are
    pinned to the oracle in test_gpu_configs.py)."""
    sc = scenes.hopper_settle()
    steps = 2000
    got = []
    for form in (1, 3):
        with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
            _tuned(s)
            s.set_tuning(use_graph=1, sort_variant=form)
            s.prepare(steps)
            s.step(steps)
            s.synchronize()
            got.append(s.state())
    assert_bitwise(got[1][0], got[0][0], "hopper pos, incremental vs full")
    assert_bitwise(got[1][1], got[0][1], "hopper vel, incremental vs full")
