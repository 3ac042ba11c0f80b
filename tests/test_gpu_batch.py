"""Batched scenes: one launch per kernel per step over scenes of uneven size, each with its own
constants. Every scene's trajectory must equal its plain single-scene run bit for bit, and the
oracle's."""
import numpy as np
import pytest

from helpers import as_dtype, assert_bitwise, dense_box, hip_run, oracle_run, small_cylinder
from psx.sim import Params, Scene, mass_of

pytestmark = pytest.mark.gpu


def _scenes(dtype):
    out = []
    for k, (n_side, spring, es) in enumerate([(5, 3e3, 0.3), (9, 1e4, 0.5), (3, 500.0, 0.9),
                                              (7, 2e3, 0.7)]):
        sc = dense_box(n_side=n_side, seed=10 + k, dtype=dtype, spring=spring, es=es,
                       box=(n_side + 2.0, n_side + 1.0, n_side + 3.0))
        out.append(sc)
    out.append(small_cylinder(n=300, seed=5, dtype=dtype))
    return out


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("graph,staged", [(False, False), (True, False), (True, True)])
def test_batch_equals_plain_and_oracle(psx_mod, oracle, dtype, graph, staged):
    scenes = _scenes(dtype)
    steps = 40
    with psx_mod.Batch(scenes) as b:
        if graph:
            b.set_tuning(use_graph=1)
        if staged:  # the LDS-staged pair kernel on the scenes' cell segments
            b.set_tuning(pair_variant=32, block_pairs=128)
        b.step(steps)
        b.synchronize()
        bp, bv = b.state()
        assert b.num_scenes == len(scenes)
    off = 0
    for k, sc in enumerate(scenes):
        n = sc.radius.shape[0]
        p, v = hip_run(sc, steps)
        assert_bitwise(bp[off:off + n], p, f"scene {k} pos vs plain")
        assert_bitwise(bv[off:off + n], v, f"scene {k} vel vs plain")
        rp, rv = oracle_run(oracle, sc, steps)
        assert_bitwise(as_dtype(bp[off:off + n], dtype), rp, f"scene {k} pos vs oracle")
        assert_bitwise(as_dtype(bv[off:off + n], dtype), rv, f"scene {k} vel vs oracle")
        off += n


def test_batch_bucketing_segments(psx_mod, oracle):
    scenes = _scenes("f32")
    with psx_mod.Batch(scenes) as b:
        got = b.bucketing()
        off = 0
        for k, sc in enumerate(scenes):
            n = sc.radius.shape[0]
            g = b.grid(k)
            if sc.mesh is not None:
                v = sc.mesh.reshape(-1, 3)
                mn, mx = v.min(0), v.max(0)
            else:
                mn, mx = sc.params.box_min, sc.params.box_max
            ref = oracle.bucketing(float(np.float32(sc.radius).max()), mn, mx,
                                   np.float32(sc.pos), dtype="f32")
            base = g["cell_base"]
            nc = int(np.prod(g["n"]))
            assert_bitwise(got["cell"][off:off + n] - base, ref["cell"], f"scene {k} cell")
            assert_bitwise(got["count"][base:base + nc], ref["count"], f"scene {k} count")
            seg = got["offset"][base:base + nc + 1] - got["offset"][base]
            assert_bitwise(seg, ref["offset"], f"scene {k} offset")
            assert_bitwise(got["perm"][off:off + n] - off, ref["perm"], f"scene {k} perm")
            off += n


def _crowded_scene(dtype):
    """A box with one cell of 300 spheres and 40 cells of 9-12 (as in the overflow parity test)."""
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
    vel = rng.normal(0.0, 0.1, size=(n, 3))
    return Scene(pos, vel, r, prm, mass_of(r))


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("chunk", [4, 2])
def test_batch_crowded_scene_queue_tail(psx_mod, oracle, dtype, chunk):
    """The crowded scene between two dense boxes in one batch: in its 300-sphere cell the pair
    kernel's per-lane queue of near candidates fills and the lanes test the rest in line, over the
    workgroup table. Each scene equals its plain run and the oracle bit for bit."""
    scenes = [dense_box(n_side=5, dtype=dtype), _crowded_scene(dtype), dense_box(n_side=4, seed=3, dtype=dtype)]
    steps = 3
    with psx_mod.Batch(scenes) as b:
        b.set_tuning(pair_variant=1, pair_chunk=chunk, block_pairs=128)
        b.step(steps)
        b.synchronize()
        bp, bv = b.state()
    off = 0
    for k, sc in enumerate(scenes):
        n = sc.radius.shape[0]
        p, v = hip_run(sc, steps, pair_variant=1, pair_chunk=chunk, block_pairs=128)
        assert_bitwise(bp[off:off + n], p, f"scene {k} pos vs plain")
        assert_bitwise(bv[off:off + n], v, f"scene {k} vel vs plain")
        rp, rv = oracle_run(oracle, sc, steps)
        assert_bitwise(as_dtype(bp[off:off + n], dtype), rp, f"scene {k} pos vs oracle")
        assert_bitwise(as_dtype(bv[off:off + n], dtype), rv, f"scene {k} vel vs oracle")
        off += n
