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
