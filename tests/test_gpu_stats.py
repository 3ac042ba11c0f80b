"""Statistics of the HIP backend (GPU, O(N) on the grid) against the oracle's O(N^2) restatement of
ParticleStatistics::calculateStatistics: AABB, collided count and maximum overlap exact; the
kinetic energy (a parallel sum in a fixed order, not the reference's sequential one) to 1e-12."""
import numpy as np
import pytest

from helpers import as_dtype, dense_box, small_cylinder

pytestmark = pytest.mark.gpu


def _check(sim, ref_stats, dtype, rtol_ke=1e-12):
    got = sim.statistics()
    assert len(got) == len(ref_stats)
    for g, r in zip(got, ref_stats):
        assert g["n_collided"] == r["n_collided"]
        assert g["max_overlap"] == r["max_overlap"]
        np.testing.assert_array_equal(g["min"], r["min"])
        np.testing.assert_array_equal(g["max"], r["max"])
        assert abs(g["kinetic_energy"] - r["kinetic_energy"]) <= rtol_ke * abs(r["kinetic_energy"])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("scene", ["box", "cylinder"])
def test_statistics_match_oracle(psx_mod, oracle, dtype, scene):
    sc = dense_box(dtype=dtype) if scene == "box" else small_cylinder(dtype=dtype)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        s.step(40)  # a packed state with overlaps
        pos, vel = s.state()
        ref = oracle.statistics(as_dtype(pos, dtype), as_dtype(vel, dtype),
                                as_dtype(sc.radius, dtype), as_dtype(sc.mass, dtype))
        assert ref["n_collided"] > 0
        _check(s, [ref], dtype)
        # the bucketing pass of the statistics leaves the trajectory unchanged
        p1, v1 = s.state()
        np.testing.assert_array_equal(p1, pos)
        np.testing.assert_array_equal(v1, vel)


def test_statistics_per_scene_of_a_batch(psx_mod, oracle):
    from psx.sim import Scene
    scs = [dense_box(n_side=n, dtype="f64", seed=k) for k, n in enumerate((5, 7, 4))]
    with psx_mod.Batch([Scene(sc.pos, sc.vel, sc.radius, sc.params, sc.mass) for sc in scs]) as b:
        b.step(30)
        pos, vel = b.state()
        refs, k0 = [], 0
        for sc in scs:
            k1 = k0 + sc.radius.shape[0]
            refs.append(oracle.statistics(pos[k0:k1], vel[k0:k1], sc.radius, sc.mass))
            k0 = k1
        _check(b, refs, "f64")


def test_statistics_before_the_first_step(psx_mod, oracle):
    # the state is still in input order: the statistics bucket it themselves
    sc = dense_box(dtype="f64")
    with psx_mod.Simulation(sc.pos, np.zeros_like(sc.vel), sc.radius, sc.mass, sc.params) as s:
        ref = oracle.statistics(sc.pos, np.zeros_like(sc.vel), sc.radius, sc.mass)
        got = s.statistics()[0]
        assert got["kinetic_energy"] == 0.0 and ref["kinetic_energy"] == 0.0
        assert got["n_collided"] == ref["n_collided"] and got["max_overlap"] == ref["max_overlap"]
        np.testing.assert_array_equal(got["min"], ref["min"])
        np.testing.assert_array_equal(got["max"], ref["max"])


def test_statistics_between_steps_leave_the_trajectory(psx_mod):
    sc = dense_box(dtype="f64", seed=3)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as a, \
            psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as b:
        for _ in range(4):
            a.step(7)
            b.step(7)
            b.statistics()
        pa, va = a.state()
        pb, vb = b.state()
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(va, vb)
