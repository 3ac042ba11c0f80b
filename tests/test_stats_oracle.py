"""Pins of the oracle's statistics restatement (ParticleStatistics.cpp:3-43) by closed forms."""
import numpy as np

from oracle import ref


def test_two_overlapping_spheres():
    pos = np.array([[0.0, 0.0, 0.0], [0.75, 0.0, 0.0], [5.0, 5.0, 5.0]])
    vel = np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 0.0], [0.0, -3.0, 4.0]])
    r = np.array([0.5, 0.5, 0.25])
    m = np.array([2.0, 1.0, 4.0])
    s = ref.statistics(pos, vel, r, m)
    # 0.5 * 2 * 9 + 0 + 0.5 * 4 * 25
    assert s["kinetic_energy"] == 9.0 + 50.0
    np.testing.assert_array_equal(s["min"], [-0.5, -0.5, -0.5])
    np.testing.assert_array_equal(s["max"], [5.25, 5.25, 5.25])
    # only the lower index of an overlapping pair counts as collided
    assert s["n_collided"] == 1
    assert s["max_overlap"] == 1.0 - 0.75


def test_touching_is_not_collided_and_order_matters():
    # exactly touching: r_i + r_j - d == 0, not > 0
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]])
    s = ref.statistics(pos, np.zeros((2, 3)), [0.5, 0.5], [1.0, 1.0])
    assert s["n_collided"] == 0 and s["max_overlap"] == 0.0
    # a chain of three: particles 0 and 1 have a higher-index partner, particle 2 does not
    pos = np.array([[0.0, 0.0, 0.0], [0.9, 0.0, 0.0], [1.8, 0.0, 0.0]])
    s = ref.statistics(pos, np.zeros((3, 3)), [0.5, 0.5, 0.5], [1.0, 1.0, 1.0])
    assert s["n_collided"] == 2
    assert np.isclose(s["max_overlap"], 0.1)


def test_brute_force_count_matches_definition():
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, 4, size=(60, 3))
    r = rng.uniform(0.2, 0.5, 60)
    s = ref.statistics(pos, np.zeros((60, 3)), r, np.ones(60))
    d = np.sqrt(((pos[None, :, :] - pos[:, None, :]) ** 2).sum(-1))
    ov = r[:, None] + r[None, :] - d
    upper = np.triu(ov > 0, k=1)
    assert s["n_collided"] == int(upper.any(axis=1).sum())
    assert np.isclose(s["max_overlap"], ov[upper].max())
