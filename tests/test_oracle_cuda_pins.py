"""Pins the oracle's restatement of the reference CUDA backend's step (semantics "cuda", the
product's PSX_FLAG_SEMANTICS_CUDA mode) at every point where it departs from the CPU backend:

- cell width r_max * 2.0 * 1.05 with an exact 1.05 (ParticleModelCUDA.cpp:84), which changes the
  cell count of a box whose length falls between the two widths' multiples;
- the pair kick's spring term divided by m_i (ParticleModel_kernels.cu:195): for two spheres of
  masses 0.3 and 7 the kicks are k delta / m_i, and the pair's momentum is conserved (the CPU
  backend's kicks have equal size k delta);
- container before bucketing (PhysicsEngineCUDA.cpp:38-42): a sphere that a wall pushes across a
  cell face is listed in its new cell, so its neighbour on that side finds it;
- mesh corrections computed from the incoming position and summed
  (PolygonContainerCUDA_kernels.cu:254-294), on a 60-degree crease where the sequential
  corrections of the CPU backend land elsewhere;
- the edge point picked by dot(center, q) (PolygonContainerCUDA_kernels.cu:125-154), which differs
  from the closest edge point when the edges tie on that measure, and the hit test |q - x|^2 < r^2
  that follows it.

Every expectation is a closed form of the geometry or the conservation law, or a bit pattern of a
one-line IEEE expression, not the oracle's code retyped.
"""
import math

import numpy as np

from psx.sim import Params

from test_oracle_pins import _crease_60, _floor


def _run(oracle, prm, x, v, r, m, steps=1, mesh=None):
    return oracle.run(prm.as_array(), x, v, r, m, steps, mesh=mesh, semantics="cuda")


# ---------------------------------------------------------------------------------------------
# grid (ParticleModelCUDA.cpp:62-99)
# ---------------------------------------------------------------------------------------------
def test_cuda_grid_width_is_exact_105(oracle):
    h, n = oracle.grid(0.55, (0, 0, 0), (20, 20, 60), semantics="cuda")
    assert h == 0.55 * 2.0 * 1.05 and h != 0.55 * 2.0 * float(np.float32(1.05))
    assert list(n) == [int(20 / h) + 1, int(20 / h) + 1, int(60 / h) + 1]
    # a box 10.4999999 long: 9.9999999 cells of the exact width, 10.0000004 of the CPU backend's
    hc, nc = oracle.grid(0.5, (0, 0, 0), (10.4999999, 5, 5), semantics="cuda")
    hp, np_ = oracle.grid(0.5, (0, 0, 0), (10.4999999, 5, 5))
    assert hc == 1.05 and hp == float(np.float32(1.05))
    assert nc[0] == 10 and np_[0] == 11


# ---------------------------------------------------------------------------------------------
# pair kick (ParticleModel_kernels.cu:163-199)
# ---------------------------------------------------------------------------------------------
def test_cuda_spring_kick_divided_by_mass_conserves_momentum(oracle):
    """Two spheres at rest in contact (delta = r_i + r_j - |d|), restitution 0: each is kicked by
    k delta / m_i along its own outward normal, so m_0 dv_0 + m_1 dv_1 = 0."""
    k = 250.0
    m = (0.3, 7.0)
    prm = Params(gravity=(0, 0, 0), dt=2.0 ** -40, box_min=(0, 0, 0), box_max=(10, 10, 10),
                 spring=k, restitution=0.0)
    x = np.array([[4.6, 5.0, 5.0], [5.3, 5.1, 5.05]])
    v = np.zeros((2, 3))
    p, vn = _run(oracle, prm, x, v, [0.5, 0.5], list(m))
    d = p[0] - p[1]
    nrm = d / np.linalg.norm(d)
    delta = 1.0 - np.linalg.norm(d)
    assert delta > 0.05
    assert np.allclose(vn[0], k * delta / m[0] * nrm, rtol=1e-13, atol=1e-12)
    assert np.allclose(vn[1], -k * delta / m[1] * nrm, rtol=1e-13, atol=1e-12)
    assert np.allclose(m[0] * vn[0] + m[1] * vn[1], 0.0, atol=1e-10)
    # the CPU backend's kicks have the same size k delta
    _, vcpu = oracle.run(prm.as_array(), x, v, [0.5, 0.5], list(m), 1)
    assert np.allclose(vcpu[0], k * delta * nrm, rtol=1e-13, atol=1e-12)


def test_cuda_restitution_term_two_body(oracle):
    """Head-on approach along x (spring 0): dv_i = -2 m_j e (v_rel . n) n / (m_i + m_j); with
    e = 1 and equal masses the spheres swap their normal velocities."""
    prm = Params(gravity=(0, 0, 0), dt=2.0 ** -40, box_min=(0, 0, 0), box_max=(10, 10, 10),
                 spring=0.0, restitution=1.0)
    x = np.array([[4.6, 5.0, 5.0], [5.5, 5.0, 5.0]])
    v = np.array([[1.0, 0.25, 0.0], [-2.0, 0.0, 0.5]])
    _, vn = _run(oracle, prm, x, v, [0.5, 0.5], [1.0, 1.0])
    assert np.allclose(vn[0], [-2.0, 0.25, 0.0], rtol=0, atol=1e-9)
    assert np.allclose(vn[1], [1.0, 0.0, 0.5], rtol=0, atol=1e-9)
    # unequal masses, e = 0.5: dv_i = -2 m_j e u / (m_i + m_j) along x (u = v_i,x - v_j,x = 3),
    # and the momentum is conserved
    prm.restitution = 0.5
    m = (2.0, 5.0)
    _, vn = _run(oracle, prm, x, v, [0.5, 0.5], list(m))
    u = 3.0
    assert np.allclose(vn[0, 0], 1.0 - 2.0 * m[1] * 0.5 * u / (m[0] + m[1]), atol=1e-9)
    assert np.allclose(vn[1, 0], -2.0 + 2.0 * m[0] * 0.5 * u / (m[0] + m[1]), atol=1e-9)
    assert np.allclose(m[0] * vn[0] + m[1] * vn[1], m[0] * v[0] + m[1] * v[1], atol=1e-9)


# ---------------------------------------------------------------------------------------------
# stage order (PhysicsEngineCUDA.cpp:38-42)
# ---------------------------------------------------------------------------------------------
def test_cuda_container_before_bucketing(oracle):
    """The CPU pin's set-up (cells 1.05 wide at this r_max, box end 10.6): sphere A starts at
    x = 10.7 and the wall puts it at 10.1; B rests at 9.3, 0.8 from A. Bucketed after the wall, A
    is listed in cell 9, so B's search (cells 7-9) finds it and both are kicked, by k delta / m
    each (the CPU backend lists A in cell 10 and leaves B alone)."""
    k = 2.0
    prm = Params(gravity=(0, 0, 0), walls=(0.5,) * 6, dt=2.0 ** -40, box_min=(0, 0, 0),
                 box_max=(10.6, 10.0, 10.0), spring=k, restitution=0.0)
    h, n3 = oracle.grid(0.5, prm.box_min, prm.box_max, semantics="cuda")
    assert h == 1.05 and n3[0] == 11
    x = np.array([[10.7, 5.0, 5.0], [9.3, 5.0, 5.0]])
    v = np.zeros((2, 3))
    m = (1.0, 3.0)
    p, vn = _run(oracle, prm, x, v, [0.5, 0.5], list(m))
    assert p[0, 0] == 10.6 - 0.5 and p[1, 0] == 9.3
    delta = 1.0 - (p[0, 0] - p[1, 0])
    assert np.allclose(vn[0], [k * delta / m[0], 0.0, 0.0], rtol=1e-12)
    assert np.allclose(vn[1], [-k * delta / m[1], 0.0, 0.0], rtol=1e-12)
    _, vcpu = oracle.run(prm.as_array(), x, v, [0.5, 0.5], list(m), 1)
    assert np.array_equal(vcpu[1], [0.0, 0.0, 0.0])


# ---------------------------------------------------------------------------------------------
# mesh container (PolygonContainerCUDA_kernels.cu:69-294)
# ---------------------------------------------------------------------------------------------
def test_cuda_mesh_corrections_summed_from_incoming_position(oracle):
    """The crease of the CPU pin (floor z = 0 with normal -z, wall with outward normal
    n2 = (-sqrt(3)/2, 0, -1/2)): a sphere at x = (0.1, 0, 0.3) penetrates both. Each correction is
    computed from x: the floor's lifts by r - z0, the wall's moves by -(r - s) n2 with
    s = sqrt(3)/2 x0 + z0/2 its distance to the wall plane; they are summed. The velocity is the
    mean of the two reflections times e."""
    tri, n2 = _crease_60(oracle)
    r, e = 0.5, 0.5
    x0, z0 = 0.1, 0.3
    prm = Params(gravity=(0, 0, 0), walls=(e,) * 6, dt=2.0 ** -40, spring=0.0, restitution=0.0)
    v0 = np.array([[0.25, 0.5, -1.0]])
    p, v = _run(oracle, prm, np.array([[x0, 0.0, z0]]), v0, [r], [1.0], mesh=tri)
    s2 = math.sqrt(3.0) / 2.0 * x0 + z0 / 2.0
    summed = np.array([x0, 0.0, z0]) + (r - z0) * np.array([0, 0, 1.0]) - (r - s2) * n2
    assert np.allclose(p[0], summed, atol=1e-9)
    n1 = np.array([0.0, 0.0, -1.0])
    refl = lambda vv, nn: vv - 2.0 * (vv @ nn) * nn  # noqa: E731
    assert np.allclose(v[0], e * (refl(v0[0], n1) + refl(v0[0], n2)) / 2.0, rtol=1e-14)
    pc, _ = oracle.run(prm.as_array(), np.array([[x0, 0.0, z0]]), v0, [r], [1.0], 1, mesh=tri)
    assert np.linalg.norm(pc[0] - p[0]) > 0.05  # the CPU backend's sequential corrections


def test_cuda_mesh_plane_and_unhit_particles(oracle):
    """A sphere penetrating the floor by 0.25 ends r above it with its normal velocity reflected
    and scaled by e; a sphere that touches nothing keeps its integrated position and velocity."""
    tri = _floor()
    e = 0.37
    prm = Params(gravity=(0, 0, 0), walls=(e,) * 6, dt=2.0 ** -30, spring=0.0, restitution=0.0)
    x0 = np.array([[3.0, -4.0, 0.25], [-3.0, 4.0, 2.0]])
    v0 = np.array([[1.0, 0.5, -2.0], [0.5, -0.5, 0.25]])
    p, v = _run(oracle, prm, x0, v0, [0.5, 0.5], [1.0, 1.0], mesh=tri)
    assert p[0, 2] == 0.5
    assert np.allclose(v[0], e * np.array([1.0, 0.5, 2.0]), rtol=1e-15)
    assert np.array_equal(p[1], x0[1] + v0[1] * 2.0 ** -30) and np.array_equal(v[1], v0[1])


def _segment_point(p, a, b):
    ab, ap = b - a, p - a
    t = min(max((ap @ ab) / (ab @ ab), 0.0), 1.0)
    return a + ab * t


def test_cuda_edge_point_picked_by_dot(oracle):
    """Triangle (0,0,0), (1,0,0), (0,1,0) and p = (2, 2, 0.1), whose projection lies outside: the
    segment points are (1,0,0) on AB, (0.5,0.5,0) on BC and (0,1,0) on CA. Their dot products with
    p are all 2, so the first (AB's) is kept, while the closest point is BC's."""
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    p = np.array([2.0, 2.0, 0.1])
    qs = [_segment_point(p, a, b), _segment_point(p, b, c), _segment_point(p, c, a)]
    assert np.allclose(qs, [[1, 0, 0], [0.5, 0.5, 0], [0, 1, 0]])
    assert [q @ p for q in qs] == [2.0, 2.0, 2.0]
    assert np.allclose(oracle.closest_point(a, b, c, p, semantics="cuda"), [1.0, 0.0, 0.0])
    assert np.allclose(oracle.closest_point(a, b, c, p), [0.5, 0.5, 0.0])
    # p = (0.2, 3, 0.1): dot picks AB's (0.2, 0, 0) (0.04) over CA's (0, 1, 0) (3)
    p2 = np.array([0.2, 3.0, 0.1])
    assert np.allclose(oracle.closest_point(a, b, c, p2, semantics="cuda"), [0.2, 0.0, 0.0])
    assert np.allclose(oracle.closest_point(a, b, c, p2), [0.0, 1.0, 0.0])


def test_cuda_edge_pick_decides_the_hit(oracle):
    """The triangle above (normal +z by its winding) plus a far one that only widens the bounds:
    a sphere of radius 2.2 at (2, 2, 0.1) is 2.12 from the closest point (0.5, 0.5, 0) but 2.24
    from (1, 0, 0), so the CPU backend counts a hit and the CUDA backend's test
    |q - x|^2 = 5.01 < 4.84 fails: its velocity stays as integrated."""
    tri = np.array([[0.0, 0, 0, 1, 0, 0, 0, 1, 0], [10.0, 10, -5, 11, 10, -5, 10, 11, 5]])
    prm = Params(gravity=(0, 0, 0), walls=(0.5,) * 6, dt=2.0 ** -40, spring=0.0, restitution=0.0)
    x = np.array([[2.0, 2.0, 0.1]])
    v = np.array([[0.0, 0.0, -1.0]])
    _, vc = _run(oracle, prm, x, v, [2.2], [1.0], mesh=tri)
    _, vp = oracle.run(prm.as_array(), x, v, [2.2], [1.0], 1, mesh=tri)
    assert np.array_equal(vc[0], v[0])
    assert not np.array_equal(vp[0], v[0])
