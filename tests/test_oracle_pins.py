"""Pins the oracle (CPU restatement of the reference's CPU step) without the reference binary.

The reference has no tests, fixtures or golden vectors for this path, and its CPU engine cannot be
built here (glm, spdlog, assimp and SimView come from an empty submodule). Every oracle function is
therefore pinned by what the mathematics fixes: known grid answers, closed forms (free fall, wall
reflection, two-body restitution and spring kicks, plane and crease contact), invariants
(momentum of the restitution term, bucketing = stable counting sort), brute force (O(N^2) pair
enumeration, dense sampling of a triangle) and a textbook routine (Ericson's closest point).
"""
import math

import numpy as np
import pytest

from psx import meshes, scenes
from psx.sim import Params, mass_of


# ---------------------------------------------------------------------------------------------
# grid geometry and cell hash (ParticleBucket.cpp:17-38, 66-106)
# ---------------------------------------------------------------------------------------------
def test_grid_known_answers(oracle):
    # h = r_max * 2 * (double)1.05f
    h, n = oracle.grid(0.55, (0, 0, 0), (20, 20, 60))
    assert h == 0.55 * 2.0 * float(np.float32(1.05))
    assert abs(h - 1.1549999475) < 1e-9 and list(n) == [18, 18, 52]
    h, n = oracle.grid(0.8, (-40, -40, 0), (40, 40, 300))
    assert abs(h - 1.6799999237) < 1e-9 and list(n) == [48, 48, 179]
    tri = meshes.hopper().reshape(-1, 3)
    h, n = oracle.grid(0.8, tri.min(0), tri.max(0))
    assert list(n) == [179, 179, 358]


def test_grid_of_generated_workloads(oracle):
    c1 = scenes.box_settle()
    _, n = oracle.grid(float(c1.radius.max()), c1.params.box_min, c1.params.box_max)
    assert list(n) == [18, 18, 52]
    c2 = scenes.cylinder_settle()
    v = c2.mesh.reshape(-1, 3)
    h, n = oracle.grid(float(np.float32(c2.radius.max())), v.min(0), v.max(0))
    assert list(n) == [48, 48, 179] and c2.mesh.shape[0] == 2176


def test_axis_index_special_cases(oracle):
    h = 0.5  # dyadic: (x - min)/h is exact for dyadic x
    ai = oracle.axis_index
    assert ai(0.0, 0.0, h, 10) == 0
    assert ai(-0.0, 0.0, h, 10) == 0
    assert ai(0.5, 0.0, h, 10) == 1          # exactly on a face -> upper cell
    assert ai(np.nextafter(0.5, 0), 0.0, h, 10) == 0
    assert ai(4.75, 0.0, h, 10) == 9
    assert ai(5.0, 0.0, h, 10) == 9          # at max -> clamped to n-1
    assert ai(1e9, 0.0, h, 10) == 9
    assert ai(-0.25, 0.0, h, 10) == 0        # truncation toward zero
    assert ai(-3.0, 0.0, h, 10) == 0         # clamp below
    assert ai(float("nan"), 0.0, h, 10) == 0
    # x86-64 cvttsd2si gives INT64_MIN for |q| >= 2^63 and for inf: the clamp then yields 0
    assert ai(float("inf"), 0.0, h, 10) == 0
    assert ai(1e300, 0.0, h, 10) == 0
    assert ai(-float("inf"), 0.0, h, 10) == 0


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_bucketing_is_stable_counting_sort(oracle, dtype):
    rng = np.random.default_rng(3)
    pos = rng.uniform(-2, 14, size=(3000, 3))
    res = oracle.bucketing(0.5, (0, 0, 0), (10, 11, 12), pos, dtype=dtype)
    cell, count, offset, perm = res["cell"], res["count"], res["offset"], res["perm"]
    ncell = int(np.prod(res["n"]))
    assert count.shape == (ncell,) and offset.shape == (ncell + 1,)
    assert np.array_equal(count, np.bincount(cell, minlength=ncell))
    assert np.array_equal(offset[:-1], np.concatenate([[0], np.cumsum(count)[:-1]]))
    assert offset[-1] == len(pos)
    assert np.array_equal(perm, np.argsort(cell, kind="stable"))
    # z-fastest linear ids, each coordinate clamped into the grid
    ix, rem = np.divmod(cell, res["n"][1] * res["n"][2])
    iy, iz = np.divmod(rem, res["n"][2])
    assert (ix < res["n"][0]).all() and (iy < res["n"][1]).all() and (iz < res["n"][2]).all()
    inside = ((pos >= 0) & (pos < [10, 11, 12])).all(1)
    h = res["h"]
    p = pos.astype(np.float32).astype(np.float64) if dtype == "f32" else pos
    assert np.array_equal(ix[inside], np.floor(p[inside, 0] / h).astype(int))


# ---------------------------------------------------------------------------------------------
# integrate (PhysicsEngine.cpp:19-26): free-fall closed form, exact for dyadic inputs
# ---------------------------------------------------------------------------------------------
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_free_fall_closed_form(oracle, dtype):
    dt, g = 2.0 ** -8, (0.0, 0.0, -8.0)
    x0 = np.array([[10.0, 10.0, 50.0]])
    v0 = np.array([[1.0, -2.0, 0.5]])
    prm = Params(gravity=g, dt=dt, box_min=(-100, -100, -100), box_max=(100, 100, 100),
                 spring=0.0, restitution=0.0)
    for n in (1, 7, 64):
        p, v = oracle.run(prm.as_array(), x0, v0, [0.5], [1.0], n, dtype=dtype)
        xn = x0 + n * dt * v0 + np.asarray(g) * dt * dt * n * (n - 1) / 2
        vn = v0 + n * np.asarray(g) * dt
        assert np.array_equal(p.astype(np.float64), xn) and np.array_equal(v.astype(np.float64), vn)


# ---------------------------------------------------------------------------------------------
# box walls (BoxContainer.cpp:22-65): per face, distinct coefficients
# ---------------------------------------------------------------------------------------------
WALLS = (0.11, 0.12, 0.13, 0.14, 0.15, 0.16)  # zMin, zMax, xMin, xMax, yMin, yMax


@pytest.mark.parametrize("axis,side,coeff", [(0, 0, 0.13), (0, 1, 0.14), (1, 0, 0.15),
                                             (1, 1, 0.16), (2, 0, 0.11), (2, 1, 0.12)])
def test_wall_face_closed_form(oracle, axis, side, coeff):
    lo, hi = np.zeros(3), np.array([4.0, 5.0, 6.0])
    x0 = np.array([2.0, 2.5, 3.0])
    v0 = np.array([0.25, -0.5, 0.75])
    r = 0.5
    x0[axis] = lo[axis] + 0.25 if side == 0 else hi[axis] - 0.25   # penetrating this face only
    prm = Params(gravity=(0, 0, 0), walls=WALLS, dt=2.0 ** -20, box_min=lo, box_max=hi,
                 spring=0.0, restitution=0.0)
    p, v = oracle.run(prm.as_array(), x0[None], v0[None], [r], [1.0], 1)
    x1 = x0 + v0 * 2.0 ** -20
    exp_x, exp_v = x1.copy(), v0.copy()
    exp_x[axis] = lo[axis] + r if side == 0 else hi[axis] - r
    exp_v[axis] = -v0[axis] * coeff
    assert np.array_equal(p[0], exp_x) and np.array_equal(v[0], exp_v)


def test_wall_both_faces_max_applied_last(oracle):
    prm = Params(gravity=(0, 0, 0), walls=WALLS, dt=2.0 ** -20, box_min=(0, 0, 0),
                 box_max=(1.0, 10.0, 10.0), spring=0.0, restitution=0.0)
    x0 = np.array([[0.5, 5.0, 5.0]])
    v0 = np.array([[2.0, 0.0, 0.0]])
    p, v = oracle.run(prm.as_array(), x0, v0, [0.75], [1.0], 1)
    assert p[0, 0] == 1.0 - 0.75
    assert v[0, 0] == -(-2.0 * 0.13) * 0.14


# ---------------------------------------------------------------------------------------------
# pair contact (ParticleModel.cpp:64-144): two-body closed forms
# ---------------------------------------------------------------------------------------------
def _two_body(oracle, es, k, m=(1.0, 1.0), v1=(1.0, 0.3, 0.0), v2=(-1.0, 0.0, 0.2)):
    prm = Params(gravity=(0, 0, 0), dt=2.0 ** -40, box_min=(0, 0, 0), box_max=(10, 10, 10),
                 spring=k, restitution=es)
    x = np.array([[4.6, 5.0, 5.0], [5.3, 5.1, 5.05]])
    v = np.array([v1, v2], float)
    p, vn = oracle.run(prm.as_array(), x, v, [0.5, 0.5], list(m), 1)
    d = p[0] - p[1]
    nrm = d / np.linalg.norm(d)
    return p, v, vn, nrm, d


def test_restitution_half_equal_masses_stops_normal_motion(oracle):
    p, v, vn, nrm, _ = _two_body(oracle, es=0.5, k=0.0)
    rel0 = (v[0] - v[1]) @ nrm
    assert abs(rel0) > 0.5
    assert abs((vn[0] - vn[1]) @ nrm) < 1e-13
    # tangential relative velocity untouched
    t0 = (v[0] - v[1]) - rel0 * nrm
    t1 = (vn[0] - vn[1]) - ((vn[0] - vn[1]) @ nrm) * nrm
    assert np.allclose(t0, t1, atol=1e-13)


def test_restitution_one_reverses_normal_motion(oracle):
    p, v, vn, nrm, _ = _two_body(oracle, es=1.0, k=0.0)
    assert abs((vn[0] - vn[1]) @ nrm + (v[0] - v[1]) @ nrm) < 1e-13


def test_spring_kick_along_normal(oracle):
    k = 1234.5
    p, v, vn, nrm, d = _two_body(oracle, es=0.0, k=k)
    delta = 1.0 - np.linalg.norm(d)
    assert delta > 0
    assert np.allclose(vn[0] - v[0], k * delta * nrm, rtol=1e-13, atol=1e-12)
    assert np.allclose(vn[1] - v[1], -k * delta * nrm, rtol=1e-13, atol=1e-12)


def test_restitution_conserves_momentum_unequal_masses(oracle):
    m = (0.7, 2.9)
    p, v, vn, nrm, _ = _two_body(oracle, es=0.8, k=0.0, m=m)
    dp = m[0] * (vn[0] - v[0]) + m[1] * (vn[1] - v[1])
    assert np.abs(dp).max() < 1e-12
    assert np.abs(vn[0] - v[0]).max() > 0.1


def test_no_contact_no_kick(oracle):
    prm = Params(gravity=(0, 0, 0), dt=2.0 ** -40, box_min=(0, 0, 0), box_max=(10, 10, 10),
                 spring=1e4, restitution=0.5)
    x = np.array([[4.0, 5.0, 5.0], [5.0000001, 5.0, 5.0]])  # distance just above r1 + r2
    v = np.array([[1.0, 0, 0], [-1.0, 0, 0]])
    _, vn = oracle.run(prm.as_array(), x, v, [0.5, 0.5], [1.0, 1.0], 1)
    assert np.array_equal(vn, v)


def _kick(xi, xj, vi, vj, ri, rj, mi, mj, k, es):
    """One term of the velocity kick in the reference's expression order, Python floats."""
    d = [xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]]
    dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]
    ln = math.sqrt(dd)
    ov = ri + rj - ln
    dv = [vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]]
    s = (2.0 * mj / (mi + mj)) * (((dv[0] * d[0] + dv[1] * d[1]) + dv[2] * d[2]) / dd)
    inv = 1.0 / math.sqrt(dd)
    return [(-((s * d[c]) * es)) + (k * ov) * (d[c] * inv) for c in range(3)]


def test_pairs_equal_bruteforce_in_grid_order(oracle):
    """O(N^2) enumeration finds exactly the grid's contacts, summed in the grid's order."""
    rng = np.random.default_rng(11)
    n = 260
    pos = rng.uniform(2.0, 8.0, size=(n, 3))
    vel = rng.normal(0, 1, size=(n, 3))
    r = rng.uniform(0.3, 0.6, n)
    m = mass_of(r)
    k, es = 777.0, 0.6
    prm = Params(gravity=(0, 0, 0), dt=2.0 ** -60, box_min=(0, 0, 0), box_max=(10, 10, 10),
                 spring=k, restitution=es)
    p1, v1 = oracle.run(prm.as_array(), pos, vel, r, m, 1)
    # positions after integrate (no walls are hit)
    assert ((pos - r[:, None] > 0) & (pos + r[:, None] < 10)).all()
    x = [list(map(float, row)) for row in p1]
    h, n3 = oracle.grid(float(r.max()), (0, 0, 0), (10, 10, 10))
    cells = np.clip(np.floor(p1 / h).astype(int), 0, n3 - 1)
    expected = []
    contacts = 0
    for i in range(n):
        terms = []
        for j in range(n):
            if j == i:
                continue
            e = [x[j][c] - x[i][c] for c in range(3)]
            if math.sqrt((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]) < r[i] + r[j]:
                dc = cells[j] - cells[i]
                assert (np.abs(dc) <= 1).all()          # completeness of the 27-cell search
                order = (dc[0] + 1) * 9 + (dc[1] + 1) * 3 + (dc[2] + 1)
                terms.append((order, j))
        contacts += len(terms)
        acc = list(map(float, vel[i]))
        for _, j in sorted(terms):
            t = _kick(x[i], x[j], vel[i], vel[j], float(r[i]), float(r[j]), float(m[i]),
                      float(m[j]), k, es)
            acc = [acc[c] + t[c] for c in range(3)]
        expected.append(acc)
    assert contacts > 100
    assert np.array_equal(v1, np.asarray(expected))


# ---------------------------------------------------------------------------------------------
# mesh (PolygonContainer.cpp:52-286)
# ---------------------------------------------------------------------------------------------
def _ericson(p, a, b, c):
    """Closest point on triangle abc to p (Ericson, Real-Time Collision Detection, 5.1.5)."""
    ab, ac, ap = b - a, c - a, p - a
    d1, d2 = ab @ ap, ac @ ap
    if d1 <= 0 and d2 <= 0:
        return a
    bp = p - b
    d3, d4 = ab @ bp, ac @ bp
    if d3 >= 0 and d4 <= d3:
        return b
    vc = d1 * d4 - d3 * d2
    if vc <= 0 and d1 >= 0 and d3 <= 0:
        return a + d1 / (d1 - d3) * ab
    cp = p - c
    d5, d6 = ab @ cp, ac @ cp
    if d6 >= 0 and d5 <= d6:
        return c
    vb = d5 * d2 - d1 * d6
    if vb <= 0 and d2 >= 0 and d6 <= 0:
        return a + d2 / (d2 - d6) * ac
    va = d3 * d6 - d5 * d4
    if va <= 0 and (d4 - d3) >= 0 and (d5 - d6) >= 0:
        return b + (d4 - d3) / ((d4 - d3) + (d5 - d6)) * (c - b)
    den = 1.0 / (va + vb + vc)
    return a + ab * (vb * den) + ac * (vc * den)


def test_closest_point_matches_textbook(oracle):
    rng = np.random.default_rng(5)
    for _ in range(400):
        a, b, c = rng.normal(size=(3, 3)) * rng.uniform(0.2, 3)
        p = rng.normal(size=3) * 2
        q = oracle.closest_point(a, b, c, p)
        qe = _ericson(p, a, b, c)
        assert abs(np.linalg.norm(q - p) - np.linalg.norm(qe - p)) < 1e-9 * (1 + np.linalg.norm(p))
        assert np.linalg.norm(q - qe) < 1e-6 * (1 + np.linalg.norm(p))


def test_closest_point_dense_sampling(oracle):
    a, b, c = np.array([0.0, 0, 0]), np.array([2.0, 0.2, 0]), np.array([0.3, 1.5, 0.4])
    u, v = np.meshgrid(np.linspace(0, 1, 301), np.linspace(0, 1, 301))
    keep = u + v <= 1
    samples = a + u[keep, None] * (b - a) + v[keep, None] * (c - a)
    rng = np.random.default_rng(6)
    for p in rng.uniform(-2, 3, size=(40, 3)):
        q = oracle.closest_point(a, b, c, p)
        dmin = np.linalg.norm(samples - p, axis=1).min()
        assert np.linalg.norm(q - p) <= dmin + 1e-12
        assert np.linalg.norm(q - p) >= dmin - 0.02


def test_generated_meshes_have_outward_normals(oracle):
    cyl = meshes.open_cylinder()
    n = oracle.mesh_normals(cyl)
    ctr = cyl.reshape(-1, 3, 3).mean(1)
    side = np.abs(n[:, 2]) < 0.5
    rad = ctr[side, :2] / np.linalg.norm(ctr[side, :2], axis=1, keepdims=True)
    assert side.sum() == 2048 and ((n[side, :2] * rad).sum(1) > 0.99).all()
    assert np.allclose(n[~side], [0, 0, -1])
    box = meshes.box_mesh()
    nb = oracle.mesh_normals(box)
    cb = box.reshape(-1, 3, 3).mean(1) - np.array([20.0, 20.0, 30.0])
    assert box.shape[0] == 80 and ((nb * cb).sum(1) > 0).all()
    hop = meshes.hopper()
    nh = oracle.mesh_normals(hop)
    ch = hop.reshape(-1, 3, 3).mean(1)
    cone = (ch[:, 2] > 1e-9) & (ch[:, 2] < 199.0)
    radial = (nh[cone, :2] * ch[cone, :2]).sum(1)
    assert hop.shape[0] == 8192 and (radial > 0).all() and (nh[cone, 2] < 0).all()
    assert np.allclose(oracle.mesh_normals(hop[-256:]), [0, 0, -1])


def _floor(size=10.0):
    a, b, c, d = (np.array(x, float) for x in ([-size, -size, 0], [size, -size, 0],
                                               [size, size, 0], [-size, size, 0]))
    # winding seen from below: normal -z
    return np.array([np.concatenate([a, c, b]), np.concatenate([a, d, c])])


def test_mesh_floor_plane_closed_form(oracle):
    tri = _floor()
    assert np.allclose(oracle.mesh_normals(tri), [0, 0, -1])
    e = 0.37
    prm = Params(gravity=(0, 0, 0), walls=(e,) * 6, dt=2.0 ** -30, spring=0.0, restitution=0.0)
    x0 = np.array([[3.0, -4.0, 0.25]])        # inside the first triangle only, penetrating
    v0 = np.array([[1.0, 0.5, -2.0]])
    p, v = oracle.run(prm.as_array(), x0, v0, [0.5], [1.0], 1, mesh=tri)
    assert p[0, 2] == 0.5
    assert np.allclose(p[0, :2], x0[0, :2] + v0[0, :2] * 2.0 ** -30, rtol=0, atol=1e-15)
    assert np.allclose(v[0], e * np.array([1.0, 0.5, 2.0]), rtol=1e-15)


def test_mesh_crease_averages_reflections(oracle):
    e = 0.5
    floor = _floor()
    # wall x = 0 facing -x (particles live at x > 0)
    a, b, c, d = (np.array(x, float) for x in ([0, -10, -1], [0, 10, -1], [0, 10, 10], [0, -10, 10]))
    wall = np.array([np.concatenate([a, c, b]), np.concatenate([a, d, c])])
    nw = oracle.mesh_normals(wall)
    assert np.allclose(nw, [-1, 0, 0])
    tri = np.concatenate([floor, wall])
    prm = Params(gravity=(0, 0, 0), walls=(e,) * 6, dt=2.0 ** -40, spring=0.0, restitution=0.0)
    x0 = np.array([[0.3, 2.0, 0.4]])
    v0 = np.array([[-1.0, 0.25, -3.0]])
    p, v = oracle.run(prm.as_array(), x0, v0, [0.5], [1.0], 1, mesh=tri)
    rf = v0[0] * [1, 1, -1]
    rw = v0[0] * [-1, 1, 1]
    assert np.allclose(v[0], e * (rf + rw) / 2, rtol=1e-14)
    assert np.allclose(p[0], [0.5, 2.0, 0.5], atol=1e-9)


def test_mass_formula(oracle):
    for r in (0.3, 0.5, 1.0, 2.0, 5.0):
        m = oracle.mass(1.0, r)
        assert math.isclose(m, 4.0 / 3.0 * math.pi * r ** 3, rel_tol=4e-16)
        assert m == float(mass_of([r])[0])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fast_mode_equals_reference_loops(oracle, dtype):
    """The oracle's fast mode (triangle-grid index for the mesh loop, threaded pair loop), used for
    the long test runs, gives bit for bit the results of the reference's loops: a mesh scene with
    wall and floor contacts, a hopper cone, and a dense box."""
    from helpers import dense_box, small_cylinder
    from psx.sim import Scene
    cases = [(small_cylinder(dtype=dtype), 150)]
    rng = np.random.default_rng(4)
    hop = meshes.hopper(outlet_radius=3.0, upper_radius=12.0, half_angle_deg=35.0,
                        upper_height=20.0, n_az=48, n_cone=6, n_cyl=4, n_floor=48)
    hp = rng.uniform([-4, -4, 6], [4, 4, 14], size=(600, 3))
    hr = rng.uniform(0.3, 0.6, 600)
    cases.append((Scene(hp, rng.normal(0, 1, (600, 3)), hr,
                        Params(gravity=(0, 0, -9.8), walls=(0.4,) * 6, spring=200.0,
                               restitution=0.4, dt=1e-3, dtype=dtype), mass_of(hr), hop), 300))
    cases.append((dense_box(n_side=8, dtype=dtype), 100))
    for sc, steps in cases:
        a = oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                       mesh=sc.mesh, dtype=dtype)
        b = oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                       mesh=sc.mesh, dtype=dtype, fast=True)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert not np.array_equal(a[0], np.asarray(sc.pos, a[0].dtype))


# ---------------------------------------------------------------------------------------------
# the points where the reference's CPU and CUDA engines differ (PhysicsEngine.cpp:38-48 against
# PhysicsEngineCUDA.cpp:36-43): each closed form below holds for the CPU engine and fails for the
# CUDA engine's variant
# ---------------------------------------------------------------------------------------------
def test_spring_kick_carries_no_mass_factor(oracle):
    """ParticleModel.cpp:87-89: the spring term is k * delta * normalize(d), whatever the masses
    (the CUDA kernel divides it by m_i, ParticleModel_kernels.cu:195). With masses 0.3 and 7 both
    spheres receive kicks of the same size k * delta, in opposite directions."""
    k = 250.0
    m = (0.3, 7.0)
    p, v, vn, nrm, d = _two_body(oracle, es=0.0, k=k, m=m)
    delta = 1.0 - np.linalg.norm(d)
    assert delta > 0.05
    assert np.allclose(vn[0] - v[0], k * delta * nrm, rtol=1e-13, atol=1e-12)
    assert np.allclose(vn[1] - v[1], -k * delta * nrm, rtol=1e-13, atol=1e-12)
    # the mass-divided variant would differ by factors 1/0.3 and 1/7
    assert not np.allclose(vn[0] - v[0], k * delta * nrm / m[0], rtol=1e-3)


def _tilted_wall(n2):
    """Two large triangles through the origin with outward normal n2 (in the x-z plane)."""
    u = np.array([0.0, 1.0, 0.0])
    w = np.cross(u, n2)  # in the plane, perpendicular to u
    c = [s1 * 10.0 * u + s2 * 10.0 * w for s1, s2 in ((-1, -1), (1, -1), (1, 1), (-1, 1))]
    tri = np.array([np.concatenate([c[0], c[1], c[2]]), np.concatenate([c[0], c[2], c[3]])])
    return tri


def _crease_60(oracle):
    """Floor z = 0 (normal -z, triangles 0 and 1) and a wall through the origin whose outward
    normal n2 = (-sqrt(3)/2, 0, -1/2) makes 60 degrees with the floor's (triangles 2 and 3): a
    120-degree wedge along the y axis."""
    n2 = np.array([-math.sqrt(3.0) / 2.0, 0.0, -0.5])
    wall = _tilted_wall(n2)
    nw = oracle.mesh_normals(wall)
    if not np.allclose(nw, n2):
        wall = wall.reshape(-1, 3, 3)[:, [0, 2, 1], :].reshape(-1, 9)
    assert np.allclose(oracle.mesh_normals(wall), n2)
    return np.concatenate([_floor(), wall]), n2


def test_mesh_corrections_are_sequential(oracle):
    """PolygonContainer.cpp:206-210: each hit moves the running position y so that it ends r
    inside that triangle's plane, starting from where the previous hit left it (the CUDA kernels
    sum corrections computed from the original x, PolygonContainerCUDA_kernels.cu:261-294). For a
    sphere at x = (x0, 0, z0) penetrating the floor, then the 60-degree wall:
      floor hit:  y = (x0, 0, r)
      wall hit:   m = r - (sqrt(3)/2 x0 + r/2),  y = (x0 + sqrt(3)/2 m, 0, r + m/2)
    so the sphere ends exactly r inside the wall and more than r above the floor."""
    tri, n2 = _crease_60(oracle)
    r, e = 0.5, 0.5
    x0, z0 = 0.1, 0.3
    prm = Params(gravity=(0, 0, 0), walls=(e,) * 6, dt=2.0 ** -40, spring=0.0, restitution=0.0)
    v0 = np.array([[0.25, 0.5, -1.0]])
    p, v = oracle.run(prm.as_array(), np.array([[x0, 0.0, z0]]), v0, [r], [1.0], 1, mesh=tri)
    m = r - (math.sqrt(3.0) / 2.0 * x0 + r / 2.0)
    seq = np.array([x0 + math.sqrt(3.0) / 2.0 * m, 0.0, r + m / 2.0])
    assert np.allclose(p[0], seq, atol=1e-9)
    # the summed corrections of the CUDA kernels would land elsewhere
    s2 = math.sqrt(3.0) / 2.0 * x0 + z0 / 2.0
    summed = np.array([x0, 0.0, z0]) + (r - z0) * np.array([0, 0, 1.0]) - (r - s2) * n2
    assert np.linalg.norm(summed - seq) > 0.05
    # distance r to the wall's plane, more than r above the floor
    assert abs(-(p[0] @ n2) - r) < 1e-9 and p[0, 2] > r + 0.05
    # velocity: the average of the two reflections times e
    n1 = np.array([0.0, 0.0, -1.0])
    refl = lambda vv, nn: vv - 2.0 * (vv @ nn) * nn  # noqa: E731
    assert np.allclose(v[0], e * (refl(v0[0], n1) + refl(v0[0], n2)) / 2.0, rtol=1e-14)


def test_mesh_contact_tested_at_the_original_position(oracle):
    """PolygonContainer.cpp:184-193: whether a triangle is hit is decided from the original x,
    not from the running position. At x = (0.4, 0, 0.1) the floor hit lifts y to (0.4, 0, r),
    which is more than r from the wall's plane; the wall still counts as hit (tested at x), so its
    correction m = r - (sqrt(3)/2 0.4 + r/2) < 0 pulls y back to r from the wall, and the velocity
    is the average of both reflections."""
    tri, n2 = _crease_60(oracle)
    r, e = 0.5, 0.5
    x0, z0 = 0.4, 0.1
    prm = Params(gravity=(0, 0, 0), walls=(e,) * 6, dt=2.0 ** -40, spring=0.0, restitution=0.0)
    v0 = np.array([[-0.75, 0.25, -2.0]])
    p, v = oracle.run(prm.as_array(), np.array([[x0, 0.0, z0]]), v0, [r], [1.0], 1, mesh=tri)
    m = r - (math.sqrt(3.0) / 2.0 * x0 + r / 2.0)
    assert m < 0.0
    assert np.allclose(p[0], [x0 + math.sqrt(3.0) / 2.0 * m, 0.0, r + m / 2.0], atol=1e-9)
    n1 = np.array([0.0, 0.0, -1.0])
    refl = lambda vv, nn: vv - 2.0 * (vv @ nn) * nn  # noqa: E731
    assert np.allclose(v[0], e * (refl(v0[0], n1) + refl(v0[0], n2)) / 2.0, rtol=1e-14)
    # a test at the running position would have counted the floor only
    assert not np.allclose(v[0], e * refl(v0[0], n1), rtol=1e-3)


def test_bucketing_before_walls(oracle):
    """PhysicsEngine.cpp:40-47: the cell lists are built from the positions before the container
    acts, while each sphere's neighbour search starts from the cell of its position after it
    (ParticleModel.cpp:106-110); the CUDA engine does both after the walls. Cells are 1.05 wide
    (r = 0.5) and the box ends at x = 10.6, so cell 9 spans [9.45, 10.5) and cell 10 [10.5, 11.55).
    Sphere A starts outside the box at x = 10.7 (listed in cell 10) and the wall puts it at 10.1
    (home cell 9); sphere B rests at x = 9.3 (cell 8), 0.8 from A's new position: in contact.
    A's search (cells 8-10) finds B and A is kicked by k * delta along +x; B's search (cells 7-9)
    does not find A, which is listed in cell 10, so B is not kicked at all."""
    k = 2.0
    prm = Params(gravity=(0, 0, 0), walls=(0.5,) * 6, dt=2.0 ** -40, box_min=(0, 0, 0),
                 box_max=(10.6, 10.0, 10.0), spring=k, restitution=0.0)
    h, n3 = oracle.grid(0.5, prm.box_min, prm.box_max)
    assert n3[0] == 11
    assert oracle.axis_index(10.7, 0.0, h, 11) == 10 and oracle.axis_index(10.1, 0.0, h, 11) == 9
    assert oracle.axis_index(9.3, 0.0, h, 11) == 8
    x = np.array([[10.7, 5.0, 5.0], [9.3, 5.0, 5.0]])
    v = np.zeros((2, 3))
    p, vn = oracle.run(prm.as_array(), x, v, [0.5, 0.5], [1.0, 1.0], 1)
    assert p[0, 0] == 10.6 - 0.5 and p[1, 0] == 9.3
    delta = 1.0 - (p[0, 0] - p[1, 0])
    assert np.allclose(vn[0], [k * delta, 0.0, 0.0], rtol=1e-12)
    assert np.array_equal(vn[1], [0.0, 0.0, 0.0])
