"""Parity of the HIP backend with the oracle (CPU restatement of the reference's CPU step).

Bar: bucketing (cell ids, counts, offsets, permutation) bit-exact; trajectories bit-exact against
the oracle evaluated in the same scalar type (binary64 = the reference's arithmetic; binary32 = the
same expression order in float).
"""
import numpy as np
import pytest

from helpers import (as_dtype, assert_bitwise, dense_box, hip_run, oracle_run, small_cylinder)

pytestmark = pytest.mark.gpu


def _bucket_compare(oracle, sim, pos, dtype, rmax, mn, mx):
    ref = oracle.bucketing(rmax, mn, mx, as_dtype(pos, dtype), dtype=dtype)
    # per-cell id blocks, dense lists, and the full bucketing of the incremental form's handles
    # (the look-back scan on every grid)
    for variant in (1, 2, 3):
        sim.set_tuning(sort_variant=variant)
        got = sim.bucketing()
        for k in ("cell", "count", "offset", "perm"):
            assert_bitwise(got[k], ref[k], f"{k} sort_variant={variant}")
    g = sim.grid()
    assert g["h"] == ref["h"] and np.array_equal(g["n"], ref["n"])


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_bucketing_box(psx_mod, oracle, dtype):
    sc = dense_box(dtype=dtype)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        rmax = float(as_dtype(sc.radius, dtype).max())
        _bucket_compare(oracle, s, sc.pos, dtype, rmax, sc.params.box_min, sc.params.box_max)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_bucketing_edge_positions(psx_mod, oracle, dtype):
    """Cell faces, outside the domain on every side, -0.0, huge and non-finite coordinates."""
    rng = np.random.default_rng(7)
    h_r = 0.5
    from psx.sim import Params
    prm = Params(box_min=(0.0, -2.0, 1.0), box_max=(8.0, 6.0, 9.0), dtype=dtype, dt=1e-4)
    h, n3 = oracle.grid(h_r, prm.box_min, prm.box_max)
    faces = []
    for k in range(int(n3[0]) + 2):
        faces.append([k * h, 1.0, 2.0])
        faces.append([np.nextafter(k * h, -np.inf), 1.0, 2.0])
    special = [[-0.0, -2.0, 1.0], [-1e-300, 0.0, 5.0], [-5.0, -7.0, -3.0], [50.0, 60.0, 70.0],
               [1e30, -1e30, 4.0], [np.inf, 0.0, 2.0], [-np.inf, 1.0, 2.0], [np.nan, 1.0, 2.0],
               [3.0, np.nan, np.nan], [8.0, 6.0, 9.0], [0.0, -2.0, 1.0]]
    rand = rng.uniform(-3, 12, size=(400, 3))
    pos = np.concatenate([np.asarray(faces), np.asarray(special), rand])
    n = pos.shape[0]
    r = np.full(n, h_r)
    with psx_mod.Simulation(np.nan_to_num(pos), np.zeros((n, 3)), r, None, prm) as s:
        s.set_state(pos, np.zeros((n, 3)))
        _bucket_compare(oracle, s, pos, dtype, h_r, prm.box_min, prm.box_max)


# Launch shapes: workgroup sizes and lanes per particle in the pair kernel (pair_variant). Every
# shape must reproduce the oracle bit for bit.
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_bucketing_cell_overflow(psx_mod, oracle, dtype):
    """Cells holding more particles than the per-cell id block (8 slots): one cell with 300, many
    with 9-12, the rest sparse. The overflow list must give the same stable order; then a few steps
    through the same path."""
    from psx.sim import Params, mass_of
    rng = np.random.default_rng(11)
    h_r = 0.5
    prm = Params(box_min=(0.0, 0.0, 0.0), box_max=(12.0, 12.0, 12.0), dtype=dtype, dt=1e-4,
                 spring=5.0, restitution=0.3, gravity=(0.0, 0.0, -1.0))
    h = h_r * 2.0 * float(np.float32(1.05))
    parts = [rng.uniform(2 * h + 0.01, 3 * h - 0.01, size=(300, 3))]  # one cell
    for k in range(40):
        corner = rng.integers(4, 10, size=3) * h
        parts.append(corner + rng.uniform(0.01, h - 0.01, size=(int(rng.integers(9, 13)), 3)))
    parts.append(rng.uniform(0.5, 11.5, size=(500, 3)))
    pos = np.concatenate(parts)
    rng.shuffle(pos)  # ids in no particular spatial order
    n = pos.shape[0]
    r = rng.uniform(0.3, h_r, n)
    r[0] = h_r
    vel = rng.normal(0.0, 0.1, size=(n, 3))
    with psx_mod.Simulation(pos, vel, r, None, prm) as s:
        rmax = float(as_dtype(r, dtype).max())
        _bucket_compare(oracle, s, pos, dtype, rmax, prm.box_min, prm.box_max)
        got = s.bucketing()
        assert got["count"].max() >= 300
    from psx.sim import Scene
    sc = Scene(pos, vel, r, prm, mass_of(r))
    ref_p, ref_v = oracle_run(oracle, sc, 3)
    for variant in (1, 2):  # id blocks with their overflow list, dense lists
        p, v = hip_run(sc, 3, sort_variant=variant)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos sort_variant={variant}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel sort_variant={variant}")
    # the one-lane kernel's near-candidate queue overflows in the 300-sphere cell: the lanes test
    # the candidates after the queue's last one in line (four- and two-wide chunks)
    for chunk in (4, 2):
        p, v = hip_run(sc, 3, pair_variant=1, pair_chunk=chunk)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos pair_chunk={chunk}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel pair_chunk={chunk}")
    # the LDS-staged pair kernel: the 300-sphere cell overfills a workgroup's LDS share, so some
    # workgroups read from memory and the others from LDS
    p, v = hip_run(sc, 3, pair_variant=32, block_pairs=64)
    assert_bitwise(as_dtype(p, dtype), ref_p, "pos staged")
    assert_bitwise(as_dtype(v, dtype), ref_v, "vel staged")


VARIANTS = [dict(block_pairs=256, block_integrate=256),
            dict(block_pairs=128, block_integrate=64, pair_variant=2, sort_variant=2),
            dict(block_pairs=64, block_integrate=1024, block_sort=64, pair_variant=4, sort_variant=1),
            dict(block_pairs=1024, block_integrate=128, block_sort=1024, pair_variant=8,
                 sort_variant=2),
            dict(block_pairs=512, pair_variant=16, sort_variant=1),
            dict(block_pairs=256, block_sort=128, pair_variant=32, sort_variant=1),
            dict(block_pairs=128, pair_variant=1, pair_chunk=2, sort_variant=1),
            dict(block_pairs=1024, pair_chunk=2, sort_variant=2),
            dict(block_pairs=64, block_integrate=512, pair_variant=32, sort_variant=2),
            dict(block_pairs=256, pair_variant=64, sort_variant=1),
            dict(block_pairs=512, block_sort=512, pair_variant=64, sort_variant=2)]


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("steps", [1, 60])
@pytest.mark.parametrize("tune", range(len(VARIANTS)))
def test_box_steps_bitexact(psx_mod, oracle, dtype, steps, tune):
    sc = dense_box(dtype=dtype)
    ref_p, ref_v = oracle_run(oracle, sc, steps)
    tune_kw = dict(VARIANTS[tune])
    if dtype == "f64":  # the binary64 rank/gather kernel takes workgroups of at most 512
        tune_kw["block_sort"] = min(tune_kw.get("block_sort", 256), 512)
    p, v = hip_run(sc, steps, **tune_kw)
    assert_bitwise(as_dtype(p, dtype), ref_p, "pos")
    assert_bitwise(as_dtype(v, dtype), ref_v, "vel")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_box_long_bitexact(psx_mod, oracle, dtype):
    """600 steps of a settling box at moderate stiffness."""
    sc = dense_box(n_side=12, spacing=1.05, dtype=dtype, spring=300.0, vscale=0.5,
                   box=(14.0, 14.0, 16.0))
    ref_p, ref_v = oracle_run(oracle, sc, 600)
    p, v = hip_run(sc, 600)
    assert_bitwise(as_dtype(p, dtype), ref_p, "pos")
    assert_bitwise(as_dtype(v, dtype), ref_v, "vel")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_graph_replay_equals_eager(psx_mod, dtype):
    sc = dense_box(dtype=dtype)
    p0, v0 = hip_run(sc, 37, graph=False)
    p1, v1 = hip_run(sc, 37, graph=True)
    assert_bitwise(p1, p0, "pos")
    assert_bitwise(v1, v0, "vel")


RUNS = (16, 16, 64, 83, 64, 5, 16)  # 83 = 64 + 16 + 3 single steps


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_graph_replays_interleaved_with_eager_work(psx_mod, oracle, dtype):
    """Graphs replayed repeatedly (16-step and 64-step graphs, and runs that mix both with single
    steps), with the program's own eager work between the replays (state copies, a statistics pass
    that buckets the state on the same arrays, a synchronisation): every replay's state equals that
    of an eager handle stepped alongside, bit for bit, on a box scene and a mesh scene, and the whole
    run equals the oracle's (the statistics passes leave the trajectory alone)."""
    for sc in (dense_box(dtype=dtype), small_cylinder(n=400, dtype=dtype)):
        with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as e, \
                psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as g:
            e.set_tuning(use_graph=0)
            g.set_tuning(use_graph=1)
            for r, k in enumerate(RUNS):
                e.step(k)
                g.step(k)
                e.synchronize()
                g.synchronize()
                pe, ve = e.state()
                pg, vg = g.state()
                assert_bitwise(pg, pe, f"pos replay {r}")
                assert_bitwise(vg, ve, f"vel replay {r}")
                se, sg = e.statistics()[0], g.statistics()[0]
                assert se["n_collided"] == sg["n_collided"]
                assert se["kinetic_energy"] == sg["kinetic_energy"]
        ref_p, ref_v = oracle_run(oracle, sc, sum(RUNS))
        assert_bitwise(as_dtype(pg, dtype), ref_p, "pos after the runs")
        assert_bitwise(as_dtype(vg, dtype), ref_v, "vel after the runs")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("steps", [1, 150])
def test_mesh_steps_bitexact(psx_mod, oracle, dtype, steps):
    """Triangle-grid mesh contact against the reference's brute-force O(N*T) loop."""
    sc = small_cylinder(dtype=dtype)
    ref_p, ref_v = oracle_run(oracle, sc, steps)
    p, v = hip_run(sc, steps)
    assert_bitwise(as_dtype(p, dtype), ref_p, "pos")
    assert_bitwise(as_dtype(v, dtype), ref_v, "vel")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("lanes", [2, 4, 8, 16, 32, 64, "chunk2"])
def test_mesh_pair_lanes_bitexact(psx_mod, oracle, dtype, lanes):
    """Multi-lane pair kernel (and the LDS-staged ones, pair_variant 32 per workgroup and 64 per
    wave, and the one-lane one with two candidates per round) on the mesh scene, eager and
    graph-replayed; container pushes move some homes out of their staged ranges."""
    sc = small_cylinder(dtype=dtype)
    ref_p, ref_v = oracle_run(oracle, sc, 40)
    shape = dict(pair_variant=1, pair_chunk=2) if lanes == "chunk2" else dict(pair_variant=lanes)
    for graph in (False, True):
        p, v = hip_run(sc, 40, graph=graph, block_pairs=256, **shape)
        assert_bitwise(as_dtype(p, dtype), ref_p, f"pos lanes={lanes} graph={graph}")
        assert_bitwise(as_dtype(v, dtype), ref_v, f"vel lanes={lanes} graph={graph}")


def test_tuning_rejects_bad_lanes(psx_mod):
    sc = dense_box(dtype="f64")
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        for bad in (3, 48, 128, -1):
            with pytest.raises(psx_mod.PsxError):
                s.set_tuning(pair_variant=bad)
        with pytest.raises(psx_mod.PsxError, match="512"):
            s.set_tuning(pair_variant=64, block_pairs=1024)  # binary64: two windows per wave
        for bad in (1, 3, 8):
            with pytest.raises(psx_mod.PsxError, match="pair_chunk"):
                s.set_tuning(pair_chunk=bad)
        with pytest.raises(psx_mod.PsxError, match="256"):
            s.set_tuning(pair_variant=32, block_pairs=512)  # binary64: LDS share too large
        s.set_tuning(pair_variant=4)
        assert s.tuning["pair_variant"] == 4
        with pytest.raises(psx_mod.PsxError, match="512"):
            s.set_tuning(block_sort=1024)  # binary64 rank/gather: at most 512
        s.set_tuning(block_sort=512)
        for bad in (-1, 4):
            with pytest.raises(psx_mod.PsxError, match="sort_variant"):
                s.set_tuning(sort_variant=bad)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_mesh_fan_center_many_hits(psx_mod, oracle, dtype):
    """Spheres resting on the centre of a 128-sliver fan floor: cell lists of ~128 triangles (the
    wave-cooperative path) and particles touching dozens of triangles at once."""
    from psx import meshes
    from psx.sim import Params, Scene, mass_of
    rng = np.random.default_rng(9)
    mesh = meshes.open_cylinder(30.0, 30.0, 128, 2, 128)
    n = 400
    p = np.concatenate([rng.uniform([-2, -2, 0.1], [2, 2, 0.6], size=(n // 2, 3)),
                        rng.uniform([-12, -12, 0.2], [12, 12, 3.0], size=(n // 2, 3))])
    r = rng.uniform(0.3, 0.6, n)
    v = rng.normal(0.0, 1.0, size=(n, 3))
    v[:, 2] -= 3.0
    prm = Params(gravity=(0, 0, -9.8), walls=(0.4,) * 6, spring=80.0, restitution=0.5, dt=1e-3,
                 dtype=dtype)
    sc = Scene(p, v, r, prm, mass_of(r), mesh)
    for steps in (1, 40):
        ref_p, ref_v = oracle_run(oracle, sc, steps)
        hp, hv = hip_run(sc, steps)
        assert_bitwise(as_dtype(hp, dtype), ref_p, f"pos after {steps}")
        assert_bitwise(as_dtype(hv, dtype), ref_v, f"vel after {steps}")


def test_state_roundtrip(psx_mod):
    sc = dense_box(dtype="f64")
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        s.step(3)
        s.bucketing()
        p, v = s.state()
        s.set_state(p, v)
        p2, v2 = s.state()
        assert_bitwise(p2, p, "pos")
        assert_bitwise(v2, v, "vel")


def test_nonfinite_flag(psx_mod):
    sc = dense_box(dtype="f64")
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        s.step(1)
        s.synchronize()
        p, v = s.state()
        v[3, 0] = np.inf
        s.set_state(p, v)
        s.step(1)
        with pytest.raises(psx_mod.PsxError) as e:
            s.synchronize()
        assert e.value.code == 5


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_autotuned_run_bitexact(psx_mod, oracle, dtype, tmp_path):
    """The autotuner's sweep leaves the state as it was and its chosen shape reproduces the oracle."""
    from psx.tune import autotune
    sc = small_cylinder(dtype=dtype)
    ref_p, ref_v = oracle_run(oracle, sc, 30)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        out = autotune(s, steps=2, rounds_per_shape=1, cache=tmp_path / "t.json")
        assert s.tuning["pair_variant"] == out["tuning"]["pair_variant"]
        s.step(30)
        s.synchronize()
        p, v = s.state()
    assert_bitwise(as_dtype(p, dtype), ref_p, "pos")
    assert_bitwise(as_dtype(v, dtype), ref_v, "vel")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("n", [1, 2, 3, 5])
def test_few_spheres_bitexact(psx_mod, oracle, dtype, n):
    """A handful of touching spheres in a box: the chunked candidate reads run into the arrays'
    spare entries and the workgroups are almost empty. Both chunk widths, eager and graph-replayed,
    bit-exact against the oracle; and the same scene as a one-of-three scene batch."""
    from psx.sim import Params, Scene, mass_of
    rng = np.random.default_rng(100 + n)
    pos = np.array([[2.0 + 0.8 * k, 2.0 + 0.1 * k, 1.0 + 0.05 * k] for k in range(n)])
    vel = rng.normal(0.0, 1.0, size=(n, 3))
    r = rng.uniform(0.45, 0.55, n)
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=1e4, restitution=0.5, dt=1e-4,
                 box_min=(0.0, 0.0, 0.0), box_max=(6.0, 5.0, 7.0), dtype=dtype)
    sc = Scene(pos, vel, r, prm, mass_of(r))
    ref_p, ref_v = oracle_run(oracle, sc, 30)
    for shape in (dict(pair_chunk=4), dict(pair_chunk=2), dict(pair_variant=4)):
        for graph in (False, True):
            p, v = hip_run(sc, 30, graph=graph, **shape)
            assert_bitwise(as_dtype(p, dtype), ref_p, f"pos n={n} {shape} graph={graph}")
            assert_bitwise(as_dtype(v, dtype), ref_v, f"vel n={n} {shape} graph={graph}")
    big = dense_box(n_side=4, dtype=dtype)
    with psx_mod.Batch([big, sc, big]) as b:
        b.step(30)
        b.synchronize()
        bp, bv = b.state()
    off = big.radius.shape[0]
    assert_bitwise(as_dtype(bp[off:off + n], dtype), ref_p, f"batched pos n={n}")
    assert_bitwise(as_dtype(bv[off:off + n], dtype), ref_v, f"batched vel n={n}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_large_grid_steps_bitexact(psx_mod, oracle, dtype):
    """1,000 spheres in a box of 104^3 = 1.12M cells (more than the direct scan's 2^20): the
    look-back scan over almost empty tiles, steps eager and graph-replayed, bit-exact against the
    oracle, with both cell-list forms."""
    sc = dense_box(dtype=dtype, box=(120.0, 120.0, 120.0))
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        assert s.num_cells > (1 << 20)
    ref_p, ref_v = oracle_run(oracle, sc, 20)
    for variant in (1, 2):
        for graph in (False, True):
            p, v = hip_run(sc, 20, graph=graph, sort_variant=variant)
            assert_bitwise(as_dtype(p, dtype), ref_p, f"pos sort_variant={variant} graph={graph}")
            assert_bitwise(as_dtype(v, dtype), ref_v, f"vel sort_variant={variant} graph={graph}")


@pytest.mark.parametrize("dtype", ["f64", "f32"])
@pytest.mark.parametrize("form", [1, 3])
def test_prepared_tail_graphs_equal_eager(psx_mod, oracle, dtype, form):
    """Runs of 1, 7, 31 (16 + 15), 69 (64 + 5), 7 and 1 steps, each prepared first, so that its last
    1-15 steps replay a graph of that length (an odd length ends on the other position buffer, and the
    next run starts from it): every run's state equals an eager handle's stepped alongside, bit for
    bit, and the whole run equals the oracle's."""
    sc = dense_box(dtype=dtype)
    runs = (1, 7, 31, 69, 7, 1)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as g, \
            psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as e:
        g.set_tuning(use_graph=1, sort_variant=form)
        e.set_tuning(use_graph=0, sort_variant=form)
        for r, k in enumerate(runs):
            g.prepare(k)
            g.step(k)
            e.step(k)
            pg, vg = g.state()
            pe, ve = e.state()
            assert_bitwise(pg, pe, f"pos run {r} ({k} steps)")
            assert_bitwise(vg, ve, f"vel run {r} ({k} steps)")
    ref_p, ref_v = oracle_run(oracle, sc, sum(runs))
    assert_bitwise(as_dtype(pg, dtype), ref_p, "pos vs oracle")
    assert_bitwise(as_dtype(vg, dtype), ref_v, "vel vs oracle")
