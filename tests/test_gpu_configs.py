"""The standard workloads' step paths at full size, bit-exact against the oracle:

- the 4,194,304-sphere hopper (8,192-triangle mesh) with its tuned launch shape (two-wide candidate
  chunks), from its lattice and from a contact-rich state pushed into the cone wall, eager and
  replayed from prepared graphs;
- the whole 256-scene batch, every scene against its plain run and a sample against the oracle;
- a crowded bimodal box of 1,048,576 spheres whose cells hold up to 27 members (overflow chains),
  bit-exact and with the rank kernel's time bounded against the lattice state;
- graphs prepared before a run (psx_prepare) give the run's results bit for bit.

The oracle runs in its fast mode (the results of its reference loops, pinned in
test_oracle_pins.py)."""
import math

import numpy as np
import pytest

from helpers import as_dtype, assert_bitwise
from psx import scenes
from psx.sim import Params, Scene, mass_of

pytestmark = pytest.mark.gpu

TUNE_CACHE = scenes.__file__.replace("scenes.py", "tuning_mi355x.json")


def _tuned(sim):
    from psx.tune import autotune
    return autotune(sim, cache=TUNE_CACHE)["tuning"]


def _oracle(oracle, sc, steps):
    return oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                      mesh=sc.mesh, dtype=sc.params.dtype, fast=True)


def _hopper_radius_at(z):
    t = math.tan(math.radians(35.0))
    return np.where(z >= 140.0 / t, 150.0, 10.0 + z * t)


def _hopper_against_the_wall():
    """The hopper scene with its heights compressed by 0.6 (vertical spacing 1.08: large spheres
    touch their vertical neighbours) and every sphere that the compression carried outside the
    cone brought back to 0.3 outside the wall at its height: a ring of spheres pressed into the
    8,192-triangle mesh, several per cell."""
    sc = scenes.hopper_settle()
    p = sc.pos.copy()
    p[:, 2] = 1.8 + (p[:, 2] - 1.8) * 0.6
    rho = np.hypot(p[:, 0], p[:, 1])
    lim = _hopper_radius_at(p[:, 2]) + 0.3
    out = rho > lim
    p[out, :2] *= (lim[out] / rho[out])[:, None]
    sc.pos = p
    return sc, int(out.sum())


def test_hopper_lattice_steps_bitexact(psx_mod, oracle):
    sc, steps = scenes.hopper_settle(), 3
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        t = _tuned(s)
        assert t["pair_chunk"] == 2 and t["pair_variant"] == 1  # the shape the bench runs
        s.set_tuning(use_graph=0)
        s.step(steps)
        p, v = s.state()
    rp, rv = _oracle(oracle, sc, steps)
    assert_bitwise(as_dtype(p, "f32"), rp, "hopper lattice pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "hopper lattice vel")


def test_hopper_against_the_wall_bitexact(psx_mod, oracle):
    """One eager step, then 16 steps replayed from a prepared graph, tuned shape; after the first
    step the container has put every sphere that started outside the cone back inside."""
    sc, n_out = _hopper_against_the_wall()
    assert n_out > 10000
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        _tuned(s)
        s.set_tuning(use_graph=1)
        s.step(1)
        p1, v1 = s.state()
        s.prepare(16)
        s.step(16)
        p, v = s.state()
    rp, rv = _oracle(oracle, sc, 1)
    assert_bitwise(as_dtype(p1, "f32"), rp, "hopper wall pos, 1 step")
    assert_bitwise(as_dtype(v1, "f32"), rv, "hopper wall vel, 1 step")
    rho = np.hypot(p1[:, 0], p1[:, 1])
    assert (rho <= _hopper_radius_at(p1[:, 2]) + 1e-3).mean() > 0.9999
    rp, rv = _oracle(oracle, sc, 17)
    assert_bitwise(as_dtype(p, "f32"), rp, "hopper wall pos, 17 steps")
    assert_bitwise(as_dtype(v, "f32"), rv, "hopper wall vel, 17 steps")


def test_full_batch_equals_plain_runs_and_oracle(psx_mod, oracle):
    """All 256 scenes (4.6M spheres, 35.5M cells) in one batch with the tuned shape, replayed
    from a prepared 16-step graph: every scene equals its own plain run bit for bit, and a sample
    of scenes (the smallest, the largest and six more) equals the oracle."""
    scs = scenes.batched_boxes()
    steps = 16
    with psx_mod.Batch(scs) as b:
        _tuned(b)
        b.set_tuning(use_graph=1)
        b.prepare(steps)
        b.step(steps)
        b.synchronize()
        bp, bv = b.state()
    sizes = np.array([sc.radius.shape[0] for sc in scs])
    offs = np.concatenate([[0], np.cumsum(sizes)])
    assert offs[-1] == bp.shape[0] > 4_000_000
    for k, sc in enumerate(scs):
        with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
            s.step(steps)
            p, v = s.state()
        assert_bitwise(bp[offs[k]:offs[k + 1]], p, f"scene {k} pos vs plain")
        assert_bitwise(bv[offs[k]:offs[k + 1]], v, f"scene {k} vel vs plain")
    rng = np.random.default_rng(1)
    sample = {int(np.argmin(sizes)), int(np.argmax(sizes))} | set(rng.choice(len(scs), 6).tolist())
    for k in sorted(sample):
        rp, rv = _oracle(oracle, scs[k], steps)
        assert_bitwise(as_dtype(bp[offs[k]:offs[k + 1]], "f32"), rp, f"scene {k} pos vs oracle")
        assert_bitwise(as_dtype(bv[offs[k]:offs[k + 1]], "f32"), rv, f"scene {k} vel vs oracle")


def crowded_box(n_side=(128, 128, 64), spacing=0.8, seed=5, dtype="f32") -> Scene:
    """A bimodal pack (70 % r = 0.4, 30 % r = 0.8) on a lattice of spacing 0.8: with cells of
    2 r_max 1.05 = 1.68, two or three lattice planes cross every cell along each axis, so cells
    hold 8 to 27 spheres and most of them overflow their 8 id slots. A soft spring keeps the
    overlapping pack in place over the steps timed."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n_side
    n = nx * ny * nz
    k = np.arange(n)
    p = np.stack([k % nx, (k // nx) % ny, k // (nx * ny)], 1) * spacing + 1.0
    p = p + rng.uniform(-0.02, 0.02, size=(n, 3))
    r = np.where(rng.uniform(size=n) < 0.7, 0.4, 0.8)
    v = rng.normal(0.0, 0.01, size=(n, 3))
    hi = (nx * spacing + 2.0, ny * spacing + 2.0, nz * spacing + 2.0)
    prm = Params(gravity=(0.0, 0.0, -9.8), walls=(0.5,) * 6, spring=10.0, restitution=0.4,
                 dt=5e-5, box_min=(0.0, 0.0, 0.0), box_max=hi, dtype=dtype)
    return Scene(p, v, r, prm, mass_of(r))


def test_crowded_cells_bitexact_and_rank_cost_bounded(psx_mod, oracle):
    sc = crowded_box()
    n = sc.radius.shape[0]
    steps = 3
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        counts = s.bucketing()["count"]
        s.set_state(sc.pos, sc.vel)
        over = int((counts > 8).sum())
        assert over > 20000 and counts.max() >= 18
        s.step(steps)
        p, v = s.state()
        s.set_state(sc.pos, sc.vel)
        crowded_ms = s.profile(20)["k_rank_gather"]["ms"]
    rp, rv = _oracle(oracle, sc, steps)
    assert_bitwise(as_dtype(p, "f32"), rp, "crowded pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "crowded vel")
    # the same number of spheres on a lattice of spacing 1.8 (at most one per cell)
    lat = crowded_box(spacing=1.8)
    with psx_mod.Simulation(lat.pos, lat.vel, lat.radius, lat.mass, lat.params) as s:
        lattice_ms = s.profile(20)["k_rank_gather"]["ms"]
    ratio = crowded_ms / lattice_ms
    print(f"k_rank_gather, {n:,} spheres: crowded {1e3 * crowded_ms:.1f} us ({over:,} cells over "
          f"8 members, up to {int(counts.max())}), lattice {1e3 * lattice_ms:.1f} us, ratio {ratio:.2f}")
    assert ratio < 3.0


def test_prepared_graphs_equal_eager_run(psx_mod, oracle):
    """psx_prepare captures the graphs of a run without stepping; the prepared run and an eager
    run give the same state bit for bit, and preparing twice or with graphs off changes nothing."""
    sc = scenes.box_settle(dtype="f32")
    steps = 64 + 16 + 5
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        s.set_tuning(use_graph=1)
        s.prepare(steps)
        s.prepare(steps)
        p0, v0 = s.state()
        assert_bitwise(p0, np.float32(sc.pos).astype(np.float64), "prepare does not step")
        s.step(steps)
        pg, vg = s.state()
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) as s:
        s.set_tuning(use_graph=0)
        s.prepare(steps)
        s.step(steps)
        pe, ve = s.state()
    assert_bitwise(pg, pe, "pos")
    assert_bitwise(vg, ve, "vel")
    rp, rv = _oracle(oracle, sc, steps)
    assert_bitwise(as_dtype(pg, "f32"), rp, "pos vs oracle")
