"""Full-size workloads: bucketing bit-exact against the oracle at 262,144 and 4,194,304 spheres,
a few bit-exact steps of the 262,144-sphere cylinder settle, the binary32 settle within 1e-3 of a
radius of the binary64 oracle over 1,000 steps, a full box settle's packing fraction and height
histogram, and size-independent properties. The long oracle runs use its fast mode (same results
as its reference loops, pinned in test_oracle_pins.py)."""
import numpy as np
import pytest

from helpers import as_dtype, assert_bitwise, oracle_run
from psx import scenes

pytestmark = pytest.mark.gpu


def _bucket_check(psx_mod, oracle, sc):
    v = sc.mesh.reshape(-1, 3)
    rmax = float(np.float32(sc.radius).max())
    ref = oracle.bucketing(rmax, v.min(0), v.max(0), np.float32(sc.pos), dtype="f32")
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        for variant in (1, 2):  # per-cell id blocks, dense lists
            s.set_tuning(sort_variant=variant)
            got = s.bucketing()
            for k in ("cell", "count", "offset", "perm"):
                assert_bitwise(got[k], ref[k], f"{k} sort_variant={variant}")
        nc = s.num_cells
    assert got["offset"][-1] == sc.radius.shape[0] and nc == got["count"].shape[0]
    return got


def test_cylinder_bucketing_bitexact(psx_mod, oracle):
    _bucket_check(psx_mod, oracle, scenes.cylinder_settle())


def test_hopper_bucketing_bitexact(psx_mod, oracle):
    got = _bucket_check(psx_mod, oracle, scenes.hopper_settle())
    perm = got["perm"]
    assert np.array_equal(np.sort(perm), np.arange(perm.shape[0]))     # a permutation
    cells_sorted = got["cell"][perm]
    assert (np.diff(cells_sorted) >= 0).all()                          # sorted by cell
    same = np.diff(cells_sorted) == 0
    assert (np.diff(perm)[same] > 0).all()                              # stable inside a cell


def test_cylinder_steps_bitexact(psx_mod, oracle):
    sc = scenes.cylinder_settle()
    steps = 3
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        s.set_tuning(use_graph=1)
        s.step(steps)
        s.synchronize()
        p, v = s.state()
    rp, rv = oracle_run(oracle, sc, steps)
    assert_bitwise(as_dtype(p, "f32"), rp, "pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "vel")


@pytest.mark.parametrize("chunk", [4, 2])
def test_compressed_cylinder_steps_bitexact(psx_mod, oracle, chunk):
    """The 262,144-sphere cylinder with its heights compressed to 0.6 (vertical spacing 1.02,
    0.88 contacts per sphere, and the vertical neighbours of every sphere within the prefilter's
    reach): the pair kernel's rare path runs on most spheres. Binary32, bit-exact against the
    oracle, with four- and two-wide chunks."""
    sc = scenes.cylinder_packed()
    steps = 3
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        s.set_tuning(use_graph=1, pair_variant=1, pair_chunk=chunk)
        s.step(steps)
        p, v = s.state()
    rp, rv = oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                        mesh=sc.mesh, dtype="f32", fast=True)
    assert_bitwise(as_dtype(p, "f32"), rp, "pos")
    assert_bitwise(as_dtype(v, "f32"), rv, "vel")


def test_cylinder_f32_within_1e3_radius_of_f64_oracle_1000_steps(psx_mod, oracle):
    """The headline workload in binary32 over 1,000 steps against the reference algorithm in
    binary64 (the oracle): every sphere within 1e-3 of its radius. The binary64 GPU run of the
    same scene is bit-exact against the same oracle run."""
    sc = scenes.cylinder_settle()
    steps = 1000
    rp, rv = oracle.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps,
                        mesh=sc.mesh, dtype="f64", fast=True)
    with psx_mod.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh) as s:
        s.set_tuning(use_graph=1)
        s.step(steps)
        p32, _ = s.state()
    dev = np.linalg.norm(p32 - rp, axis=1) / sc.radius
    print(f"max |x_f32 - x_f64| / r over {steps} steps: {dev.max():.3e}")
    assert dev.max() <= 1e-3
    sc64 = scenes.cylinder_settle(dtype="f64")
    with psx_mod.Simulation(sc64.pos, sc64.vel, sc64.radius, sc64.mass, sc64.params, sc64.mesh) as s:
        s.set_tuning(use_graph=1)
        s.step(steps)
        p64, v64 = s.state()
    assert_bitwise(p64, rp, "pos f64")
    assert_bitwise(v64, rv, "vel f64")


def _packing(pos, r, area, top_q=95):
    """Packing fraction under the 95th percentile of the spheres' tops, and a 30-bin height
    histogram (fractions of the spheres) over the box height."""
    top = np.percentile(pos[:, 2] + r, top_q)
    phi = (4.0 / 3.0 * np.pi * r ** 3).sum() / (area * top)
    hist, _ = np.histogram(pos[:, 2], bins=30, range=(0.0, 60.0))
    return phi, hist / pos.shape[0], pos[:, 2].mean()


def test_box_settle_packing_matches_f64_oracle(psx_mod, oracle):
    """The 4,096-sphere box settle (20,000 steps, the lattice falls, lands and packs): the binary32
    GPU run and the binary64 oracle end in the same packing fraction and height histogram (the
    trajectories themselves diverge, as chaotic contact dynamics do), and the binary64 GPU run is
    bit-exact against the oracle."""
    steps = 20000
    sc64 = scenes.box_settle(dtype="f64")
    rp, rv = oracle.run(sc64.params.as_array(), sc64.pos, sc64.vel, sc64.radius, sc64.mass, steps,
                        dtype="f64", fast=True)
    with psx_mod.Simulation(sc64.pos, sc64.vel, sc64.radius, sc64.mass, sc64.params) as s:
        s.set_tuning(use_graph=1)
        s.step(steps)
        p64, v64 = s.state()
    assert_bitwise(p64, rp, "pos f64")
    assert_bitwise(v64, rv, "vel f64")
    sc32 = scenes.box_settle(dtype="f32")
    with psx_mod.Simulation(sc32.pos, sc32.vel, sc32.radius, sc32.mass, sc32.params) as s:
        s.set_tuning(use_graph=1)
        s.step(steps)
        p32, _ = s.state()
    phi_r, hist_r, zm_r = _packing(rp, sc64.radius, 20.0 * 20.0)
    phi_g, hist_g, zm_g = _packing(p32, sc32.radius, 20.0 * 20.0)
    l1 = np.abs(hist_g - hist_r).sum()
    print(f"packing fraction f32 {phi_g:.5f} f64 oracle {phi_r:.5f}; histogram L1 {l1:.4f}; "
          f"mean height {zm_g:.4f} / {zm_r:.4f}")
    assert np.percentile(rp[:, 2], 95) < 20.0  # landed: the lattice started at z >= 20
    assert abs(phi_g / phi_r - 1.0) <= 0.01
    assert l1 <= 0.05
    assert abs(zm_g - zm_r) <= 0.05
