"""Post-run outputs of the library, on the host (no GPU): the overlap post-process
(psx_resolve_overlaps, ParticleModel.cpp:154-252) against the oracle's literal restatement, and the
sphere OBJ export (psx_write_obj_spheres, ObjData.cpp:8-140) against the oracle's template.

The oracle itself is pinned here by closed forms: two stacked spheres end at
x_j + sqrt((r_i + r_j)^2 - d0^2 - d1^2), a chain of three resolves bottom-up, the overlap check
equals the statistics oracle's independent count, and the sphere template lies on the unit sphere
with its first vertex at the pole."""
import math

import numpy as np
import pytest

from oracle import ref
from psx import io


def _pile(n, seed, spread=6.0, height=4.0, rmin=0.3, rmax=0.6):
    rng = np.random.default_rng(seed)
    pos = np.stack([rng.uniform(0, spread, n), rng.uniform(0, spread, n),
                    rng.uniform(0, height, n)], axis=1)
    radius = rng.uniform(rmin, rmax, n)
    return pos, radius


# ---- the oracle, pinned -------------------------------------------------------------------------

def test_oracle_two_spheres_closed_form():
    pos = np.array([[1.0, 2.0, 3.0], [1.3, 1.8, 3.5]])
    r = np.array([0.5, 0.4])
    out, nc, mo = ref.resolve_overlaps(pos, r, 2)
    lift = 3.0 + math.sqrt((0.9 * 0.9 - 0.3 * 0.3) - 0.2 * 0.2)
    assert out[1, 2] == pytest.approx(lift, rel=1e-15)
    assert np.array_equal(out[0], pos[0]) and np.array_equal(out[1, :2], pos[1, :2])
    assert nc == 0 or mo < 1e-12
    # along x: the sphere with the larger x moves
    out, _, _ = ref.resolve_overlaps(pos, r, 0)
    assert out[1, 0] == pytest.approx(1.0 + math.sqrt((0.81 - 0.04) - 0.25), rel=1e-15)


def test_oracle_chain_resolves_bottom_up():
    # three unit spheres on one vertical line, each overlapping the one below
    pos = np.array([[0.0, 0.0, 1.5], [0.0, 0.0, 0.0], [0.0, 0.0, 0.8]])
    r = np.ones(3)
    out, nc, mo = ref.resolve_overlaps(pos, r, 2)
    assert out[:, 2].tolist() == [4.0, 0.0, 2.0]
    assert nc == 0 and mo == 0.0


def test_oracle_overlap_check_matches_statistics_oracle():
    pos, r = _pile(400, 3)
    out, nc, mo = ref.resolve_overlaps(pos, r, 1)
    st = ref.statistics(out, np.zeros_like(out), r, np.ones_like(r))
    assert (nc, mo) == (st["n_collided"], st["max_overlap"])


def test_oracle_sphere_template_on_unit_sphere():
    ndivs, tpl = ref.obj_sphere_template(40000)
    assert ndivs == 6 and tpl.shape == (6 * 5 * 11, 3)
    assert np.array_equal(tpl[0], np.float32([0.0, 0.0, 1.0]))
    assert np.allclose(np.linalg.norm(tpl.astype(np.float64), axis=1), 1.0, atol=1e-6)
    # float64 restatement of the same vertices
    th = np.arange(5) * math.pi / 5
    ph = np.arange(11) * 2 * math.pi / 11
    q0 = np.stack([np.outer(np.sin(th), np.cos(ph)), np.outer(np.sin(th), np.sin(ph)),
                   np.repeat(np.cos(th)[:, None], 11, 1)], -1).reshape(-1, 3)
    assert np.allclose(tpl.reshape(-1, 6, 3)[:, 0], q0, atol=2e-6)


# ---- the library against the oracle --------------------------------------------------------------

@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("n,seed", [(300, 1), (1500, 2), (3000, 7)])
def test_resolve_overlaps_bitexact(n, seed, axis):
    pos, r = _pile(n, seed)
    want, nc_w, mo_w = ref.resolve_overlaps(pos, r, axis)
    got, nc, mo = io.resolve_overlaps(pos, r, axis)
    assert np.array_equal(got, want)
    assert (nc, mo) == (nc_w, mo_w)


@pytest.mark.parametrize("axis", [0, 2])
def test_resolve_overlaps_dense_heap_cascades(axis):
    # more sphere volume than space: lifts cascade and re-sort over long rank ranges
    pos, r = _pile(2000, 11, spread=5.0, height=2.0, rmin=0.25, rmax=0.5)
    want, nc_w, mo_w = ref.resolve_overlaps(pos, r, axis)
    got, nc, mo = io.resolve_overlaps(pos, r, axis)
    assert np.array_equal(got, want) and (nc, mo) == (nc_w, mo_w)
    assert (got[:, axis] - pos[:, axis]).max() > 5.0


def test_resolve_overlaps_far_outlier_without_grid():
    # one particle 1e13 away: too many cells for the grid, every rank is tested
    pos, r = _pile(500, 5)
    pos[17] = [1e13, 0.0, 0.0]
    want, nc_w, mo_w = ref.resolve_overlaps(pos, r, 2)
    got, nc, mo = io.resolve_overlaps(pos, r, 2)
    assert np.array_equal(got, want) and (nc, mo) == (nc_w, mo_w)


def test_resolve_overlaps_equal_radii_lattice_with_ties():
    # a jittered lattice of equal spheres with many equal coordinates (ties in the ranking)
    g = np.arange(8) * 0.9
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos[::3, 2] += 0.05
    r = np.full(pos.shape[0], 0.5)
    want, nc_w, mo_w = ref.resolve_overlaps(pos, r, 2)
    got, nc, mo = io.resolve_overlaps(pos, r, 2)
    assert np.array_equal(got, want) and (nc, mo) == (nc_w, mo_w)


def test_resolve_overlaps_lift_above_every_later_rank():
    # the middle sphere is lifted above the top one: no re-sort, the ranking stays out of order
    pos = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.5], [5.0, 0.0, 0.9], [5.0, 0.2, 1.3],
                    [0.1, 0.0, 1.2]])
    r = np.full(5, 0.5)
    want, nc_w, mo_w = ref.resolve_overlaps(pos, r, 2)
    got, nc, mo = io.resolve_overlaps(pos, r, 2)
    assert np.array_equal(got, want) and (nc, mo) == (nc_w, mo_w)
    assert got[1, 2] == 1.0


def test_resolve_overlaps_edge_cases():
    got, nc, mo = io.resolve_overlaps(np.zeros((0, 3)), np.zeros(0), 2)
    assert got.shape == (0, 3) and nc == 0 and mo == 0.0
    got, nc, mo = io.resolve_overlaps([[1.0, 2.0, 3.0]], [0.5], 0)
    assert got.tolist() == [[1.0, 2.0, 3.0]] and nc == 0
    with pytest.raises(RuntimeError, match="axis"):
        io.resolve_overlaps([[1.0, 2.0, 3.0]], [0.5], 3)
    with pytest.raises(RuntimeError, match="finite"):
        io.resolve_overlaps([[1.0, float("nan"), 3.0]], [0.5], 2)
    with pytest.raises(RuntimeError, match="non-negative"):
        io.resolve_overlaps([[1.0, 2.0, 3.0]], [-0.5], 2)
    # zero radii: nothing can overlap
    pos = np.zeros((4, 3))
    got, nc, mo = io.resolve_overlaps(pos, np.zeros(4), 2)
    assert np.array_equal(got, pos) and nc == 0 and mo == 0.0


# ---- sphere OBJ export ----------------------------------------------------------------------------

def test_write_obj_spheres_matches_template(tmp_path):
    pos = np.array([[1.25, -2.5, 3.0], [10.0, 20.0, 30.5]])
    r = np.array([0.75, 1.5])
    path = tmp_path / "sub" / "dir" / "spheres.obj"
    io.write_obj_spheres(path, pos, r)
    ndivs, tpl = ref.obj_sphere_template(2)
    nv = tpl.shape[0]
    nf = nv // 3
    assert ndivs == 354 and nf == 2 * (ndivs - 1) * (2 * ndivs - 1)
    lines = path.read_text().split("\n")
    assert lines[-1] == "" and lines[0] == "o Generated particles"
    lines = lines[1:-1]
    assert len(lines) == 2 * nv + nv + 2 * nf
    v, vn, f = lines[:2 * nv], lines[2 * nv:3 * nv], lines[3 * nv:]
    rng = np.random.default_rng(0)
    for k in np.concatenate([np.arange(50), rng.integers(0, nv, 400), np.arange(nv - 50, nv)]):
        for p in range(2):
            want = "v " + " ".join(ref.stream_float(x) for x in ref.obj_vertex(pos[p], r[p], tpl[k]))
            assert v[p * nv + k] == want, (p, k)
        assert vn[k] == "vn " + " ".join(ref.stream_float(float(x)) for x in tpl[k])
    for t in np.concatenate([np.arange(20), rng.integers(0, nf, 200)]):
        for p in range(2):
            o = p * nv + 3 * t
            want = f"f {o + 1}//{3 * t + 1} {o + 2}//{3 * t + 2} {o + 3}//{3 * t + 3}"
            assert f[p * nf + t] == want
    assert all(line.startswith("v ") for line in v) and all(line.startswith("f ") for line in f)


def test_write_obj_spheres_errors(tmp_path):
    blocker = tmp_path / "file"
    blocker.write_text("x")
    with pytest.raises(RuntimeError):
        io.write_obj_spheres(blocker / "a.obj", [[0.0, 0.0, 0.0]], [1.0])
