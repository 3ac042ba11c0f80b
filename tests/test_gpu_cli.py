"""The CLI end to end on the GPU: the reference's flags, the sample CSV, CSV bytes identical to the
oracle's result written by the same writer."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from psx import io
from psx._lib import BIN_PATH
from psx.sim import Params

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("backend,semantics", [("--hip", "cpu"), ("--cuda", "cuda")])
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_cli_sample_matches_oracle(oracle, tmp_path, dtype, backend, semantics):
    """--hip runs the reference CPU engine's step, --cuda its CUDA engine's step
    (PSX_FLAG_SEMANTICS_CUDA)."""
    out = tmp_path / "out" / "result.csv"
    args = [str(BIN_PATH), str(GOLDEN / "sample.csv"), backend, "--dtype", dtype,
            "--initial-velocity", "0", "0", "-10", "--gravity", "0", "0", "-10",
            "--box-min", "-20", "-20", "-20", "--box-max", "20", "20", "20",
            "--coeff-of-spring", "50", "--coeff-of-restitution-sphere", "0.3",
            "--coeff-of-restitution-wall", "0.5", "0.5", "0.5", "0.5", "0.5", "0.5",
            "--dt", "1e-3", "--max-steps", "200", "--out-csv-path", str(out)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    assert "Simulation completed successfully." in r.stdout
    d = io.read_csv(GOLDEN / "sample.csv", velocity=(0, 0, -10), density=1.0)
    prm = Params(gravity=(0, 0, -10), walls=(0.5,) * 6, spring=50.0, restitution=0.3, dt=1e-3,
                 box_min=(-20, -20, -20), box_max=(20, 20, 20))
    rp, _ = oracle.run(prm.as_array(), d["position"], d["velocity"], d["radius"], d["mass"], 200,
                       dtype=dtype, semantics=semantics)
    expect = tmp_path / "expect.csv"
    io.write_csv(expect, rp.astype(np.float64), d["radius"])
    assert out.read_bytes() == expect.read_bytes()


def test_cli_mesh_container(oracle, tmp_path):
    from psx import meshes
    obj = tmp_path / "cyl.obj"
    meshes.write_obj(obj, meshes.open_cylinder(25.0, 60.0, 32, 4, 32))
    csv = tmp_path / "p.csv"
    rng = np.random.default_rng(2)
    pts = rng.uniform([-15, -15, 2], [15, 15, 40], size=(500, 3))
    io.write_csv(csv, pts, rng.uniform(0.4, 0.6, 500))
    out = tmp_path / "r.csv"
    r = subprocess.run([str(BIN_PATH), str(csv), "--container", str(obj), "--max-steps", "100",
                        "--dt", "1e-3", "--coeff-of-spring", "100", "--initial-velocity", "1", "0",
                        "-5", "--out-csv-path", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    d = io.read_csv(csv, velocity=(1, 0, -5))
    prm = Params(gravity=(0, 0, -10), walls=(0.5,) * 6, spring=100.0, restitution=0.1, dt=1e-3)
    rp, _ = oracle.run(prm.as_array(), d["position"], d["velocity"], d["radius"], d["mass"], 100,
                       mesh=io.read_obj(obj))
    expect = tmp_path / "e.csv"
    io.write_csv(expect, rp, d["radius"])
    assert out.read_bytes() == expect.read_bytes()


def test_cli_post_run_outputs(oracle, tmp_path):
    """--resolve-overlap-axis then --out-csv-path (ParticleStabilizerMain.cpp:275-291): the CSV
    equals the oracle's trajectory lifted by the oracle's resolveOverlaps, and the logged overlap
    check is the oracle's."""
    from oracle import ref
    csv = tmp_path / "r.csv"
    r = subprocess.run([str(BIN_PATH), str(GOLDEN / "sample.csv"), "--max-steps", "300",
                        "--dt", "1e-3", "--coeff-of-spring", "20", "--box-min", "-20", "-20", "-20",
                        "--box-max", "20", "20", "20", "--resolve-overlap-axis", "2",
                        "--out-csv-path", str(csv)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    d = io.read_csv(GOLDEN / "sample.csv", velocity=(0, 0, -10))
    prm = Params(gravity=(0, 0, -10), walls=(0.5,) * 6, spring=20.0, restitution=0.1, dt=1e-3,
                 box_min=(-20, -20, -20), box_max=(20, 20, 20))
    rp, rv = oracle.run(prm.as_array(), d["position"], d["velocity"], d["radius"], d["mass"], 300)
    # the statistics logged at the last step (PhysicsEngine.cpp:54-64), std::to_string format
    st = ref.statistics(rp, rv, d["radius"], d["mass"])
    fmt = lambda v: "%f" % v  # noqa: E731
    assert (f"Number of collided particles : {st['n_collided']} / {len(rp)}") in r.stdout
    assert ("Min position                 : (" + ", ".join(map(fmt, st["min"])) + ")") in r.stdout
    assert ("Max position                 : (" + ", ".join(map(fmt, st["max"])) + ")") in r.stdout
    assert f"Max collision distance       : {fmt(st['max_overlap'])}" in r.stdout
    ke = float(r.stdout.split("Total kinetic energy         : ")[1].split()[0])
    assert ke == pytest.approx(st["kinetic_energy"], rel=1e-6, abs=1e-6)
    lifted, nc, mo = ref.resolve_overlaps(rp, d["radius"], 2)
    assert not np.array_equal(lifted, rp)
    assert f"Number of collided particles: {nc} / {len(lifted)}" in r.stdout
    assert f"Max collision distance: {mo:.15f}" in r.stdout
    expect = tmp_path / "e.csv"
    io.write_csv(expect, lifted, d["radius"])
    assert csv.read_bytes() == expect.read_bytes()


def test_cli_obj_output(tmp_path):
    """--out-obj-path (ParticleStabilizerMain.cpp:293-296, ObjData.cpp:8-140) for two particles:
    line counts and sampled vertices on the oracle's sphere template at the final positions."""
    from oracle import ref
    src, csv, obj = tmp_path / "p.csv", tmp_path / "r.csv", tmp_path / "o" / "spheres.obj"
    io.write_csv(src, np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]), np.array([0.5, 0.75]))
    r = subprocess.run([str(BIN_PATH), str(src), "--max-steps", "10", "--out-csv-path", str(csv),
                        "--out-obj-path", str(obj)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    fin = io.read_csv(csv)
    ndivs, tpl = ref.obj_sphere_template(2)
    nv = tpl.shape[0]
    with open(obj) as f:
        lines = f.read().split("\n")
    assert len(lines) == 2 + 2 * nv + nv + 2 * nv // 3 and lines[0] == "o Generated particles"
    # the CSV holds 10-character fixed-point values, so the vertices are checked to that precision
    for p in range(2):
        for k in (0, 1, nv // 2, nv - 1):
            got = np.array(lines[1 + p * nv + k].split()[1:], float)
            want = np.array(ref.obj_vertex(fin["position"][p], fin["radius"][p], tpl[k]))
            assert np.allclose(got, want, atol=1e-4), (p, k)
