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


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_cli_sample_matches_oracle(oracle, tmp_path, dtype):
    out = tmp_path / "out" / "result.csv"
    args = [str(BIN_PATH), str(GOLDEN / "sample.csv"), "--hip", "--dtype", dtype,
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
                       dtype=dtype)
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
