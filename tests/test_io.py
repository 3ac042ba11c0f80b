"""Host-side formats of the library (no GPU): CSV particles (ParticleIO.cpp:8-144), OBJ
containers, and the CLI's argument handling."""
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from psx import io, meshes
from psx._lib import BIN_PATH, PsxError

GOLDEN = Path(__file__).resolve().parent / "golden"


def _expected_field(v: float) -> str:
    dec = 10 - len(str(int(abs(v)))) - (1 if v < 0 else 0) - 1
    return f"{v:.{max(dec, 0)}f}"


def test_csv_writer_format(tmp_path):
    rng = np.random.default_rng(0)
    pos = np.concatenate([rng.uniform(-30, 30, size=(50, 3)),
                          [[0.0, -0.0, 1e-9], [123456.789, -9.99999999, 99999999.5],
                           [-0.5, 5.0, 12345678901.0]]])
    r = np.concatenate([rng.uniform(0.3, 5, 50), [1.0, 2.0, 0.5]])
    f = tmp_path / "sub" / "dir" / "out.csv"
    io.write_csv(f, pos, r)
    lines = f.read_text().split("\n")
    assert lines[0] == "x,y,z,radius" and lines[-1] == ""
    for row, (p, rr) in enumerate(zip(pos, r)):
        fields = lines[row + 1].split(",")
        assert fields == [_expected_field(float(x)) for x in (*p, rr)]
        for x, fs in zip((*p, rr), fields):
            if abs(x) < 1e8:
                assert len(fs) <= 11  # 10, or 11 when rounding carries into a new digit / -0.0
                dec = len(fs.split(".")[1]) if "." in fs else 0
                assert abs(float(fs) - x) <= 0.5 * 10.0 ** (-dec) * (1 + 1e-12)


def test_csv_reader_reference_sample():
    d = io.read_csv(GOLDEN / "sample.csv", velocity=(0.0, 0.0, -10.0), density=1.0)
    assert d["position"].shape == (2410, 3) and d["skipped"] == 0
    vals, counts = np.unique(d["radius"], return_counts=True)
    assert dict(zip(vals.tolist(), counts.tolist())) == {1.0: 2127, 2.0: 266, 5.0: 17}
    assert np.array_equal(d["position"][0], [8.85649, -13.00788, -1.18328])
    assert (d["velocity"] == [0.0, 0.0, -10.0]).all()
    assert np.allclose(d["mass"], 4.0 / 3.0 * math.pi * d["radius"] ** 3, rtol=1e-15)


def test_csv_reader_token_rules(tmp_path):
    f = tmp_path / "p.csv"
    f.write_text("x,y,z,r\n"
                 "1,2,3,0.5\n"
                 " 4.5 , -1e-3,  7 ,1\n"        # stod skips leading blanks
                 "1,2,3\n"                      # 3 tokens: skipped
                 "1,,2,3,0.25\n"                # empty token dropped -> 4 tokens
                 "null,1,2,3,0.75\n"            # "null" dropped -> 4 tokens
                 "\n"                           # 0 tokens: skipped
                 "9,8,7,2,1\n")                 # 5 tokens: skipped
    d = io.read_csv(f, velocity=(1.0, 2.0, 3.0), density=2.0)
    assert d["skipped"] == 3
    assert np.array_equal(d["position"], [[1, 2, 3], [4.5, -1e-3, 7], [1, 2, 3], [1, 2, 3]])
    assert np.array_equal(d["radius"], [0.5, 1.0, 0.25, 0.75])
    assert np.allclose(d["mass"], 2.0 * 4 / 3 * math.pi * d["radius"] ** 3, rtol=1e-15)
    bad = tmp_path / "bad.csv"
    bad.write_text("x,y,z,r\n1,2,abc,0.5\n")
    with pytest.raises(PsxError) as e:
        io.read_csv(bad)
    assert e.value.code == 6
    with pytest.raises(PsxError):
        io.read_csv(tmp_path / "missing.csv")


def test_csv_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    pos = rng.uniform(-19, 19, size=(200, 3))
    r = rng.uniform(0.5, 5, 200)
    f = tmp_path / "rt.csv"
    io.write_csv(f, pos, r)
    d = io.read_csv(f)
    assert np.allclose(d["position"], pos, atol=5e-8) and np.allclose(d["radius"], r, atol=5e-9)


def test_obj_roundtrip_generated(tmp_path):
    for tri in (meshes.open_cylinder(), meshes.box_mesh(), meshes.hopper(n_az=16, n_cone=2, n_cyl=2, n_floor=16)):
        f = tmp_path / "m.obj"
        meshes.write_obj(f, tri)
        assert np.array_equal(io.read_obj(f), tri)


def test_obj_polygons_and_index_forms(tmp_path):
    f = tmp_path / "q.obj"
    f.write_text("# quad + triangle\no box\nv 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvn 0 0 1\n"
                 "f 1//1 2//1 3//1 4//1\nf -4/1/1 -2/1/1 -1/1/1\n")
    t = io.read_obj(f)
    assert t.shape == (3, 9)
    assert np.array_equal(t[0], [0, 0, 0, 1, 0, 0, 1, 1, 0])
    assert np.array_equal(t[1], [0, 0, 0, 1, 1, 0, 0, 1, 0])
    assert np.array_equal(t[2], [0, 0, 0, 1, 1, 0, 0, 1, 0])
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(PsxError):
        io.read_obj(bad)


def _cli(*args):
    return subprocess.run([str(BIN_PATH), *map(str, args)], capture_output=True, text=True,
                          timeout=60)


def test_cli_help_and_errors():
    for h in ("-h", "--help"):
        r = _cli(h)
        assert r.returncode == 0 and "--coeff-of-restitution-wall" in r.stdout
    r = _cli()
    assert r.returncode == 0 and "usage" in r.stdout
    r = _cli("--help")
    assert "--cuda" in r.stdout and "CUDA engine's step" in r.stdout
    r = _cli(GOLDEN / "sample.csv", "--gravity", "0", "0")
    assert r.returncode == 1 and "needs 3" in r.stderr
    r = _cli(GOLDEN / "sample.csv", "--dt", "abc")
    assert r.returncode == 1 and "not a number" in r.stderr
