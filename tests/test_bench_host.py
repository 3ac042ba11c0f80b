"""Host logic of bench.py (no GPU): the algorithmic bytes behind roofline.achieved, the measured
traffic read from the committed rocprofv3 summary, the CPU description, the length of the CPU
baseline's sample and the one-GPU rule."""
import subprocess
import sys
from pathlib import Path

import bench

REPO = Path(__file__).resolve().parent.parent


def test_pair_kernel_bytes_of_the_262k_cylinder():
    # 262,144 spheres, 48 x 48 x 179 cells, binary32, id slots: 11 s N for the pair kick, 8 C for
    # the offset reads, 12 N for the hash pass's ids and cells and 8 N for its count updates (one
    # count read and written per sphere, not every cell of the grid)
    b = bench.algorithmic_bytes(262144, 48 * 48 * 179, 4, 1)
    assert b["k_pairs"] == 44 * 262144 + 8 * 412416 + 20 * 262144 == 20076544
    assert b["k_scan"] == 12 * 412416 and b["k_scatter_ids"] == 0
    dense = bench.algorithmic_bytes(262144, 412416, 4, 2)
    assert dense["k_scan"] == 8 * 412416
    assert dense["k_scatter_ids"] == 16 * 262144 + 8 * 412416
    assert dense["k_pairs"] == b["k_pairs"] - 4 * 262144  # no id filed in the dense form
    # binary64 moves twice the scalar bytes of every per-sphere array and the same index bytes
    b64 = bench.algorithmic_bytes(262144, 412416, 8, 1)
    assert b64["k_pairs"] - b["k_pairs"] == 11 * 4 * 262144
    assert b64["k_rank_gather"] - b["k_rank_gather"] == 16 * 4 * 262144


def test_pair_kernel_bytes_of_the_hopper():
    # 4,194,304 spheres on the 179 x 179 x 358 grid: the per-launch figure behind the headline
    # roofline (64 N + 8 C)
    n, c = 4194304, 179 * 179 * 358
    assert bench.algorithmic_bytes(n, c, 4, 1)["k_pairs"] == 64 * n + 8 * c == 360200880
    assert bench.algorithmic_bytes(n, c, 4, 0) == bench.algorithmic_bytes(n, c, 4, 1)
    # the incremental form: the hash pass writes the cells only (4 N), the scan reads the counts
    # and writes the offsets (8 C), the rank kernel reads a mark and a kept rank and writes a rank
    # per sphere (one byte each)
    inc = bench.algorithmic_bytes(n, c, 4, 3)
    assert inc["k_pairs"] == 48 * n + 8 * c == 293092016
    assert inc["k_scan"] == 8 * c
    assert inc["k_rank_gather"] == bench.algorithmic_bytes(n, c, 4, 1)["k_rank_gather"] + 3 * n


def test_cpu_threads_follow_the_cpu_share(monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "3")
    t, host = bench.cpu_threads()
    assert t == min(3, len(__import__("os").sched_getaffinity(0)))
    assert host["nproc"] >= 1 and host["affinity_cores"] >= t and host["model"]
    monkeypatch.delenv("OMP_NUM_THREADS")
    t, host = bench.cpu_threads()
    assert t == host["affinity_cores"] and host["omp_num_threads"] is None


def test_traffic_read_from_the_committed_summary():
    text = (REPO / "profiles" / "pmc_k_pairs_cylinder_settle.txt").read_text()
    line = [r for r in text.splitlines() if r.startswith("traffic_bytes_per_launch:")][0]
    assert bench.measured_traffic("cylinder_settle", "k_pairs") == float(line.split(":")[1])
    assert bench.measured_traffic("no_such_workload", "k_pairs") is None


def test_cpu_description():
    assert isinstance(bench.cpu_model(), str) and bench.cpu_model()


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_cpu_baseline_sized_to_the_requested_seconds():
    """The CPU baseline's timed run is sized by a first run to about the requested seconds; with
    no target it times the given steps as they are."""
    from helpers import dense_box
    sc = dense_box(n_side=6, dtype="f64")
    fixed = bench.cpu_baseline(sc, 2)
    assert fixed["sample"].startswith("2 float64 steps") and "length from" not in fixed["sample"]
    assert fixed["kind"] == "port" and fixed["value"] > 0
    sized = bench.cpu_baseline(sc, 2, seconds=0.3)
    steps = int(sized["sample"].split()[0])
    assert steps > 2 and "length from a 2-step run" in sized["sample"]
    secs = float(sized["sample"].split("), ")[1].split(" s")[0])
    assert 0.05 < secs < 3.0
