"""Host logic of the autotuner (psx/tune.py) against a stand-in handle: the fastest shape of each
kernel is chosen, the state is restored, and the JSON cache is written and honoured."""
import json

import numpy as np
import pytest

from psx import tune


class FakeSim:
    """Duck-typed handle whose kernel times are a known function of the launch shape."""

    def __init__(self):
        self.n = 1000
        self.dtype = np.float32
        self.num_cells = 4096
        self.num_scenes = 1
        self.tuning = {"block_integrate": 256, "block_sort": 256, "block_pairs": 256,
                       "use_graph": 0, "pair_variant": 0, "sort_variant": 0, "pair_chunk": 0}
        self.pos = np.arange(3000.0).reshape(1000, 3)
        self.vel = -self.pos
        self.profiled = 0
        self.restored = None

    def set_tuning(self, **kw):
        self.tuning.update(kw)

    def state(self):
        return self.pos.copy(), self.vel.copy()

    def set_state(self, p, v):
        self.restored = (p, v)

    def profile(self, steps):
        self.profiled += 1
        t = self.tuning
        g = max(1, t["pair_variant"])
        dense = t["sort_variant"] == 2
        return {
            "k_integrate": {"ms": 0.01 + abs(t["block_integrate"] - 512) * 1e-5 - 0.004 * dense,
                            "per_step": 1},
            "k_scan": {"ms": 0.004, "per_step": 1},
            "k_rank_gather": {"ms": 0.005 + abs(t["block_sort"] - 128) * 1e-6, "per_step": 1},
            "k_pairs": {"ms": 0.02 + abs(g - 4) * 1e-3 + abs(t["block_pairs"] - 256) * 1e-6
                        + (0.0 if t["pair_chunk"] in (0, 4) else 2e-4), "per_step": 1},
            "k_scatter_ids": {"ms": 0.003 if dense else 0.0, "per_step": 1 if dense else 0},
        }


def test_picks_fastest_shapes_and_restores_state():
    sim = FakeSim()
    out = tune.autotune(sim, steps=2)
    # the hash kernel only starts a run: dense lists cost the scatter kernel's 0.003 ms per step
    assert out["tuning"] == {"block_integrate": 512, "block_sort": 128, "block_pairs": 256,
                             "pair_variant": 4, "pair_chunk": 4, "sort_variant": 1, "use_graph": 1}
    assert sim.tuning["pair_variant"] == 4 and sim.tuning["use_graph"] == 1
    assert sim.restored is not None
    np.testing.assert_array_equal(sim.restored[0], sim.pos)
    np.testing.assert_array_equal(sim.restored[1], sim.vel)
    # every pair shape and every block size was timed
    assert sim.profiled == len(tune.PAIR_SHAPES) * 2 * 3


def test_pair_kernel_cost_taken_per_cell_list_form():
    """k_integrate only starts a run (few launches per step) and the pair kernel's time depends on
    the cell-list form (it does the next step's hash pass): the plan weighs each kernel by its
    launches per step and takes the pair kernel's time of the same form."""
    class RunSim(FakeSim):
        def profile(self, steps):
            out = super().profile(steps)
            dense = self.tuning["sort_variant"] == 2
            out["k_integrate"]["per_step"] = 1.0 / steps
            out["k_integrate"]["ms"] += 0.0 if dense else 0.05  # id blocks: a slow hash kernel
            out["k_pairs"]["ms"] += 0.0 if not dense else 0.005  # dense: slower pair kernel
            return out

    out = tune.autotune(RunSim(), steps=100)
    # id blocks: a slower hash kernel that a steady run does not launch, 0.008 ms less per step
    # (0.005 pair kernel + 0.003 scatter kernel)
    assert out["tuning"]["sort_variant"] == 1


def test_dense_lists_win_on_a_cheaper_pair_kernel():
    class DenseSim(FakeSim):
        def profile(self, steps):
            out = super().profile(steps)
            if self.tuning["sort_variant"] != 2:
                out["k_pairs"]["ms"] += 0.004  # id blocks: the hash pass waits on its atomics
            return out

    assert tune.autotune(DenseSim(), steps=4)["tuning"]["sort_variant"] == 2


def test_incremental_form_wins_on_a_cheaper_pair_kernel_and_scan():
    """The incremental id slots (sort_variant 3): a pair kernel without most of its hash pass's
    atomics and a scan that leaves the counts alone, against a slower rank kernel."""
    class IncSim(FakeSim):
        def profile(self, steps):
            out = super().profile(steps)
            if self.tuning["sort_variant"] == 3:
                out["k_pairs"]["ms"] -= 0.004
                out["k_rank_gather"]["ms"] += 0.002
            return out

    out = tune.autotune(IncSim(), steps=4)
    assert out["tuning"]["sort_variant"] == 3
    assert out["kernel_ms"]["step_incremental"] < out["kernel_ms"]["step_id_blocks"]

    class SlowRankSim(IncSim):
        def profile(self, steps):
            out = super().profile(steps)
            if self.tuning["sort_variant"] == 3:
                out["k_rank_gather"]["ms"] += 0.003  # more than the pair kernel saves
            return out

    assert tune.autotune(SlowRankSim(), steps=4)["tuning"]["sort_variant"] == 1


def test_two_wide_chunks_chosen_when_faster():
    class ChunkSim(FakeSim):
        def profile(self, steps):
            out = super().profile(steps)
            t = self.tuning
            if t["pair_variant"] in (0, 1) and t["pair_chunk"] == 2:
                out["k_pairs"]["ms"] = 0.001  # the one-lane, two-wide kernel wins
            return out

    tuning = tune.autotune(ChunkSim(), steps=2)["tuning"]
    assert tuning["pair_variant"] == 1 and tuning["pair_chunk"] == 2


def test_cache_written_and_reused(tmp_path):
    cache = tmp_path / "tune.json"
    sim = FakeSim()
    first = tune.autotune(sim, steps=1, cache=cache)
    assert not first["cached"]
    data = json.loads(cache.read_text())
    assert tune.signature(sim) in data
    assert "float32" in tune.signature(sim)
    sim2 = FakeSim()
    again = tune.autotune(sim2, cache=cache)
    assert again["cached"] and sim2.profiled == 0
    assert again["tuning"] == first["tuning"]
    assert sim2.tuning["pair_variant"] == first["tuning"]["pair_variant"]
    forced = tune.autotune(FakeSim(), steps=1, cache=cache, force=True)
    assert not forced["cached"]


def test_signature_separates_workloads():
    a, b = FakeSim(), FakeSim()
    b.num_cells = 8192
    assert tune.signature(a) != tune.signature(b)


@pytest.mark.parametrize("bad", ["{not json", ""])
def test_unreadable_cache_is_ignored(tmp_path, bad):
    cache = tmp_path / "tune.json"
    cache.write_text(bad)
    out = tune.autotune(FakeSim(), steps=1, cache=cache)
    assert not out["cached"]
