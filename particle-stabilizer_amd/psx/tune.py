"""Autotuner for the launch shapes of a simulation handle.

The kernels of one step are independent in their launch shapes, so every candidate shape of every
kernel is timed in the same sweep: round k applies the k-th candidate of each kernel (cycling
through the shorter lists), profiles a few steps with the kernels' own dispatch timestamps, and
each kernel keeps the shape with the smallest median time over its rounds. Tuned knobs:

- block_integrate: workgroup size of the integrate / hash / container kernel;
- block_sort: workgroup size of the rank / gather / container kernel (at most 512 in binary64);
- (block_pairs, pair_variant): workgroup size and lanes per particle of the pair kernel.

The cell size stays what the reference's grid rule fixes (it changes the neighbour search, not
only the speed), and every launch shape gives bit-identical results, so tuning never changes a
trajectory. The state is saved before the sweep and restored after it. The chosen shape can be
kept in a JSON file keyed by the workload's signature (dtype, particles, cells, scenes, device).
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np

BLOCKS = (64, 128, 256, 512, 1024)
PAIR_SHAPES = tuple((b, g) for g in (1, 2, 4, 8, 16) for b in (64, 128, 256, 512, 1024))


def signature(sim) -> str:
    """Key of a workload in the tuning cache."""
    dev = os.environ.get("PSX_TUNE_DEVICE", "gfx950")
    return f"{dev}|{np.dtype(sim.dtype).name}|n={sim.n}|cells={sim.num_cells}|scenes={sim.num_scenes}"


def _load(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def autotune(sim, steps: int = 8, rounds_per_shape: int = 2, use_graph: bool = True,
             cache: str | os.PathLike | None = None, force: bool = False) -> dict:
    """Times the candidate launch shapes on `sim`, applies the fastest and returns
    {"tuning": {...}, "kernel_ms": {...}, "cached": bool}.

    steps: steps profiled per round. cache: optional JSON path; a stored entry for the same
    signature is applied without timing unless force is set.
    """
    key = signature(sim)
    path = Path(cache) if cache else None
    if path is not None and not force:
        entry = _load(path).get(key)
        if entry:
            sim.set_tuning(**entry["tuning"])
            return {"tuning": dict(entry["tuning"]), "kernel_ms": entry.get("kernel_ms", {}),
                    "cached": True}

    p0, v0 = sim.state()
    best: dict[str, tuple[float, object]] = {}
    n_rounds = max(len(BLOCKS), len(PAIR_SHAPES)) * rounds_per_shape
    samples: dict[tuple[str, object], list] = {}
    # the binary64 rank/gather kernel takes workgroups of at most 512
    sort_blocks = BLOCKS if np.dtype(sim.dtype) == np.float32 else tuple(b for b in BLOCKS if b <= 512)
    for k in range(n_rounds):
        bi = BLOCKS[k % len(BLOCKS)]
        bs = sort_blocks[(k // len(BLOCKS)) % len(sort_blocks)]
        bp, g = PAIR_SHAPES[k % len(PAIR_SHAPES)]
        sim.set_tuning(block_integrate=bi, block_sort=bs, block_pairs=bp, pair_variant=g,
                       use_graph=0)
        prof = sim.profile(steps)
        for name, d in prof.items():
            shape = {"k_integrate": bi, "k_rank_gather": bs, "k_pairs": (bp, g)}.get(name)
            if shape is None:
                continue
            samples.setdefault((name, shape), []).append(d["ms"])
    for (name, shape), ms in samples.items():
        t = sorted(ms)[len(ms) // 2]  # median over the rounds that used this shape
        if name not in best or t < best[name][0]:
            best[name] = (t, shape)
    bs_best = best["k_rank_gather"][1]
    bp_best, g_best = best["k_pairs"][1]
    tuning = {"block_integrate": int(best["k_integrate"][1]), "block_sort": int(bs_best),
              "block_pairs": int(bp_best), "pair_variant": int(g_best),
              "use_graph": 1 if use_graph else 0}
    sim.set_tuning(**tuning)
    sim.set_state(p0, v0)
    kernel_ms = {k: best[k][0] for k in ("k_integrate", "k_rank_gather", "k_pairs")}
    if path is not None:
        data = _load(path)
        data[key] = {"tuning": tuning, "kernel_ms": kernel_ms}
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    return {"tuning": tuning, "kernel_ms": kernel_ms, "cached": False}
