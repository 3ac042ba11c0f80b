"""Autotuner for the launch shapes of a simulation handle.

The kernels of one step are independent in their launch shapes, so every candidate shape of every
kernel is timed in the same sweep: round k applies the k-th candidate of each kernel (cycling
through the shorter lists), profiles a few steps with the kernels' own dispatch timestamps, and
each kernel keeps the shape with the smallest median time over its rounds. Tuned knobs:

- block_integrate: workgroup size of the integrate / hash / container kernel;
- block_sort: workgroup size of the rank / gather / container kernel (at most 512 in binary64);
- (block_pairs, pair_variant, pair_chunk): workgroup size and lanes per particle of the pair
  kernel, or its LDS-staged form (pair_variant 32, workgroups of 64 to 256), and the one-lane
  kernel's candidates per round (4 or 2);
- sort_variant: per-cell id slots, dense lists or incremental id slots for the cell lists (a third
  of the rounds each; the form whose kernels take less time per step wins, the pair kernel
  included, since it does the next step's hash pass).

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
PAIR_SHAPES = (tuple((b, g, 4) for g in (1, 2, 4, 8, 16) for b in (64, 128, 256, 512, 1024))
               + tuple((b, 1, 2) for b in (64, 128, 256, 512, 1024))
               + tuple((b, 32, 4) for b in (64, 128, 256)))


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
    forms = (1, 2, 3)  # id slots, dense lists, incremental id slots
    n_form = max(len(BLOCKS), len(PAIR_SHAPES)) * rounds_per_shape
    samples: dict[tuple, list] = {}
    launches: dict[tuple, list] = {}  # (kernel, sort variant) -> launches per step
    # the binary64 rank/gather kernel takes workgroups of at most 512
    sort_blocks = BLOCKS if np.dtype(sim.dtype) == np.float32 else tuple(b for b in BLOCKS if b <= 512)
    for k in range(len(forms) * n_form):  # the rounds of each cell-list form in turn
        sv = forms[k // n_form]
        bi = BLOCKS[k % len(BLOCKS)]
        bs = sort_blocks[(k // len(BLOCKS)) % len(sort_blocks)]
        bp, g, w = PAIR_SHAPES[k % len(PAIR_SHAPES)]
        sim.set_tuning(block_integrate=bi, block_sort=bs, block_pairs=bp, pair_variant=g,
                       pair_chunk=w, sort_variant=sv, use_graph=0)
        prof = sim.profile(steps)
        for name, d in prof.items():
            # the pair kernel also does the next step's hash pass, whose cost depends on the
            # cell-list form: its samples are kept per form
            shape = {"k_integrate": (bi, sv), "k_scan": (0, sv), "k_rank_gather": (bs, sv),
                     "k_scatter_ids": (bs, sv), "k_pairs": (bp, g, w, sv)}.get(name)
            per = d.get("per_step", 1)
            if shape is None or per == 0:
                continue
            samples.setdefault((name,) + shape, []).append(d["ms"])
            launches.setdefault((name, sv), []).append(per)

    def med(name, *key):
        ms = samples.get((name,) + key)
        return sorted(ms)[len(ms) // 2] if ms else None

    def rate(name, sv):
        """Launches per step. k_integrate only starts a run of steps (after the state or the
        tuning changes, as at the start of each profiled round here): none in a steady run."""
        if name == "k_integrate":
            return 0.0
        w = launches.get((name, sv))
        return sum(w) / len(w) if w else 0.0

    # per cell-list form: the fastest shape of each kernel, weighted by its launches per step
    plans = {}
    for sv in forms:
        bi_ms = {bi: med("k_integrate", bi, sv) for bi in BLOCKS}
        bi_ms = {b: t for b, t in bi_ms.items() if t is not None}
        bs_ms = {}
        for bs in sort_blocks:
            r = med("k_rank_gather", bs, sv)
            sc = med("k_scatter_ids", bs, sv) if sv == 2 else 0.0
            if r is not None and sc is not None:
                bs_ms[bs] = r * rate("k_rank_gather", sv) + sc * rate("k_scatter_ids", sv)
        pair_sv = {(bp, g, w): med("k_pairs", bp, g, w, sv) for bp, g, w in PAIR_SHAPES}
        pair_sv = {k: t for k, t in pair_sv.items() if t is not None}
        scan = (med("k_scan", 0, sv) or 0.0) * rate("k_scan", sv)
        if bi_ms and bs_ms and pair_sv:
            bi = min(bi_ms, key=bi_ms.get)
            bs = min(bs_ms, key=bs_ms.get)
            bp, g, w = min(pair_sv, key=pair_sv.get)
            total = (bi_ms[bi] * rate("k_integrate", sv) + bs_ms[bs] + scan
                     + pair_sv[(bp, g, w)] * rate("k_pairs", sv))
            plans[sv] = (total, bi, bs, bp, g, w, pair_sv[(bp, g, w)])
    sv_best = min(plans, key=lambda v: plans[v][0])
    _, bi_best, bs_best, bp_best, g_best, w_best, pair_best = plans[sv_best]
    tuning = {"block_integrate": int(bi_best), "block_sort": int(bs_best),
              "block_pairs": int(bp_best), "pair_variant": int(g_best),
              "pair_chunk": int(w_best),
              "sort_variant": int(sv_best), "use_graph": 1 if use_graph else 0}
    sim.set_tuning(**tuning)
    sim.set_state(p0, v0)
    kernel_ms = {"step_id_blocks": plans.get(1, (None,))[0],
                 "step_dense_lists": plans.get(2, (None,))[0],
                 "step_incremental": plans.get(3, (None,))[0],
                 "k_pairs": pair_best}
    if path is not None:
        data = _load(path)
        data[key] = {"tuning": tuning, "kernel_ms": kernel_ms}
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(data, indent=1, sort_keys=True))
    return {"tuning": tuning, "kernel_ms": kernel_ms, "cached": False}


def _main() -> int:
    """python -m psx.tune --out FILE [--workloads ...]: tunes the standard workloads of
    psx.scenes on this GPU and stores their launch shapes in FILE."""
    import argparse
    import sys

    from . import scenes
    from .sim import Batch, Simulation

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--out", required=True)
    ap.add_argument("--workloads", nargs="*",
                    default=["cylinder_settle", "hopper_settle", "batched_boxes",
                             "calibration_scene"])
    args = ap.parse_args()
    for name in args.workloads:
        made = getattr(scenes, name)()
        if isinstance(made, list):
            sim = Batch(made)
        else:
            sim = Simulation(made.pos, made.vel, made.radius, made.mass, made.params, made.mesh)
        out = autotune(sim, cache=args.out, force=True)
        print(f"{name}: {signature(sim)} -> {out['tuning']}", file=sys.stderr, flush=True)
        sim.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
