#!/usr/bin/env python3
"""Throughput of the per-step hot path on one MI355X.

Workload (default): the 262,144-sphere settle in an open-top cylinder mesh (2,176 triangles),
float32, inputs resident in HBM. A step is one pass of the hot path over all spheres: integrate,
bucket, container contact, 27-cell pair contact.

Prints ONE JSON line with particle-steps/s, the roofline of the dominant kernel (HIP events around
each launch on the library's stream, replaying exactly the timed steps) and the CPU baseline (the
oracle, a float64 restatement of the reference's CPU step, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "particle-stabilizer_amd"))
sys.path.insert(0, str(REPO))

import numpy as np  # noqa: E402

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
METRIC = "particle-steps/s on 1 GPU (plain vs batched); adjoint/forward step time ratio"


def algorithmic_bytes(n: int, cells: int, s: int) -> dict:
    """Bytes each kernel must move per launch (each array touched once; s = scalar bytes)."""
    return {
        "k_integrate": n * (13 * s + 4),
        "k_scan_reduce": 4 * cells,
        "k_scan_tiles": 0,
        "k_scan_down": 12 * cells,
        "k_scatter_ids": 16 * n,
        "k_rank_gather": n * (16 * s + 12),
        "k_pairs": 11 * s * n + 8 * cells,
    }


def cpu_baseline(sc, steps: int) -> dict:
    from oracle import ref
    threads = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    ref.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps, mesh=sc.mesh,
            dtype="f64", nthreads=threads)
    dt = time.perf_counter() - t0
    n = sc.radius.shape[0]
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{steps} float64 steps of the same {n:,}-sphere scene from its initial "
                      f"state (reference CPU algorithm: serial pair loop, OpenMP brute-force "
                      f"mesh loop), {dt:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="cylinder_settle",
                    choices=["cylinder_settle", "hopper_settle", "box_settle"])
    ap.add_argument("--dtype", default=None)
    ap.add_argument("--profile-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the path would shard by spatial slabs, see DESIGN.md)", file=sys.stderr)
        return 2

    import psx
    from psx import scenes

    gen = getattr(scenes, args.workload)
    kw = {"dtype": args.dtype} if args.dtype else {}
    sc = gen(**kw)
    n = sc.radius.shape[0]
    dtype = sc.params.dtype
    s_bytes = 4 if dtype == "f32" else 8

    sim = psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh)
    sim.set_tuning(use_graph=0 if args.no_graph else 1)
    sim.step(args.warmup)
    sim.synchronize()
    p0, v0 = sim.state()

    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed

    # replay the first profile_steps of the timed region with an event pair around every launch
    sim.set_state(p0, v0)
    psteps = max(1, min(args.profile_steps, args.steps))
    prof = sim.profile(psteps)
    cells = sim.num_cells
    algo = algorithmic_bytes(n, cells, s_bytes)
    kernels = {}
    for name, d in prof.items():
        ms = d["ms"]
        b = algo.get(name, 0)
        kernels[name] = {"ms": round(ms, 5), "algo_bytes": b,
                         "GBps": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 and b else None}
    dom = max(prof, key=lambda k: prof[k]["ms"])
    dom_ms = prof[dom]["ms"]
    achieved = algo[dom] / (dom_ms * 1e-3)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": None, "avg_launch_ms": round(dom_ms, 5),
                "algo_bytes_per_launch": algo[dom], "profiled_steps": psteps}
    sim.close()

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n:,} spheres"
                               + (f", {sc.mesh.shape[0]:,}-triangle mesh container" if sc.mesh is not None else ", box container")
                               + f", {dtype}", "n_particles": n, "cells": cells,
                   "graph": not args.no_graph},
        "roofline": roofline,
        "kernels": kernels,
    }
    if not args.no_cpu_baseline:
        cpu_steps = args.cpu_steps or {"cylinder_settle": 20, "hopper_settle": 3,
                                       "box_settle": 400}[args.workload]
        line["cpu_baseline"] = cpu_baseline(sc, cpu_steps)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
