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
# launch shapes chosen by psx.tune.autotune on MI355X for these workloads (applied without timing;
# --retune sweeps again)
TUNE_CACHE = REPO / "particle-stabilizer_amd" / "psx" / "tuning_mi355x.json"
RETUNE = False


def tuned(sim) -> dict:
    from psx.tune import autotune
    return autotune(sim, cache=TUNE_CACHE, force=RETUNE)["tuning"]


METRIC = "particle-steps/s on 1 GPU (plain vs batched); adjoint/forward step time ratio"


def algorithmic_bytes(n: int, cells: int, s: int, blocks: bool) -> dict:
    """Bytes each kernel must move per launch (each array touched once; s = scalar bytes; blocks:
    the per-cell id-block form of the cell lists, else the dense lists). In a run of steps the pair
    kernel also does the next step's hash pass (ids read, cells written, counts updated, ids filed
    into the id blocks); k_integrate only starts a run."""
    return {
        "k_integrate": n * (13 * s + 16),
        "k_scan": (12 if blocks else 8) * cells,
        "k_scatter_ids": 0 if blocks else 16 * n + 8 * cells,
        "k_rank_gather": n * (16 * s + 24),
        "k_pairs": 11 * s * n + 8 * cells + n * (12 if blocks else 8) + 8 * cells,
    }


def measured_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the plain-copy calibration), or None."""
    f = REPO / "profiles" / f"pmc_{kernel}_{workload}.txt"
    if not f.exists():
        return None
    for row in f.read_text().splitlines():
        if row.startswith("traffic_bytes_per_launch:"):
            return float(row.split(":", 1)[1])
    return None


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        for row in open("/proc/cpuinfo"):
            if row.startswith("model name"):
                return row.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, steps: int, seconds: float = 0.0) -> dict:
    """The oracle (the reference CPU algorithm) on the host cores, from the scene's initial state:
    a run of `steps` steps, then, when `seconds` is set and that run was shorter, one run whose
    length that first run sizes to about `seconds` of work (the figure is that second run's)."""
    from oracle import ref
    threads = min(16, os.cpu_count() or 1)

    def run(k: int) -> float:
        t0 = time.perf_counter()
        ref.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, k, mesh=sc.mesh,
                dtype="f64", nthreads=threads)
        return time.perf_counter() - t0

    dt = run(steps)
    sized = ""
    if seconds > 0.0 and dt < seconds:
        first = steps
        steps = max(steps, int(round(steps * seconds / dt)))
        dt = run(steps)
        sized = f"; length from a {first}-step run"
    n = sc.radius.shape[0]
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "cpu": cpu_model(),
            "sample": f"{steps} float64 steps of the same {n:,}-sphere scene from its initial "
                      f"state (reference CPU algorithm: serial pair loop, OpenMP brute-force "
                      f"mesh loop), {dt:.2f} s{sized}"}


def _timed(sim, steps: int) -> float:
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    return time.perf_counter() - t0


def bench_hopper(steps: int = 100, warmup: int = 10) -> dict:
    """The 4,194,304-sphere hopper settle (8,192-triangle mesh), float32."""
    import psx
    from psx import scenes
    sc = scenes.hopper_settle()
    n = sc.radius.shape[0]
    sim = psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh)
    tuning = tuned(sim)
    sim.step(warmup)
    p0, v0 = sim.state()
    dt = _timed(sim, steps)
    sim.set_state(p0, v0)
    prof = sim.profile(min(steps, 50))
    sv = tuning.get("sort_variant", 0)
    algo = algorithmic_bytes(n, sim.num_cells, 4, sv != 2)
    kern = {k: {"ms": round(d["ms"], 4), "launches_per_step": round(d["per_step"], 4),
                "GBps": round(algo[k] / (d["ms"] * 1e-3) / 1e9, 1) if algo.get(k) and d["ms"] else None}
            for k, d in prof.items() if d["per_step"] > 0}
    sim.close()
    pm = prof["k_pairs"]["ms"]
    return {"workload": f"hopper_settle: {n:,} spheres, 8,192-triangle mesh, f32",
            "value": round(n * steps / dt, 1), "unit": "particle-steps/s",
            "ms_per_step": round(1e3 * dt / steps, 4), "steps": steps, "kernels": kern,
            "tuning": tuning, "k_pairs_frac_hbm": round(algo["k_pairs"] / (pm * 1e-3) / HBM_PEAK, 4)}


def bench_batched(steps: int = 100, warmup: int = 10) -> dict:
    """256 independent box scenes of uneven size: one launch per kernel per step (hipGraph)
    against the same kernels launched once per scene, on one stream and on one stream per scene."""
    import psx
    from psx import scenes
    scs = scenes.batched_boxes()
    total = sum(sc.radius.shape[0] for sc in scs)
    b = psx.Batch(scs)
    tuning = tuned(b)
    b.step(warmup)
    tb = _timed(b, steps)
    cells = b.num_cells
    b.close()
    sims = [psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) for sc in scs]
    import ctypes
    for s in sims:
        s.set_tuning(**tuning)
        s.step(warmup)
    for s in sims:
        s.synchronize()
    # (a) one stream: each scene's launches in turn, as a loop over scenes would issue them
    st = ctypes.c_void_p()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate(ctypes.byref(st))
    for s in sims:
        s.set_stream(st.value)
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in sims:
            s.step(1)
    for s in sims:
        s.synchronize()
    t_one = time.perf_counter() - t0
    # (b) one stream per scene (the scenes' launches may overlap on the GPU)
    for s in sims:
        s.set_stream(None)
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in sims:
            s.step(1)
    for s in sims:
        s.synchronize()
    t_many = time.perf_counter() - t0
    for s in sims:
        s.close()
    hip.hipStreamDestroy(st)
    rate = total * steps / tb
    return {"workload": f"batched_boxes: 256 scenes, {total:,} spheres, {cells:,} cells, f32",
            "value": round(rate, 1), "unit": "particle-steps/s", "steps": steps,
            "ms_per_step": round(1e3 * tb / steps, 4),
            "per_scene_one_stream": round(total * steps / t_one, 1),
            "per_scene_stream_each": round(total * steps / t_many, 1),
            "speedup_vs_one_stream": round(t_one / tb, 2),
            "speedup_vs_stream_each": round(t_many / tb, 2), "tuning": tuning}


def bench_adjoint(nsteps: int = 2000, every: int = 100) -> dict:
    """16,384 spheres in an 80-triangle box mesh, float32: forward window with checkpoints every
    100 steps against the backward (segment recompute + adjoint)."""
    import torch

    import psx
    from psx import scenes
    sc = scenes.calibration_scene()
    n = sc.radius.shape[0]
    sim = psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh)
    tuned(sim)
    dev = torch.device("cuda")
    gx = torch.randn((n, 3), device=dev, dtype=torch.float32)
    gv = torch.randn((n, 3), device=dev, dtype=torch.float32)
    ox = torch.empty_like(gx)
    ov = torch.empty_like(gv)
    torch.cuda.synchronize()
    sim.window_forward(min(nsteps, 200), every)  # warm-up (allocations, graph capture)
    sim.window_backward(gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
    sim.set_state(sc.pos, sc.vel)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.window_forward(nsteps, every)
    sim.synchronize()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = sim.window_backward(gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
    sim.synchronize()
    tbk = time.perf_counter() - t0
    st = sim.window_stats()
    sim.close()
    return {"workload": f"calibration_scene: {n:,} spheres, 80-triangle box mesh, f32, "
                        f"{nsteps}-step window, checkpoint every {every}",
            "forward_ms": round(1e3 * tf, 2), "backward_ms": round(1e3 * tbk, 2),
            "ratio": round(tbk / tf, 3), "recompute_ms": round(st["replay_ms"], 2),
            "adjoint_ms_per_step": round(st["adjoint_ms"] / nsteps, 5),
            "forward_ms_per_step": round(1e3 * tf / nsteps, 5),
            "grad_finite": bool(np.isfinite(g).all())}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="cylinder_settle",
                    choices=["cylinder_settle", "hopper_settle", "box_settle"])
    ap.add_argument("--dtype", default=None)
    ap.add_argument("--profile-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps of the CPU baseline (default: a run sized to --cpu-seconds)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0,
                    help="approximate length of the CPU baseline's timed run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes instead of the tuned ones")
    ap.add_argument("--retune", action="store_true",
                    help="sweep the launch shapes again instead of applying the tuning file")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the hopper, batched and adjoint measurements")
    ap.add_argument("--extras", default="hopper_settle,batched,adjoint",
                    help="comma-separated extras to run (hopper_settle, batched, adjoint)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="finish with one plain device-to-device copy of this many MiB (the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the "
              f"path would shard the cell order into x-slabs, one GPU each, exchanging a layer of "
              f"boundary cells per step)", file=sys.stderr)
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
    global RETUNE
    RETUNE = args.retune
    if not args.no_tune:
        tuned(sim)
        sim.set_tuning(use_graph=0 if args.no_graph else 1)
    tuning = sim.tuning
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
    sv = tuning.get("sort_variant", 0)
    blocks = sv != 2
    algo = algorithmic_bytes(n, cells, s_bytes, blocks)
    kernels = {}
    for name, d in prof.items():
        if d["per_step"] == 0:
            continue
        ms = d["ms"]
        b = algo.get(name, 0)
        kernels[name] = {"ms": round(ms, 5), "launches_per_step": round(d["per_step"], 4),
                         "algo_bytes": b,
                         "GBps": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 and b else None}
    dom = max(prof, key=lambda k: prof[k]["ms"] * prof[k]["per_step"])
    dom_ms = prof[dom]["ms"]
    achieved = algo[dom] / (dom_ms * 1e-3)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": measured_traffic(args.workload, dom), "avg_launch_ms": round(dom_ms, 5),
                "algo_bytes_per_launch": algo[dom], "profiled_steps": psteps}
    stats = None
    if not args.no_extras:
        # ParticleStatistics::calculateStatistics (ParticleStatistics.cpp:3-43) of the state, on
        # the GPU grid: one bucketing pass plus the statistics kernel, results copied to the host
        sim.statistics()
        t0 = time.perf_counter()
        for _ in range(20):
            st = sim.statistics()[0]
        stats = {"ms_per_call": round((time.perf_counter() - t0) / 20 * 1e3, 4),
                 "n_collided": st["n_collided"], "max_overlap": st["max_overlap"],
                 "kinetic_energy": st["kinetic_energy"]}
    sim.close()

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "particle-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"{args.workload}: {n:,} spheres"
                               + (f", {sc.mesh.shape[0]:,}-triangle mesh container" if sc.mesh is not None else ", box container")
                               + f", {dtype}", "n_particles": n, "cells": cells,
                   "graph": not args.no_graph, "tuning": tuning},
        "roofline": roofline,
        "kernels": kernels,
    }
    if not args.no_extras:
        extras = {}
        for name, fn in (("hopper_settle", bench_hopper), ("batched", bench_batched),
                         ("adjoint", bench_adjoint)):
            if name not in args.extras.split(","):
                continue
            try:
                extras[name] = fn()
            except Exception as e:  # report, keep the headline line
                extras[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"bench.py: {name} done", file=sys.stderr, flush=True)
        extras["statistics"] = stats
        line["extras"] = extras
    if not args.no_cpu_baseline:
        cpu_steps = args.cpu_steps or {"cylinder_settle": 20, "hopper_settle": 3,
                                       "box_settle": 400}[args.workload]
        line["cpu_baseline"] = cpu_baseline(sc, cpu_steps,
                                            0.0 if args.cpu_steps else args.cpu_seconds)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    if args.copy_calibration_mb:
        import torch
        nbytes = args.copy_calibration_mb << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench.py: copy calibration, {nbytes} bytes read and written", file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
