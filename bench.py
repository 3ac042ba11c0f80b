#!/usr/bin/env python3
"""Throughput of the per-step hot path on one MI355X.

Workload (default): the 4,194,304-sphere settle in a cone-bottom hopper mesh (8,192 triangles),
float32, inputs resident in HBM -- the largest single-GPU configuration. A step is one pass of the
hot path over all spheres: integrate, bucket, container contact, 27-cell pair contact.

Prints ONE JSON line with particle-steps/s, the roofline of the dominant kernel (HIP dispatch
events around each launch on the library's stream, replaying the first timed steps) and the CPU
baseline (the oracle, a float64 restatement of the reference's CPU step with its loops, timed on
this host's cores). Every hipGraph the timed run replays is captured before the clock starts.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
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

METRIC = "particle-steps/s on 1 GPU (plain vs batched); adjoint/forward step time ratio"

# CPU baseline windows (steps of the oracle from the scene's initial state): the whole 4,096-sphere
# box settle, a window of the 262k cylinder and of the 4.2M hopper
CPU_STEPS = {"box_settle": 20000, "cylinder_settle": 20, "hopper_settle": 3}


def tuned(sim) -> dict:
    from psx.tune import autotune
    return autotune(sim, cache=TUNE_CACHE, force=RETUNE)["tuning"]


def algorithmic_bytes(n: int, cells: int, s: int, form: int) -> dict:
    """Bytes each kernel must move per launch (each array touched once; s = scalar bytes; form:
    the cell-list form, sort_variant 1 or 0 per-cell id slots, 2 dense lists, 3 incremental id
    slots). In a run of steps the pair kernel also does the next step's hash pass: ids read, cells
    written, ids filed into the id slots (4 N each) and one count read and written per sphere
    (8 N); in the incremental form only the cells are written (4 N), the spheres that change cell
    (a few per thousand in a settle) being left out; k_integrate only starts a run. The pair kick
    itself reads the sorted state (8 s N) and the cell offsets (8 C) and writes the new velocities
    (3 s N). The incremental form's scan leaves the counts alone (8 C instead of 12 C) and its rank
    kernel reads each sphere's cell mark and kept rank and writes its new rank (3 N more)."""
    blocks = form != 2
    inc = form == 3
    return {
        "k_integrate": n * (13 * s + 16),
        "k_scan": (12 if blocks and not inc else 8) * cells,
        "k_scatter_ids": 0 if blocks else 16 * n + 8 * cells,
        "k_rank_gather": n * (16 * s + 24 + (3 if inc else 0)),
        "k_pairs": 11 * s * n + 8 * cells + (4 * n if inc else n * (12 if blocks else 8) + 8 * n),
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
    """The host CPU's model name (lscpu, else /proc/cpuinfo), or "unknown"."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for row in out.splitlines():
            if row.startswith("Model name:"):
                return row.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        for row in open("/proc/cpuinfo"):
            if row.startswith("model name"):
                return row.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> tuple[int, dict]:
    """Threads for the CPU baseline: the cores this process may run on, capped by
    OMP_NUM_THREADS where it is set (a GPU box grants each GPU a share of its host's cores and
    says so there); and the host facts behind that choice."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(avail, int(omp)) if omp.isdigit() and int(omp) > 0 else avail
    return max(threads, 1), {"nproc": os.cpu_count(), "affinity_cores": avail,
                             "omp_num_threads": omp or None, "model": cpu_model()}


def cpu_run(sc, steps: int, threads: int) -> float:
    """Seconds of `steps` oracle steps in float64 from the scene's initial state, with the
    reference's loops (serial pair loop, OpenMP brute-force O(N T) mesh loop)."""
    from oracle import ref
    t0 = time.perf_counter()
    ref.run(sc.params.as_array(), sc.pos, sc.vel, sc.radius, sc.mass, steps, mesh=sc.mesh,
            dtype="f64", nthreads=threads)
    return time.perf_counter() - t0


def cpu_baseline(sc, steps: int, seconds: float = 0.0, threads: int | None = None) -> dict:
    """The oracle (the reference CPU algorithm) on the host cores, from the scene's initial state:
    a run of `steps` steps, then, when `seconds` is set and that run was shorter, one run whose
    length that first run sizes to about `seconds` of work (the figure is that second run's)."""
    host = None
    if threads is None:
        threads, host = cpu_threads()
    dt = cpu_run(sc, steps, threads)
    sized = ""
    if seconds > 0.0 and dt < seconds:
        first = steps
        steps = max(steps, int(round(steps * seconds / dt)))
        dt = cpu_run(sc, steps, threads)
        sized = f"; length from a {first}-step run"
    n = sc.radius.shape[0]
    out = {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
           "sample": f"{steps} float64 steps of the same {n:,}-sphere scene from its initial "
                     f"state (reference CPU algorithm: serial pair loop, OpenMP brute-force "
                     f"mesh loop), {dt:.2f} s{sized}"}
    if host:
        out["cpu"] = host["model"]
        out["host"] = host
    return out


def _timed(sim, steps: int) -> float:
    """Wall time of `steps` steps, graphs captured beforehand."""
    sim.synchronize()
    sim.prepare(steps)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    return time.perf_counter() - t0


def bench_plain(workload: str, steps: int, warmup: int, profile_steps: int, dtype=None,
                graph: bool = True, tune: bool = True) -> dict:
    """One plain workload: warm-up, graphs prepared, `steps` timed steps; then the first
    `profile_steps` of them replayed with dispatch events around every launch."""
    import psx
    from psx import scenes
    sc = getattr(scenes, workload)(**({"dtype": dtype} if dtype else {}))
    n = sc.radius.shape[0]
    s_bytes = 4 if sc.params.dtype == "f32" else 8
    sim = psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh)
    if tune:
        tuned(sim)
    sim.set_tuning(use_graph=1 if graph else 0)
    tuning = sim.tuning
    sim.step(warmup)
    sim.synchronize()
    p0, v0 = sim.state()
    elapsed = _timed(sim, steps)
    # replay the first profile_steps of the timed region with an event pair around every launch
    sim.set_state(p0, v0)
    psteps = max(1, min(profile_steps, steps))
    prof = sim.profile(psteps)
    cells = sim.num_cells
    algo = algorithmic_bytes(n, cells, s_bytes, tuning.get("sort_variant", 0))
    kernels = {}
    for name, d in prof.items():
        if d["per_step"] == 0:
            continue
        ms, b = d["ms"], algo.get(name, 0)
        kernels[name] = {"ms": round(ms, 5), "launches_per_step": round(d["per_step"], 4),
                         "algo_bytes": b,
                         "GBps": round(b / (ms * 1e-3) / 1e9, 1) if ms > 0 and b else None,
                         "frac_hbm": round(b / (ms * 1e-3) / HBM_PEAK, 4) if ms > 0 and b else None}
    dom = max(prof, key=lambda k: prof[k]["ms"] * prof[k]["per_step"])
    dom_ms = prof[dom]["ms"]
    achieved = algo[dom] / (dom_ms * 1e-3)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved / 1e9, 1),
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": measured_traffic(workload, dom), "avg_launch_ms": round(dom_ms, 5),
                "algo_bytes_per_launch": algo[dom], "profiled_steps": psteps}
    # ParticleStatistics::calculateStatistics (ParticleStatistics.cpp:3-43) of the state on the
    # GPU grid: one bucketing pass plus the statistics kernel, results copied to the host
    sim.statistics()
    t0 = time.perf_counter()
    for _ in range(5):
        stats = sim.statistics()[0]
    stats["ms_per_call"] = round((time.perf_counter() - t0) / 5 * 1e3, 4)
    sim.close()
    desc = (f"{workload}: {n:,} spheres"
            + (f", {sc.mesh.shape[0]:,}-triangle mesh container" if sc.mesh is not None
               else ", box container") + f", {sc.params.dtype}")
    return {"scene": sc, "n": n, "cells": cells, "dtype": sc.params.dtype, "workload": desc,
            "elapsed": elapsed, "value": n * steps / elapsed, "tuning": tuning,
            "kernels": kernels, "roofline": roofline, "stats": stats}


def bench_batched(steps: int = 128, warmup: int = 10, graph: bool = True) -> dict:
    """256 independent box scenes of uneven size: one launch per kernel per step (hipGraph)
    against the same kernels run once per scene -- launched step by step (on one stream, and on
    one stream per scene) and replayed from each scene's own 64-step graphs (one stream per
    scene). Every graph is captured before its clock starts. graph=False launches every kernel
    on its own (the batch included) and skips the per-scene graphs."""
    import ctypes

    import psx
    from psx import scenes
    scs = scenes.batched_boxes()
    total = sum(sc.radius.shape[0] for sc in scs)
    b = psx.Batch(scs)
    tuning = tuned(b)
    b.set_tuning(use_graph=1 if graph else 0)
    b.step(warmup)
    tb = _timed(b, steps)
    cells = b.num_cells
    b.close()
    sims = [psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params) for sc in scs]
    for s in sims:
        s.set_tuning(**{**tuning, "use_graph": 1 if graph else 0})
        s.step(warmup)
    for s in sims:
        s.synchronize()

    def eager(order_sims) -> float:
        t0 = time.perf_counter()
        for _ in range(steps):
            for s in order_sims:
                s.step(1)
        for s in order_sims:
            s.synchronize()
        return time.perf_counter() - t0

    # (a) one stream: each scene's launches in turn, as a loop over scenes would issue them
    st = ctypes.c_void_p()
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamCreate(ctypes.byref(st))
    for s in sims:
        s.set_stream(st.value)
    t_one = eager(sims)
    # (b) one stream per scene (the scenes' launches may overlap on the GPU)
    for s in sims:
        s.set_stream(None)
    t_many = eager(sims)
    # (c) each scene replays its own captured graphs on its own stream, 64 steps per launch
    t_graph = None
    if graph:
        chunk = 64
        for s in sims:
            s.prepare(chunk)
        for s in sims:
            s.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            k = min(chunk, steps - done)
            for s in sims:
                s.step(k)
            done += k
        for s in sims:
            s.synchronize()
        t_graph = time.perf_counter() - t0
    for s in sims:
        s.close()
    hip.hipStreamDestroy(st)
    rate = total * steps / tb
    return {"workload": f"batched_boxes: 256 scenes, {total:,} spheres, {cells:,} cells, f32",
            "value": round(rate, 1), "unit": "particle-steps/s", "steps": steps,
            "ms_per_step": round(1e3 * tb / steps, 4),
            "per_scene_one_stream": round(total * steps / t_one, 1),
            "per_scene_stream_each": round(total * steps / t_many, 1),
            "per_scene_graphs_stream_each": round(total * steps / t_graph, 1) if t_graph else None,
            "speedup_vs_one_stream": round(t_one / tb, 2),
            "speedup_vs_stream_each": round(t_many / tb, 2),
            "speedup_vs_graphs_stream_each": round(t_graph / tb, 2) if t_graph else None,
            "graph": graph, "tuning": tuning}


def bench_adjoint(nsteps: int = 2000, every: int = 100, graph: bool = True) -> dict:
    """16,384 spheres in an 80-triangle box mesh, float32: forward window with checkpoints every
    100 steps against the backward (segment recompute + adjoint); graph: replay the steps and
    the backward's segments from hipGraphs (else every kernel is launched on its own)."""
    import torch

    import psx
    from psx import scenes
    sc = scenes.calibration_scene()
    n = sc.radius.shape[0]
    sim = psx.Simulation(sc.pos, sc.vel, sc.radius, sc.mass, sc.params, sc.mesh)
    tuned(sim)
    sim.set_tuning(use_graph=1 if graph else 0)
    dev = torch.device("cuda")
    gx = torch.randn((n, 3), device=dev, dtype=torch.float32)
    gv = torch.randn((n, 3), device=dev, dtype=torch.float32)
    ox = torch.empty_like(gx)
    ov = torch.empty_like(gv)
    torch.cuda.synchronize()
    # warm-up: allocations and the capture of every graph the timed windows replay
    wid = sim.window_forward(min(nsteps, 200), every)
    sim.window_backward(wid, gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
    sim.set_state(sc.pos, sc.vel)
    sim.synchronize()
    t0 = time.perf_counter()
    wid = sim.window_forward(nsteps, every)
    sim.synchronize()
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = sim.window_backward(wid, gx.data_ptr(), gv.data_ptr(), ox.data_ptr(), ov.data_ptr())
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
            "graph": graph, "grad_finite": bool(np.isfinite(g).all())}


def _plain_summary(r: dict, steps: int) -> dict:
    return {"workload": r["workload"], "value": round(r["value"], 1), "unit": "particle-steps/s",
            "ms_per_step": round(1e3 * r["elapsed"] / steps, 5), "steps": steps,
            "kernels": r["kernels"], "roofline": r["roofline"], "tuning": r["tuning"]}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="hopper_settle",
                    choices=["hopper_settle", "cylinder_settle", "box_settle"])
    ap.add_argument("--dtype", default=None)
    ap.add_argument("--profile-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="steps of the headline CPU baseline (default: the whole 4,096-sphere "
                         "box settle, 20 steps of the cylinder, 3 of the hopper)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-legs", action="store_true",
                    help="time the CPU baseline of the headline workload only")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shapes instead of the tuned ones")
    ap.add_argument("--retune", action="store_true",
                    help="sweep the launch shapes again instead of applying the tuning file")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the cylinder, batched, adjoint and statistics measurements")
    ap.add_argument("--extras", default="cylinder_settle,batched,adjoint",
                    help="comma-separated extras to run (cylinder_settle, hopper_settle, "
                         "batched, adjoint)")
    ap.add_argument("--adjoint-graph", type=int, default=1, choices=[0, 1],
                    help="1: the adjoint extra replays hipGraphs; 0: it launches every kernel")
    ap.add_argument("--batched-graph", type=int, default=1, choices=[0, 1],
                    help="1: the batched extra replays hipGraphs; 0: it launches every kernel "
                         "(and skips the per-scene graph baseline)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="finish with one plain device-to-device copy of this many MiB (the "
                         "reference kernel for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the "
              f"path would shard the cell order into x-slabs, one GPU each, exchanging a layer of "
              f"boundary cells per step)", file=sys.stderr)
        return 2

    global RETUNE
    RETUNE = args.retune
    head = bench_plain(args.workload, args.steps, args.warmup, args.profile_steps, args.dtype,
                       graph=not args.no_graph, tune=not args.no_tune)
    print(f"bench.py: {args.workload} done", file=sys.stderr, flush=True)
    sc = head["scene"]
    line = {
        "metric": METRIC, "value": round(head["value"], 1), "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1e3 * head["elapsed"] / args.steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": head["dtype"], "data": "synthetic",
        "config": {"workload": head["workload"], "n_particles": head["n"], "cells": head["cells"],
                   "graph": not args.no_graph, "tuning": head["tuning"]},
        "roofline": head["roofline"],
        "kernels": head["kernels"],
    }
    if not args.no_extras:
        extras = {}
        st = head["stats"]
        extras["statistics"] = {"ms_per_call": st["ms_per_call"], "n_collided": st["n_collided"],
                                "max_overlap": st["max_overlap"],
                                "kinetic_energy": st["kinetic_energy"]}
        for name in args.extras.split(","):
            if not name or name == args.workload:
                continue
            try:
                if name in ("cylinder_settle", "hopper_settle"):
                    k = {"cylinder_settle": 5000, "hopper_settle": 2000}[name]
                    extras[name] = _plain_summary(
                        bench_plain(name, k, 50, 1000), k)
                elif name == "batched":
                    extras[name] = bench_batched(graph=bool(args.batched_graph))
                elif name == "adjoint":
                    extras[name] = bench_adjoint(graph=bool(args.adjoint_graph))
                else:
                    extras[name] = {"error": "unknown extra"}
            except Exception as e:  # report, keep the headline line
                extras[name] = {"error": f"{type(e).__name__}: {e}"}
            print(f"bench.py: {name} done", file=sys.stderr, flush=True)
        line["extras"] = extras
    if not args.no_cpu_baseline:
        threads, host = cpu_threads()
        cpu_steps = args.cpu_steps or CPU_STEPS[args.workload]
        base = cpu_baseline(sc, cpu_steps, threads=threads)
        base["cpu"] = host["model"]
        base["host"] = host
        print("bench.py: cpu baseline done", file=sys.stderr, flush=True)
        if not args.no_cpu_legs:
            from psx import scenes
            legs = {}
            for name, steps in CPU_STEPS.items():
                if name == args.workload:
                    continue
                leg = cpu_baseline(getattr(scenes, name)(dtype="f64"), steps, threads=threads)
                legs[name] = {k: leg[k] for k in ("value", "unit", "cores", "sample")}
                print(f"bench.py: cpu leg {name} done", file=sys.stderr, flush=True)
            base["legs"] = legs
        line["cpu_baseline"] = base
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
