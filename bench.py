#!/usr/bin/env python3
"""Benchmark of the DEM time step on one MI355X.

Default workload: the 4M-disc settling case -- 4,194,304 FP32 discs on a
2048 x 2048 jittered lattice (log-normal radii, r0 = 5 mm), 4731 x 2366 =
11,193,546 cells of 2 rmax, gravity settling.  A "step" is one pass of
hash -> sort -> cell fill -> reorder/cellStart -> collide -> integrate over
all discs, replayed from a captured hipGraph, with the state resident in
HBM before the timed region starts.

Prints ONE JSON line:
  value            particle-steps/s of the timed steps (graph replays)
  roofline         the dominant kernel's algorithmic bytes per launch over
                   its average launch time (HIP events on the launch
                   stream, eager steps of the same workload)
  cpu_baseline     the CPU oracle (a port of the reference kernels) timed
                   on this host on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "particle-steps/s @4M discs; batched-ensemble and fwd+adjoint vs plain step"

# Algorithmic (compulsory) HBM bytes per particle and per cell for each
# stage, FP32 state with per-particle radius and 32-bit keys:
#   hash      R pos 8;                         W key 4, index 4
#   sort      R key+index 8;                   W key+index 8        (one pass)
#   fill      W cellStart 4 per cell
#   reorder   R key 4, index 4, state 36;      W state 36, loc 4
#             (state = pos 8, vel 8, theta 4, omega 4, id 4, type 4, radius 4)
#   collide   R pos 8, vel 8, omega 4, radius 4, type 4; W acc 8, alpha 4, pressure 4
#             + R cellStart/cellEnd 8 per cell
#   integrate R pos, vel, acc 24, theta, omega, alpha, type, radius, pressure 24;
#             W pos, vel 16, theta, omega 8
BYTES_PER_PARTICLE = {"hash": 16, "sort": 16, "cell_fill": 0, "reorder": 84, "collide": 44, "integrate": 72}
BYTES_PER_CELL = {"hash": 0, "sort": 0, "cell_fill": 4, "reorder": 0, "collide": 8, "integrate": 0}


def algorithmic_bytes(stage: str, n: int, cells: int) -> int:
    return BYTES_PER_PARTICLE[stage] * n + BYTES_PER_CELL[stage] * cells


def make_scene(name: str):
    from demso_amd import scene as sc

    if name == "settling":
        return sc.settling(), "settling: 4,194,304 log-normal discs (r0=5 mm), 4731x2366 cells, FP32"
    if name == "column_collapse":
        return sc.column_collapse(), "column collapse: 4,096 discs, 91x46 cells, FP32"
    if name == "ensemble":
        s = sc.ensemble()
        return s, f"ensemble: 256 instances, {s.n} discs, {s.total_cells} cells, FP32"
    raise SystemExit(f"unknown workload {name}")


def one_by_one(ctx, scene, steps: int) -> dict:
    """The ensemble's instances run one after another, each as its own
    single-system simulation with the plain step (graph replay)."""
    from demso_amd import api, tuning

    total = 0.0
    for k in range(len(scene.instances)):
        sim = api.Simulation(ctx, tuning.instance_scene(scene, k))
        sim.step(1)
        total += tuning.time_graph_steps(sim, steps) * steps
        sim.close()
    return {"value": scene.n * steps / total, "unit": "particle-steps/s", "seconds": total,
            "instances": len(scene.instances)}


def cpu_baseline(scene, steps: int, threads: int) -> dict:
    """The oracle (FP32 port of the reference kernels) on this host."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    used = pyoracle.set_threads(threads)
    sim = pyoracle.Sim(scene, "f32")
    secs = sim.step(steps)
    sim.close()
    return {
        "value": scene.n * steps / secs,
        "unit": "particle-steps/s",
        "cores": used,
        "kind": "port",
        "sample": f"{steps} full steps of the same {scene.n}-disc workload from its initial state "
                  f"({secs:.1f} s, OpenMP over {used} threads)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="settling", choices=["settling", "column_collapse", "ensemble"])
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="finish with a plain device copy of this many MiB (reference for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in the docs)", file=sys.stderr)
        return 2

    import torch
    from demso_amd import api

    scene, workload = make_scene(args.workload)
    ctx = api.Context(0)
    cell_tuning = None
    if args.workload == "ensemble" and not args.no_autotune:
        # the ensemble workload tunes the cell-size multiplier as well
        from demso_amd import tuning

        cell_tuning = tuning.autotune(ctx, scene, multipliers=(1.0, 1.25, 1.5, 2.0), steps=10)
        scene = tuning.with_cell_multiplier(scene, cell_tuning["multiplier"])
    sim = api.Simulation(ctx, scene)
    n, cells = scene.n, scene.total_cells

    sim.step(1)
    tuned = None if args.no_autotune else sim.autotune()
    passes = sim.sort_passes()
    sim.step_graph(max(2, args.warmup))
    torch.cuda.synchronize()
    ctx.synchronize()

    # timed region: K steps by graph replay
    t0 = time.perf_counter()
    sim.step_graph(args.steps)
    ctx.synchronize()
    torch.cuda.synchronize()
    secs = time.perf_counter() - t0
    ms_per_step = 1e3 * secs / args.steps
    value = n * args.steps / secs

    # per-kernel durations with HIP events on the launch stream; the sort is
    # a multi-kernel stage (3 kernels per radix pass) and is reported apart
    stages = sim.stage_times(max(5, min(args.steps, 50)))
    kernel_stages = {k: v for k, v in stages.items() if k in ("hash", "cell_fill", "reorder", "collide", "integrate")}
    dom = max(kernel_stages, key=kernel_stages.get)
    dom_bytes = algorithmic_bytes(dom, n, cells)
    achieved = dom_bytes / (kernel_stages[dom] * 1e-3) / 1e9
    step_bytes = sum(algorithmic_bytes(k, n, cells) for k in BYTES_PER_PARTICLE)

    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if tfile.exists():
        t = json.loads(tfile.read_text())
        traffic = t.get(dom)

    cpu = None
    if args.cpu_steps > 0:
        cpu = cpu_baseline(scene, args.cpu_steps, args.cpu_threads)

    ensemble = None
    if args.workload == "ensemble":
        sim.close()
        single = one_by_one(ctx, scene, max(10, args.steps // 4))
        ensemble = {"one_by_one": single, "speedup_vs_one_by_one": value / single["value"],
                    "cell_tuning": cell_tuning}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "particles": n, "cells": cells, "sort_passes": passes,
                   "block_sizes": tuned},
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": dom_bytes,
        },
        "step_roofline": {
            "algorithmic_bytes_per_step": step_bytes,
            "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "stage_ms": stages,
        "cpu_baseline": cpu,
    }
    if ensemble is not None:
        out["ensemble"] = ensemble
    print(json.dumps(out), flush=True)
    sim.close()
    ctx.close()
    if args.calibration_copy_mb > 0:
        # a plain copy of known size whose source is not cache-resident: the
        # source is written, then 1 GiB of other data evicts the 256 MB MALL
        nbytes = args.calibration_copy_mb << 20
        src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
        evict = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del evict
    return 0


if __name__ == "__main__":
    sys.exit(main())
