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
  dense_settling   (default run) the same discs at their mean diameter,
                   about two contacts a disc: particle-steps/s
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "particle-steps/s @4M discs; batched-ensemble and fwd+adjoint vs plain step"

# Algorithmic (compulsory) HBM bytes per particle and per cell for each
# stage of the owned simulation's re-sorted step, FP32 state and 32-bit keys.
# `radius`: a per-particle radius array (4 B where noted); `typed`: several
# material rows, so the kernels read each particle's type (4 B where noted).
#
# In place (the default split layout; the state stays in its buffer):
#   sort      the per-tile mover counts (4 B per 4,096 particles) and the
#             mover words of the tiles that have movers, plus 48 B per
#             particle whose slot changes (a handful per step on the settling
#             workload): not counted
#   reorder   only the particles whose slot changes (not counted)
#   collide   R pos 8, radius 4, type 4, vel 8, omega 4; W acc 8, alpha 4,
#             pressure 4; + R cellStart/cellEnd 8 per cell
#   integrate R pos 8, vel 8, theta 4, omega 4, acc 8, alpha 4, radius 4,
#             type 4; W pos 8, vel 8, theta 4, omega 4 (the cell of the
#             position read is recomputed, not loaded; the next key is
#             written only for the few particles that change cell)
#   (the reference algorithm's figures; in runs of 64 slots where no force
#   acted, collide writes no acc/alpha and integrate reads none, so the
#   bytes moved can be below these)
# Copy (DEM_STEP_COPY): re-sort R previous key 4, W sorted key 4, index 4 and
#   the next cellStart 4 per cell; reorder R key 4, index 4, pos 8, id 4,
#   type 4, radius 4, W pos 8, id 4, loc 4, type 4, radius 4; collide and
#   integrate also read the sorted index (4 each) and velocity, spin and
#   angle through it.
# Fused (collide + integrate in one kernel): reorder sorts every field (R key
#   4, index 4, pos 8, vel 8, theta 4, omega 4, id 4, type 4, radius 4; W the
#   same state and loc 4) and the fused kernel reads pos 8, vel 8, omega 4,
#   theta 4, radius 4, type 4 and writes acc 8, alpha 4, pressure 4, pos 8,
#   vel 8, theta 4, omega 4 (+ 8 per cell), reported under collide.
STAGES = ("hash", "sort", "cell_fill", "reorder", "collide", "integrate")


def stage_bytes(layout: str, radius: bool, typed: bool) -> tuple[dict, dict]:
    """Per-particle and per-cell algorithmic bytes of each stage."""
    r = 4 if radius else 0
    t = 4 if typed else 0
    per_cell = {"hash": 0, "sort": 0, "cell_fill": 0, "reorder": 0, "collide": 8, "integrate": 0}
    if layout == "fused":
        pp = {"hash": 0, "sort": 12, "cell_fill": 0, "reorder": 76 + 2 * r, "collide": 64 + r + t, "integrate": 0}
        per_cell["sort"] = 4

    elif layout == "copy":
        pp = {"hash": 0, "sort": 12, "cell_fill": 0, "reorder": 44 + 2 * r,
              "collide": 4 + 8 + r + t + 12 + 16, "integrate": 4 + 8 + 16 + 12 + r + t + 24}
        per_cell["sort"] = 4
    else:
        pp = {"hash": 0, "sort": 0, "cell_fill": 0, "reorder": 0,
              "collide": 8 + r + t + 12 + 16, "integrate": 8 + 16 + 12 + r + t + 24}
    return pp, per_cell


def algorithmic_bytes(stage: str, n: int, cells: int, layout: str = "inplace", radius: bool = True,
                      typed: bool = False) -> int:
    pp, pc = stage_bytes(layout, radius, typed)
    return int(pp[stage] * n + pc[stage] * cells)


def make_scene(name: str):
    from demso_amd import scene as sc

    if name == "settling":
        return sc.settling(), "settling: 4,194,304 log-normal discs (r0=5 mm), 4731x2366 cells, FP32"
    if name == "settling_dense":
        s = sc.settling_dense()
        return s, (f"settling_dense: the 4,194,304 settling discs on a lattice at their mean diameter "
                   f"(about 2 contacts a disc), {s.instances[0].grid[0]}x{s.instances[0].grid[1]} cells, FP32")
    if name == "column_collapse":
        return sc.column_collapse(), "column collapse: 4,096 discs, 91x46 cells, FP32"
    if name == "ensemble":
        s = sc.ensemble()
        return s, f"ensemble: 256 instances, {s.n} discs, {s.total_cells} cells, FP32"
    if name == "adjoint":
        s = sc.calibration_column(seed=3, cols=256, rows=256)
        return s, "adjoint: 65,536-disc column (256x256 lattice, settled 5000 steps), 2000 steps, checkpoints every 100"
    if name == "calibration":
        s = sc.ensemble(seed=4, instances=64, nmin=16384, nmax=131072)
        return s, f"calibration: 64 instances, {s.n} discs, 1000 steps fwd+adjoint per Adam iteration"
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


def adjoint_cost(ctx, scene, steps: int, K: int, settle_steps: int) -> dict:
    """Plain forward, forward with checkpoints and backward of one run of
    `steps` steps from a settled packing (wall times, after an untimed
    warm-up of each); loss = mean squared final position error against a
    target run with kn = 2e5, mu = 0.35, e = 0.6."""
    from demso_amd import api, calibrate

    s = calibrate.settle(ctx, scene, settle_steps)
    tgt = calibrate.set_params(s, [2e5], [0.35], [0.6])
    run = calibrate.Runner(ctx, s, steps, checkpoint_every=K)
    target = run.final_positions(tgt).astype("float64")
    run.run(s, target)  # warm-up
    res = run.run(s, target)
    sim = run.sim
    ctx.set_contact_params(calibrate.contact_rows(s))
    sim.upload(s)
    sim.step_graph(steps)  # warm-up
    sim.upload(s)
    ctx.synchronize()
    t0 = time.perf_counter()
    sim.step_graph(steps)
    ctx.synchronize()
    plain = time.perf_counter() - t0
    run.close()
    return {
        "steps": steps, "checkpoint_every": K, "particles": s.n,
        "plain_forward_s": plain, "forward_with_checkpoints_s": res.seconds_forward,
        "backward_s": res.seconds_backward,
        "backward_over_forward": res.seconds_backward / plain,
        "forward_plus_backward_over_forward": (res.seconds_forward + res.seconds_backward) / plain,
        "loss": res.loss, "grad_params": res.grad_params.tolist(),
    }


def ensemble_speedup(ctx, steps: int = 50) -> dict:
    """The batched-ensemble workload (256 column-collapse instances, 8.2M
    discs, its own grids) in one launch per stage, against the same
    instances run one after another with the plain single-system step."""
    from demso_amd import api, tuning
    from demso_amd import scene as sc

    scene = sc.ensemble()
    sim = api.Simulation(ctx, scene)
    sim.step(1)
    sim.autotune()
    sec = tuning.time_graph_steps(sim, steps)
    sim.close()
    single = one_by_one(ctx, scene, max(10, steps // 2))
    batched = scene.n / sec
    return {"value": batched / single["value"], "batched": batched, "one_by_one": single["value"],
            "unit": "particle-steps/s", "instances": len(scene.instances), "particles": scene.n,
            "steps": steps}


def dense_settling(ctx, steps: int = 20) -> dict:
    """The headline's 4,194,304 discs on a lattice at their mean diameter
    (about two contacts a disc, so the contact law runs for most discs):
    graph-replayed steps after the first (radix-sorted) one and autotuning,
    next to the sparse headline state."""
    from demso_amd import api, tuning
    from demso_amd import scene as sc

    scene = sc.settling_dense()
    sim = api.Simulation(ctx, scene)
    sim.step(1)
    sim.autotune()
    sec = tuning.time_graph_steps(sim, steps)
    sim.close()
    return {"value": scene.n / sec, "unit": "particle-steps/s", "ms_per_step": sec * 1e3, "steps": steps,
            "particles": scene.n, "workload": "settling_dense (bench.py --workload settling_dense)"}


def adjoint_ratio(ctx) -> dict:
    """Backward (hand-written adjoint with recomputation from checkpoints)
    over the plain forward of the 65,536-disc calibration column, 2,000
    steps, checkpoints every 100."""
    from demso_amd import scene as sc

    r = adjoint_cost(ctx, sc.calibration_column(seed=3, cols=256, rows=256), 2000, 100, 5000)
    return {"value": r["backward_over_forward"],
            "forward_plus_backward_over_forward": r["forward_plus_backward_over_forward"],
            "plain_forward_s": r["plain_forward_s"], "backward_s": r["backward_s"], "particles": r["particles"],
            "steps": r["steps"], "checkpoint_every": r["checkpoint_every"]}


def calibration_run(ctx, scene, steps: int, iters: int, K: int, settle_steps: int) -> dict:
    """Batched calibration: every instance starts from (kn, mu, e) =
    (1e5, 0.5, 0.5) and Adam recovers its own drawn parameters from the
    final positions of a target run with them (all instances in one launch
    per stage), from the lattice after `settle_steps` plain steps (0: the
    lattice as generated).  Reports the error of every instance's recovered
    parameters, and the forward + adjoint particle-steps/s against the same
    instances' plain forward run one by one."""
    import statistics

    from demso_amd import calibrate

    s = calibrate.settle(ctx, scene, settle_steps) if settle_steps else scene.copy()
    T = len(s.materials)
    truth = {"kn": [m.kn for m in s.materials], "mu": [m.mu for m in s.materials],
             "e": [float(v) for v in s.notes["restitution"]]}
    tgt = s.copy()
    start = calibrate.set_params(s, [1e5] * T, [0.5] * T, [0.5] * T)
    t0 = time.perf_counter()
    out = calibrate.calibrate(ctx, start, tgt, steps=steps, iters=iters, lr=0.1, checkpoint_every=K)
    secs = time.perf_counter() - t0
    hist = out["history"]

    def errors(kn, mu, e):
        return {"kn_rel": [abs(a / b - 1) for a, b in zip(kn, truth["kn"])],
                "mu_abs": [abs(a - b) for a, b in zip(mu, truth["mu"])],
                "e_abs": [abs(a - b) for a, b in zip(e, truth["e"])]}

    def summary(err):
        return {k: {"median": statistics.median(v), "p90": sorted(v)[int(0.9 * (len(v) - 1))], "max": max(v)}
                for k, v in err.items()}

    start_err = errors([1e5] * T, [0.5] * T, [0.5] * T)
    final_err = errors(list(out["kn"]), list(out["mu"]), list(out["e"]))
    t_fwd = sum(h["t_fwd"] for h in hist) / iters
    t_bwd = sum(h["t_bwd"] for h in hist) / iters
    plain = one_by_one(ctx, s, min(steps, 200))
    fwd_adj = s.n * steps / (t_fwd + t_bwd)
    return {
        "instances": T, "particles": s.n, "steps": steps, "iterations": iters, "checkpoint_every": K,
        "settled_first": settle_steps,
        "seconds": secs, "seconds_per_iteration": secs / iters,
        "forward_s_mean": t_fwd, "backward_s_mean": t_bwd, "backward_over_forward": t_bwd / t_fwd,
        "loss_first": hist[0]["loss"], "loss_last": hist[-1]["loss"],
        "error_start": summary(start_err), "error_final": summary(final_err),
        "per_instance": [{"instance": k, "n": int(s.segments[k + 1] - s.segments[k]),
                          "kn": [truth["kn"][k], float(out["kn"][k])], "mu": [truth["mu"][k], float(out["mu"][k])],
                          "e": [truth["e"][k], float(out["e"][k])]} for k in range(T)],
        "forward_adjoint_particle_steps_per_s": fwd_adj,
        "plain_forward_one_by_one_particle_steps_per_s": plain["value"],
        "forward_adjoint_vs_plain_one_by_one": fwd_adj / plain["value"],
        "kn_rel_err_median": statistics.median(final_err["kn_rel"]),
        "mu_abs_err_median": statistics.median(final_err["mu_abs"]),
    }


def _oracle_sample(pyoracle, scene, steps: int, seconds: float):
    """Steps of the oracle from the scene's initial state: `steps` to gauge
    the rate, then as many more as make about `seconds` in all."""
    sim = pyoracle.Sim(scene, "f32")
    secs = sim.step(steps)
    for _ in range(4):  # top up: the first steps run slower than the rest
        if secs >= 0.9 * seconds:
            break
        more = max(1, int(math.ceil((seconds - secs) * steps / secs)))
        secs += sim.step(more)
        steps += more
    sim.close()
    return steps, secs


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, steps: int, threads: int, seconds: float) -> dict:
    """The oracle (FP32 port of the reference kernels) on this host, over
    `threads` OpenMP threads for about `seconds` of work, and on one thread
    for about half that."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    used = pyoracle.set_threads(threads)
    steps_m, secs_m = _oracle_sample(pyoracle, scene, steps, seconds)
    pyoracle.set_threads(1)
    steps_1, secs_1 = _oracle_sample(pyoracle, scene, 1, seconds / 2)
    pyoracle.set_threads(used)
    return {
        "value": scene.n * steps_m / secs_m,
        "unit": "particle-steps/s",
        "cores": used,
        "kind": "port",
        "sample": f"{steps_m} full steps of the same {scene.n}-disc workload from its initial state "
                  f"({secs_m:.1f} s, OpenMP over {used} threads)",
        "single_thread": {"value": scene.n * steps_1 / secs_1, "cores": 1,
                          "sample": f"{steps_1} steps of the same workload ({secs_1:.1f} s)"},
        "cpu_model": _cpu_model(),
    }


def stage_loop(ctx, scene, steps: int, warmup: int) -> dict:
    """The reference's own step loop (SimLooping, src/main.cu:431-507)
    through the drop-in stage calls: calcHash, sortParticles (radix sort
    every step), reorderDataAndFindCellStart (full reorder into the other
    buffer set), collide and integrateSystem, then the buffer swap; eager
    launches on the context stream, as the reference issues them."""
    from demso_amd import api

    n, cells = scene.n, scene.total_cells
    v = api.allocateVectors(scene, ctx)
    v["s1"].upload(scene)
    a, b = v["s1"], v["s2"]
    h, idx = v["gridParticleHash"], v["gridParticleIndex"]
    cs, ce = v["cellStart"], v["cellEnd"]

    def step(a, b):
        api.calcHash(ctx, a.pos, idx, h, n)
        api.sortParticles(ctx, h, idx, n)
        api.reorderDataAndFindCellStart(ctx, cs, ce, b.pos, b.vel, b.theta, b.omega, b.id, b.loc, b.type, h, idx,
                                        a.pos, a.vel, a.theta, a.omega, a.id, a.type, n, cells,
                                        sortedRadius=b.radius, oldRadius=a.radius)
        api.collide(ctx, b.pos, b.vel, v["acc"], b.omega, v["alpha"], b.type, cs, ce, n, cells, v["pressure"],
                    radius=b.radius)
        api.integrateSystem(ctx, b.pos, b.vel, v["acc"], b.theta, b.omega, v["alpha"], b.type, v["pressure"], n,
                            radius=b.radius)
        return b, a

    for _ in range(warmup):
        a, b = step(a, b)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        a, b = step(a, b)
    ctx.synchronize()
    secs = time.perf_counter() - t0
    api.desAllocateVectors(v)
    return {"value": n * steps / secs, "unit": "particle-steps/s", "ms_per_step": 1e3 * secs / steps,
            "steps": steps, "sort_passes": (max(cells - 1, 1).bit_length() + 7) // 8,
            "note": "the reference's stage loop (hash, radix sort, full reorder, collide, integrate, swap) "
                    "through the C ABI, eager launches"}


def fixed_step_bytes(n: int, cells: int) -> int:
    """The fixed algorithmic figure of one full step, 212 B per particle and
    16 B per cell (hash 16, sort 16, reorder 76, collide 40, integrate 64 per
    particle; cellStart/cellEnd written once and read once per cell), FP32
    state and 32-bit keys, whatever the step does to avoid that traffic."""
    return 212 * n + 16 * cells


def cells_collide_reads(scene, sorted_keys) -> int:
    """Cells whose cellStart/cellEnd collide reads: the 3x3 neighbourhoods of
    the occupied cells, inside each instance's grid (clamp mode)."""
    import numpy as np

    keys = np.unique(np.asarray(sorted_keys, np.int64))
    base = np.array([g.cell_base for g in scene.instances], np.int64)
    order = np.argsort(base)
    k = order[np.searchsorted(base[order], keys, side="right") - 1]
    G0 = np.array([g.grid[0] for g in scene.instances], np.int64)[k]
    G1 = np.array([g.grid[1] for g in scene.instances], np.int64)[k]
    b = base[k]
    local = keys - b
    cx, cy = local % G0, local // G0
    out = []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            x, y = cx + dx, cy + dy
            ok = (x >= 0) & (x < G0) & (y >= 0) & (y < G1)
            out.append((b + y * G0 + x)[ok])
    return int(np.unique(np.concatenate(out)).size)


def full_run(sim, scene, steps: int, repeats: int) -> dict:
    """The workload's whole run of `steps` steps from its initial state (the
    first step radix-sorted, the rest re-sorted graph replays), timed
    `repeats` times after an untimed run; median and all values."""
    import statistics

    vals = []
    for r in range(repeats + 1):
        sim.upload(scene)
        sim.synchronize()
        t0 = time.perf_counter()
        sim.step_graph(steps)
        sim.synchronize()
        if r:
            vals.append(time.perf_counter() - t0)
    med = statistics.median(vals)
    return {"steps": steps, "repeats": repeats, "seconds_median": med, "seconds": vals,
            "value_median": scene.n * steps / med, "unit": "particle-steps/s",
            "from": "the initial state: first step radix-sorted, then re-sorted steps by graph replay"}


def trajectory_error(workload: str):
    """FP32-vs-FP64 trajectory error of the workload's run, from the oracle
    runs in profiles/trajectory_error.json (made by oracle/trajectory_error.py
    on the build host; the GPU's FP32 state equals the FP32 oracle's bit for
    bit in the parity tests, so these are the GPU run's errors)."""
    path = ROOT / "profiles" / "trajectory_error.json"
    name = {"settling": "settling", "column_collapse": "column_collapse", "ensemble": "ensemble",
            "adjoint": "adjoint_column", "calibration": "calibration"}.get(workload)
    if not path.exists() or name is None:
        return None
    rec = json.loads(path.read_text()).get(name)
    if rec is None:
        return None
    f = rec["final"]
    return {"steps": rec["steps"], "max_pos_err_m": f["max_pos_err_m"], "rms_pos_err_m": f["rms_pos_err_m"],
            "max_vel_err_mps": f["max_vel_err_mps"], "max_speed_mps": f["max_speed_mps"],
            "source": "profiles/trajectory_error.json (oracle FP32 vs FP64 runs of the same workload)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="settling", choices=["settling", "settling_dense", "column_collapse", "ensemble", "adjoint",
                             "calibration"])
    ap.add_argument("--adj-steps", type=int, default=0, help="steps of the adjoint / calibration run (0 = workload default)")
    ap.add_argument("--iters", type=int, default=5, help="Adam iterations of the calibration workload")
    ap.add_argument("--settle", type=int, default=2000,
                    help="calibration workload: plain steps that settle the lattice first (0: start as generated)")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps that gauge the CPU baseline's rate (0 = skip it)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU work of the whole CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0 = the CPUs this process may run on, at most 16)")
    ap.add_argument("--full-repeats", type=int, default=5,
                    help="timed repeats of the workload's whole run from its initial state (0 = skip)")
    ap.add_argument("--no-extras", action="store_true",
                    help="settling workload: skip the ensemble and adjoint comparisons of the default line")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--collide-mode", choices=["plain", "prefetch", "batch", "list"], default=None,
                    help="with --no-autotune: collide's candidate mode (default: one candidate at a time)")
    ap.add_argument("--no-one-by-one", action="store_true",
                    help="ensemble workload: skip timing the instances one by one (for profiler runs)")
    ap.add_argument("--calibration-copy-mb", type=int, default=0,
                    help="finish with a plain device copy of this many MiB (reference for HBM counters)")
    ap.add_argument("--drop-in-only", action="store_true",
                    help="time only the reference's stage loop through the drop-in calls (--steps steps after "
                         "--warmup) and print its line (for profiler runs of the radix sort and full reorder)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in the docs)", file=sys.stderr)
        return 2

    import torch
    from demso_amd import api

    scene, workload = make_scene(args.workload)
    ctx = api.Context(0)
    if args.collide_mode:
        ctx.set_step_variant(prefetch=args.collide_mode == "prefetch", batch=args.collide_mode == "batch",
                             contact_lists=args.collide_mode == "list")
    if args.drop_in_only:
        line = stage_loop(ctx, scene, args.steps, args.warmup)
        line.update(workload=workload, particles=scene.n, cells=scene.total_cells)
        print(json.dumps(line), flush=True)
        ctx.close()
        calibration_copy(args.calibration_copy_mb)
        return 0
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
    layout = "fused" if tuned and tuned.get("fused") else ("copy" if tuned and tuned.get("copy") else "inplace")
    acct = dict(layout=layout, radius=scene.radius is not None, typed=len(scene.materials) > 1)
    kernel_stages = {k: v for k, v in stages.items()
                     if k in ("cell_fill", "reorder", "collide", "integrate") and algorithmic_bytes(k, n, cells, **acct) > 0}
    dom = max(kernel_stages, key=kernel_stages.get)
    dom_bytes = algorithmic_bytes(dom, n, cells, **acct)
    achieved = dom_bytes / (kernel_stages[dom] * 1e-3) / 1e9
    step_bytes = sum(algorithmic_bytes(k, n, cells, **acct) for k in STAGES)

    full = None
    if args.full_repeats > 0:
        full = full_run(sim, scene, scene.steps if args.workload in ("settling", "settling_dense", "column_collapse") else 1000,
                        args.full_repeats)
        sim.step(1)
    fixed = fixed_step_bytes(n, cells)
    touched = None
    if "collide" in kernel_stages and layout != "fused":
        used = cells_collide_reads(scene, sim.download()["hash"])
        pp, _ = stage_bytes(layout, acct["radius"], acct["typed"])
        tb = pp["collide"] * n + 8 * used
        ach = tb / (stages["collide"] * 1e-3) / 1e9
        touched = {"cells_read": used, "bytes": tb, "achieved_GBs": ach, "frac": ach / HBM_PEAK_GBS,
                   "note": "collide's algorithmic bytes counting cellStart/cellEnd only at the cells it reads "
                           "(the 3x3 neighbourhoods of occupied cells) instead of every cell of the grid"}

    traffic = None
    tfile = ROOT / "profiles" / f"traffic_{args.workload}.json"
    if tfile.exists():
        t = json.loads(tfile.read_text())
        # per candidate mode where the file has one (collide's traffic differs a little by mode)
        mode = (args.collide_mode or "plain") if not tuned else next(
            (m for k, m in (("prefetch", "prefetch"), ("batch", "batch"), ("contact_lists", "list")) if tuned.get(k)), "plain")
        traffic = t.get(f"{dom}_{mode}", t.get(dom))

    cpu = None
    if args.cpu_steps > 0:
        share = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = args.cpu_threads or min(16, share)
        cpu = cpu_baseline(scene, args.cpu_steps, threads, args.cpu_seconds)
        cpu["host_nproc"] = os.cpu_count()
        cpu["cpus_available"] = share
        cpu["threads_note"] = ("OpenMP over the CPUs this process may use, at most 16: a one-GPU job "
                               "is given a 16-CPU share of the host")

    extra = {}
    if args.workload == "adjoint":
        sim.close()
        extra["adjoint"] = adjoint_cost(ctx, scene, args.adj_steps or 2000, 100, 5000)
    if args.workload == "calibration":
        sim.close()
        extra["calibration"] = calibration_run(ctx, scene, args.adj_steps or 1000, args.iters, 100, args.settle)

    if args.workload in ("settling", "settling_dense", "column_collapse") and not args.no_extras:
        sim.close()
        extra["drop_in_stage_loop"] = stage_loop(ctx, scene, max(10, min(args.steps, 50)), 3)
    if args.workload == "settling" and not args.no_extras:
        sim.close()
        extra["dense_settling"] = dense_settling(ctx)
        extra["ensemble_speedup"] = ensemble_speedup(ctx)
        extra["adjoint_backward_over_forward"] = adjoint_ratio(ctx)

    ensemble = None
    if args.workload == "ensemble" and not args.no_one_by_one:
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
            "layout": layout,
            "algorithmic_bytes_per_step": step_bytes,
            "achieved_GBs": step_bytes / (ms_per_step * 1e-3) / 1e9,
            "frac": step_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "step_roofline_fixed": {
            "bytes_per_step": fixed,
            "formula": "212 N + 16 C",
            "achieved_GBs": fixed / (ms_per_step * 1e-3) / 1e9,
            "frac": fixed / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "note": "charges a full sort and reorder of every particle each step; the step re-sorts in place "
                    "and moves only the particles whose slot changes, so on a state where few discs change "
                    "cell this figure can exceed the HBM peak",
        },
        "collide_touched": touched,
        "full_run": full,
        "traj_error_fp64": trajectory_error(args.workload),
        "stage_ms": stages,
        "cpu_baseline": cpu,
    }
    if ensemble is not None:
        out["ensemble"] = ensemble
    out.update(extra)
    print(json.dumps(out), flush=True)
    sim.close()
    ctx.close()
    calibration_copy(args.calibration_copy_mb)
    return 0


def calibration_copy(mb: int) -> None:
    """A plain copy of known size whose source is not cache-resident (the
    source is written, then 1 GiB of other data evicts the 256 MB MALL): the
    reference for the HBM byte counters."""
    if mb <= 0:
        return
    import torch

    nbytes = mb << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    evict = torch.ones((1 << 30) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del evict


if __name__ == "__main__":
    sys.exit(main())
