"""FP32-vs-FP64 trajectory error of the DEM step (oracle runs).

TEST INFRASTRUCTURE: runs the oracle's FP32 and FP64 instantiations from
the same initial state and reports, every `--every` steps, the maximum
position error (m), the RMS position error and the maximum velocity error
of the FP32 run against the FP64 run.  The HIP path reproduces the FP32
oracle bit for bit (see tests/test_gpu_step.py), so these are also the
HIP path's errors.

  python oracle/trajectory_error.py --workload column_collapse --steps 10000 --every 1000

Ensemble workloads run sampled instances, each alone (instances are
independent, so an instance's trajectory in the ensemble is its trajectory
alone, which tests/test_gpu_ensemble.py checks bit for bit).  The
calibration column is first settled by the FP32 run (the adjoint workload's
starting packing), then both precisions continue from it.  --out appends
the final row of the run to a JSON file keyed by workload.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
sys.path.insert(0, str(HERE.parent))

import pyoracle  # noqa: E402
from demso_amd import scene as sc  # noqa: E402


def by_id(st):
    o = np.argsort(st["id"])
    return st["pos"][o].astype(np.float64), st["vel"][o].astype(np.float64)


def run_pair(s, steps, every, settle=0):
    """FP32 and FP64 oracle runs of scene s (after `settle` FP32 steps shared
    by both) with the error rows every `every` steps."""
    if settle:
        a = pyoracle.Sim(s, "f32")
        a.step(settle)
        st = a.state()
        a.close()
        s = s.copy()
        for k in ("pos", "vel", "theta", "omega", "id", "type", "loc"):
            setattr(s, k, np.ascontiguousarray(st[k]))
        if s.radius is not None:
            s.radius = np.ascontiguousarray(st["radius"])
    a, b = pyoracle.Sim(s, "f32"), pyoracle.Sim(s, "f64")
    rows = []
    done = 0
    while done < steps:
        k = min(every, steps - done)
        a.step(k)
        b.step(k)
        done += k
        pa, va = by_id(a.state())
        pb, vb = by_id(b.state())
        dp = np.linalg.norm(pa - pb, axis=1)
        rows.append({"step": done, "max_pos_err_m": float(dp.max()), "rms_pos_err_m": float(np.sqrt((dp ** 2).mean())),
                     "max_vel_err_mps": float(np.linalg.norm(va - vb, axis=1).max()),
                     "max_speed_mps": float(np.linalg.norm(vb, axis=1).max())})
        print(json.dumps(rows[-1]), flush=True)
    a.close()
    b.close()
    return rows


def sample_instances(ens, count):
    """`count` instances spread over the ensemble's range of sizes."""
    order = np.argsort(np.diff(ens.segments), kind="stable")
    pick = np.unique(np.round(np.linspace(0, len(order) - 1, count)).astype(int))
    return [int(order[i]) for i in pick]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="column_collapse",
                    choices=["column_collapse", "settling_small", "settling", "pile", "ensemble", "calibration",
                             "adjoint_column"])
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--instances", type=int, default=8, help="instances sampled from an ensemble workload")
    ap.add_argument("--out", default=None, help="JSON file that collects the final row per workload")
    args = ap.parse_args()
    if args.threads:
        pyoracle.set_threads(args.threads)
    t0 = time.time()
    settle = 0
    if args.workload in ("ensemble", "calibration"):
        from demso_amd import tuning

        ens = sc.ensemble() if args.workload == "ensemble" else sc.ensemble(seed=4, instances=64, nmin=16384,
                                                                             nmax=131072)
        per = []
        for k in sample_instances(ens, args.instances):
            sub = tuning.instance_scene(ens, k)
            print(f"instance {k}: {sub.n} discs", flush=True)
            rows = run_pair(sub, args.steps, args.every)
            per.append({"instance": k, "n": sub.n, **rows[-1]})
        final = {"step": args.steps, "instances": per,
                 "max_pos_err_m": max(r["max_pos_err_m"] for r in per),
                 "rms_pos_err_m": float(np.sqrt(sum(r["rms_pos_err_m"] ** 2 * r["n"] for r in per)
                                                / sum(r["n"] for r in per))),
                 "max_vel_err_mps": max(r["max_vel_err_mps"] for r in per),
                 "max_speed_mps": max(r["max_speed_mps"] for r in per)}
        n = ens.n
    else:
        if args.workload == "column_collapse":
            s = sc.column_collapse()
        elif args.workload == "settling_small":
            s = sc.settling(side=256)
        elif args.workload == "settling":
            s = sc.settling()
        elif args.workload == "adjoint_column":
            s = sc.calibration_column(seed=3, cols=256, rows=256)
            settle = 5000
        else:
            s = sc.triangle_pile()
        final = run_pair(s, args.steps, args.every, settle)[-1]
        n = s.n
    res = {"workload": args.workload, "n": n, "steps": args.steps, "settled_first": settle, "final": final,
           "seconds": time.time() - t0}
    print(json.dumps(res), flush=True)
    if args.out:
        path = Path(args.out)
        data = json.loads(path.read_text()) if path.exists() else {}
        data[args.workload] = res
        path.write_text(json.dumps(data, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
