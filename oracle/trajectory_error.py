"""FP32-vs-FP64 trajectory error of the DEM step (oracle runs).

TEST INFRASTRUCTURE: runs the oracle's FP32 and FP64 instantiations from
the same initial state and reports, every `--every` steps, the maximum
position error (m), the RMS position error and the maximum velocity error
of the FP32 run against the FP64 run.  The HIP path reproduces the FP32
oracle bit for bit (see tests/test_gpu_step.py), so these are also the
HIP path's errors.

  python oracle/trajectory_error.py --workload column_collapse --steps 10000 --every 1000
"""
from __future__ import annotations

import argparse
import json
import sys
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


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="column_collapse",
                    choices=["column_collapse", "settling_small", "settling", "pile"])
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--every", type=int, default=1000)
    ap.add_argument("--threads", type=int, default=0)
    args = ap.parse_args()
    if args.threads:
        pyoracle.set_threads(args.threads)
    if args.workload == "column_collapse":
        s = sc.column_collapse()
    elif args.workload == "settling_small":
        s = sc.settling(side=256)
    elif args.workload == "settling":
        s = sc.settling()
    else:
        s = sc.triangle_pile()
    a, b = pyoracle.Sim(s, "f32"), pyoracle.Sim(s, "f64")
    rows = []
    done = 0
    while done < args.steps:
        k = min(args.every, args.steps - done)
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
    print(json.dumps({"workload": args.workload, "n": s.n, "steps": args.steps, "final": rows[-1]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
