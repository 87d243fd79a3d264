"""Writes tests/golden/column_collapse_cfg0.npz from the FP32 oracle.

The fixture is the oracle's output, so it pins nothing by itself; it records
the state the oracle (checked against the known-answer tests in
tests/test_oracle_kat.py) produced, so later changes to the oracle, the scene
generator or the HIP step that alter any bit are caught.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

FIELDS = ("hash", "index", "cellStart", "pos", "vel", "theta", "omega", "acc", "alpha", "pressure", "id", "loc")
STEPS = (1, 10)


def golden_scene():
    from demso_amd import scene as sc

    return sc.column_collapse(seed=0)


def oracle_states(scene) -> dict:
    import pyoracle

    sim = pyoracle.Sim(scene, "f32")
    out, done = {}, 0
    for k in STEPS:
        sim.step(k - done)
        done = k
        st = sim.state()
        for f in FIELDS:
            out[f"s{k}_{f}"] = st[f]
        occ = st["cellStart"] != 0xFFFFFFFF
        out[f"s{k}_cellEnd_occupied"] = st["cellEnd"][occ]
    sim.close()
    return out


def main() -> None:
    s = golden_scene()
    data = {"init_pos": s.pos, "init_radius": s.radius, "init_type": s.type}
    data.update(oracle_states(s))
    np.savez_compressed(HERE / "column_collapse_cfg0.npz", **data)
    print("wrote", HERE / "column_collapse_cfg0.npz")


if __name__ == "__main__":
    main()
