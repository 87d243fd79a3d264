"""The committed column-collapse fixture (tests/golden/column_collapse_cfg0.npz,
written by tests/golden/make_golden.py from the FP32 oracle): the scene
generator and the oracle still reproduce it (CPU), and the HIP step
reproduces it bit for bit, eager and by graph replay (GPU)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

from helpers import bits_equal

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import make_golden  # noqa: E402


@pytest.fixture(scope="module")
def fixture():
    with np.load(GOLDEN / "column_collapse_cfg0.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_scene_generator_reproduces_fixture(fixture):
    s = make_golden.golden_scene()
    assert bits_equal(s.pos, fixture["init_pos"])
    assert bits_equal(s.radius, fixture["init_radius"])
    assert np.array_equal(s.type, fixture["init_type"])


def test_oracle_reproduces_fixture(fixture, oracle):
    got = make_golden.oracle_states(make_golden.golden_scene())
    assert set(got) <= set(fixture)
    for k, v in got.items():
        assert bits_equal(v, fixture[k]), k


def check_step(d, fixture, k):
    for f in make_golden.FIELDS:
        assert bits_equal(d[f], fixture[f"s{k}_{f}"]), (k, f)
    occ = d["cellStart"] != 0xFFFFFFFF
    assert np.array_equal(d["cellEnd"][occ], fixture[f"s{k}_cellEnd_occupied"])


@pytest.mark.gpu
@pytest.mark.parametrize("graph", [False, True])
def test_hip_step_reproduces_fixture(gpu_ctx, fixture, graph):
    from demso_amd import api

    s = make_golden.golden_scene()
    sim = api.Simulation(gpu_ctx, s)
    done = 0
    for k in make_golden.STEPS:
        (sim.step_graph if graph else sim.step)(k - done)
        done = k
        check_step(sim.download(), fixture, k)
    sim.close()
