"""Batched ensemble and cell-multiplier tuning on the GPU.

The ensemble step must equal running each instance alone with the plain
single-system step: same particles in the same order, bit-identical state,
keys equal to the instance's keys offset by its cell base, permutation and
cellStart offset by its first particle."""
import numpy as np
import pytest

from demso_amd import api, tuning
from demso_amd import scene as sc
from helpers import bits_equal, rel_err

pytestmark = pytest.mark.gpu


def squeezed_ensemble():
    s = sc.ensemble(instances=4, counts=[900, 2500, 64, 1300], seed=9)
    for k in range(4):
        a, b = s.segments[k], s.segments[k + 1]
        s.pos[a:b] *= np.float32(0.86)
    rng = np.random.default_rng(5)
    s.vel[:] = (rng.normal(size=s.vel.shape) * 0.1).astype(np.float32)
    s.omega[:] = (rng.normal(size=s.n) * 10).astype(np.float32)
    return s


def check_equals_alone(gpu_ctx, s, steps):
    ens = api.Simulation(gpu_ctx, s)
    ens.step_graph(steps)
    E = ens.download()
    ens.close()
    for k in range(len(s.instances)):
        a, b = int(s.segments[k]), int(s.segments[k + 1])
        sub = tuning.instance_scene(s, k)
        one = api.Simulation(gpu_ctx, sub)
        one.step_graph(steps)
        O = one.download()
        one.close()
        g = s.instances[k]
        for key in ("pos", "vel", "theta", "omega", "radius", "acc", "alpha", "pressure"):
            assert bits_equal(E[key][a:b], O[key]), (k, key)
        assert np.array_equal(E["id"][a:b] - a, O["id"])
        assert np.array_equal(E["hash"][a:b] - g.cell_base, O["hash"])
        assert np.array_equal(E["index"][a:b] - a, O["index"])
        cs = E["cellStart"][g.cell_base:g.cell_base + g.cells]
        occ = O["cellStart"] != 0xFFFFFFFF
        assert np.array_equal(cs[occ] - a, O["cellStart"][occ])
        assert np.all(cs[~occ] == 0xFFFFFFFF)


def test_ensemble_equals_instances_run_alone(gpu_ctx):
    check_equals_alone(gpu_ctx, squeezed_ensemble(), 12)


def test_256_instances_equal_instances_run_alone(gpu_ctx):
    """256 instances (the ensemble workload's instance count and parameter
    draws, with 50-1,500 discs each): one batched step per stage equals the
    256 plain single-system steps, instance by instance."""
    counts = np.random.default_rng(12).integers(50, 1500, size=256).tolist()
    s = sc.ensemble(instances=256, counts=counts, seed=2)
    assert len(s.instances) == 256
    print(f"256 instances, {s.n} discs", flush=True)
    check_equals_alone(gpu_ctx, s, 6)


@pytest.mark.parametrize("mult", [1.0, 1.5, 2.0])
def test_cell_multiplier_matches_oracle(gpu_ctx, oracle, mult):
    s = tuning.with_cell_multiplier(sc.column_collapse(cols=24, rows=24, pitch=0.0098, jitter=0.0012), mult)
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    sim.step(10)
    ref.step(10)
    g, o = sim.download(), ref.state()
    assert np.array_equal(g["hash"], o["hash"]) and np.array_equal(g["index"], o["index"])
    for k in ("pos", "vel", "acc", "pressure"):
        assert rel_err(g[k], o[k]) <= 1e-5
        assert bits_equal(g[k], o[k])
    sim.close()


def test_autotune_picks_a_candidate(gpu_ctx):
    s = sc.column_collapse()
    res = tuning.autotune(gpu_ctx, s, multipliers=(1.0, 1.5), steps=4)
    assert res["multiplier"] in (1.0, 1.5)
    assert len(res["trials"]) == 2 and all(t["seconds_per_step"] > 0 for t in res["trials"])
