"""The ensemble workloads at their full sizes.

* 256 column-collapse instances of 2,000-60,000 discs (8.2M discs in all,
  seed 2): the segmented step against the oracle's step loop over the whole
  ensemble, bit for bit after steps 1, 2 and 10 (the last 8 by graph replay
  of re-sorted steps), with every key inside its instance's cell range.
* 64 calibration instances of 16,384-131,072 discs (4.7M discs, seed 4):
  forward + adjoint over 10 steps; the gradients do not depend on the
  checkpoint interval, and sampled instances (the smallest, the median and
  the largest) equal the FP64 host adjoint of the instance run alone.
"""
import numpy as np
import pytest

from demso_amd import api, tuning
from demso_amd import scene as sc
from test_gpu_adjoint import gpu_grads, host_grads, normwise
from test_gpu_step import compare

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def test_ensemble_256_full_size_matches_oracle(gpu_ctx, oracle):
    s = sc.ensemble()
    assert len(s.instances) == 256 and 7_500_000 < s.n < 9_000_000
    base = np.array([g.cell_base for g in s.instances], np.int64)
    cells = np.array([g.cells for g in s.instances], np.int64)
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    for chunk, graph in ((1, False), (1, False), (8, True)):
        print(f"ensemble 256: {chunk} step(s), graph={graph}", flush=True)
        sim.step_graph(chunk) if graph else sim.step(chunk)
        ref.step(chunk)
        g, o = sim.download(), ref.state()
        compare(g, o)
        h = g["hash"].astype(np.int64)
        t = g["type"].astype(np.int64)
        assert np.all(h[1:] >= h[:-1])
        assert np.all(h >= base[t]) and np.all(h < base[t] + cells[t])
        assert np.array_equal(g["loc"][g["id"]], np.arange(s.n))
    assert np.count_nonzero(g["pressure"]) > s.n // 1000     # contacts are live
    ref.close()
    sim.close()


def test_calibration_ensemble_full_size_adjoint(gpu_ctx, oracle):
    s = sc.ensemble(seed=4, instances=64, nmin=16384, nmax=131072)
    assert len(s.instances) == 64 and 4_000_000 < s.n < 5_500_000
    rng = np.random.default_rng(4)
    W = rng.normal(size=(s.n, 2))
    V = np.zeros((s.n, 2))
    steps = 10
    print("calibration ensemble: adjoint with checkpoints every 5 steps", flush=True)
    (gp0, gv0, gpar), final = gpu_grads(gpu_ctx, s, steps, W, V, K=5)
    print("calibration ensemble: adjoint with one segment", flush=True)
    (gp0b, gv0b, gparb), finalb = gpu_grads(gpu_ctx, s, steps, W, V, K=steps)
    assert np.array_equal(final, finalb)
    assert np.array_equal(gp0, gp0b) and np.array_equal(gv0, gv0b)
    assert np.allclose(gpar, gparb, rtol=1e-12, atol=0)       # FP64 atomics: order may differ
    counts = np.diff(s.segments)
    for k in sorted({int(np.argmin(counts)), int(np.argsort(counts)[len(counts) // 2]), int(np.argmax(counts))}):
        a, b = int(s.segments[k]), int(s.segments[k + 1])
        print(f"calibration ensemble: instance {k} ({b - a} discs) vs the FP64 host adjoint", flush=True)
        sub = tuning.instance_scene(s, k)
        (hp0, hv0, hpar), hfinal = host_grads(oracle, sub, steps, W[a:b], V[a:b])
        # same forward: FP32 and FP64 positions agree to a few FP32 roundings
        # of the largest coordinate (instances are up to 4 m tall)
        assert np.abs(final[a:b] - hfinal).max() <= 1e-6 * np.abs(hfinal).max()
        assert normwise(gp0[a:b], hp0) < 1e-3
        assert normwise(gv0[a:b], hv0) < 1e-3
        # parameter gradients as sensitivities to relative changes, p * dL/dp
        # (kn, kt, cn and mu have different units): normwise within 1e-3, and
        # each component that carries at least 1e-3 of the largest within
        # 1e-3 (a component a million times smaller, such as kt's over 10
        # steps of a lattice with few shearing contacts, is left to the norm)
        m = s.materials[k]
        p = np.array([m.kn, m.kt, m.cn, m.mu], np.float64)
        sg, sh = gpar[k] * p, hpar[0] * p
        assert np.linalg.norm(sg - sh) <= 1e-3 * np.linalg.norm(sh), (k, sg, sh)
        big = np.abs(sh) >= 1e-3 * np.abs(sh).max()
        assert np.allclose(sg[big], sh[big], rtol=1e-3, atol=0), (k, sg, sh)
        assert np.all(np.abs(gpar[k][sh == 0.0]) <= 1e-12)
