"""Whole-step parity: the owned HIP simulation (eager launches and hipGraph
replay) against the oracle's step loop (src/main.cu:396-507) over many
steps, bit for bit on every integer output and within 1e-5 relative on the
state; plus the 4M-disc settling workload at full size."""
import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from helpers import mixed_pack, bits_equal, compressed_pack, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-5


def compare(g, o, tol=TOL, exact=True):
    assert np.array_equal(g["hash"], o["hash"])
    assert np.array_equal(g["index"], o["index"])
    assert np.array_equal(g["cellStart"], o["cellStart"])
    occ = o["cellStart"] != 0xFFFFFFFF
    assert np.array_equal(g["cellEnd"][occ], o["cellEnd"][occ])
    for k in ("id", "type", "loc"):
        assert np.array_equal(g[k], o[k]), k
    for k in ("pos", "vel", "theta", "omega", "acc", "alpha", "pressure"):
        if np.abs(o[k]).max() > 0:
            assert rel_err(g[k], o[k]) <= tol, k
        if exact:
            assert bits_equal(g[k], o[k]), k


CASES = {
    "column_collapse_4096": (lambda: sc.column_collapse(), 40),
    "pack_1024": (lambda: compressed_pack(), 25),
    "mixed_pack_1024": (lambda: mixed_pack(), 25),
    "pile_wrap_attract": (lambda: sc.triangle_pile(), 60),
    "ensemble_5": (lambda: sc.ensemble(instances=5, counts=[1500, 800, 2100, 90, 640], seed=4), 30),
}


@pytest.mark.parametrize("name", list(CASES))
def test_steps_match_oracle(gpu_ctx, oracle, name):
    make, steps = CASES[name]
    s = make()
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    for chunk in (1, steps - 1):
        sim.step(chunk)
        ref.step(chunk)
        compare(sim.download(), ref.state())
    sim.close()


def test_graph_replay_matches_eager(gpu_ctx):
    s = compressed_pack(seed=11)
    eager = api.Simulation(gpu_ctx, s)
    graph = api.Simulation(gpu_ctx, s)
    for k in range(4):  # replays with eager work and syncs in between
        eager.step(6)
        graph.step_graph(6)
        a, b = eager.download(), graph.download()
        for key in ("pos", "vel", "theta", "omega", "acc", "alpha", "pressure", "hash", "index", "cellStart"):
            assert bits_equal(a[key], b[key]), (k, key)
    graph.step_graph(5)  # odd count: one eager step + replays
    eager.step(5)
    assert bits_equal(eager.download()["pos"], graph.download()["pos"])
    eager.close()
    graph.close()


def test_autotune_keeps_state(gpu_ctx, oracle):
    s = compressed_pack(seed=3)
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    sim.step(3)
    ref.step(3)
    choice = sim.autotune()
    fused = choice.pop("fused")
    prefetch = choice.pop("prefetch")
    batch = choice.pop("batch")
    copy = choice.pop("copy")
    lists = choice.pop("contact_lists")
    per = choice.pop("per_lane")
    assert isinstance(fused, bool) and isinstance(prefetch, bool) and isinstance(batch, bool)
    assert prefetch + batch + lists <= 1 and copy is False  # the tuner keeps the context's re-sort layout
    assert set(per) == {"hash", "reorder", "integrate"} and set(per.values()) <= {1, 2, 4}
    assert set(choice.values()) <= {64, 128, 256}
    sim.step(4)
    ref.step(4)
    compare(sim.download(), ref.state())
    sim.close()


PER_LANE = [{"hash": 2, "reorder": 4, "integrate": 2}, {"hash": 4, "reorder": 2, "integrate": 4}]


@pytest.mark.parametrize("per", PER_LANE)
@pytest.mark.parametrize("name", ["pack_1024", "ensemble_5", "pile_wrap_attract", "mixed_pack_1024"])
def test_particles_per_lane_equal(gpu_ctx, name, per):
    """Two or four particles a lane in hash, reorder and integrate (split and
    fused layouts, radix-sorted and re-sorted steps): the same bits as one a
    lane."""
    make, steps = CASES[name]
    s = make()
    for fused in (False, True):
        ctx2 = api.Context(0)
        ctx2.load_scene(s)
        gpu_ctx.load_scene(s)
        gpu_ctx.set_step_variant(fused)
        ctx2.set_step_variant(fused, per_lane=per)
        a = api.Simulation(gpu_ctx, s)
        b = api.Simulation(ctx2, s)
        for chunk in (1, steps):
            a.step_graph(chunk)
            b.step_graph(chunk)
            da, db = a.download(), b.download()
            for k in ("hash", "index", "cellStart", "cellEnd", "id", "loc", "pos", "vel", "theta", "omega", "acc",
                      "alpha", "pressure"):
                assert bits_equal(da[k], db[k]), (fused, k)
        a.close()
        b.close()
        ctx2.close()
    gpu_ctx.set_step_variant()


@pytest.mark.parametrize("variant", [(True, False), (False, True), (True, True), (False, False, True),
                                     (False, False, False, True), (False, False, True, True),
                                     (True, False, True), (False, False, False, False, True),
                                     (False, True, False, True, True), (False, False, False, False, False, True),
                                     (True, False, False, False, False, True),
                                     (False, False, False, True, False, True)])
@pytest.mark.parametrize("name", ["pack_1024", "ensemble_5", "pile_wrap_attract"])
def test_step_variants_equal(gpu_ctx, name, variant):
    """collide + integrate as one kernel or two, and each way of loading
    collide's candidates (one at a time, one ahead, the first of every row
    at once, contact lists), re-sorted steps in place or into the other buffer (copy), and
    collide with 32-bit offsets or 64-bit addresses (wide): the same state
    and cell ranges, bit for bit."""
    make, steps = CASES[name]
    s = make()
    ctx2 = api.Context(0)
    ctx2.load_scene(s)
    gpu_ctx.load_scene(s)
    gpu_ctx.set_step_variant(False, False)
    ctx2.set_step_variant(*variant)
    a = api.Simulation(gpu_ctx, s)
    b = api.Simulation(ctx2, s)
    for chunk in (1, steps):
        a.step_graph(chunk)
        b.step_graph(chunk)
        da, db = a.download(), b.download()
        for k in ("hash", "index", "cellStart", "cellEnd", "id", "loc", "pos", "vel", "theta", "omega", "acc",
                  "alpha", "pressure"):
            assert bits_equal(da[k], db[k]), k
    a.close()
    b.close()
    ctx2.close()


@pytest.mark.slow
def test_settling_4m_first_steps(gpu_ctx, oracle):
    """Full-size settling workload (4,194,304 discs, 11,193,546 cells):
    hash, stable permutation and cellStart bit-exact with the oracle after
    steps 1, 2 and 12 (eager; the last ten queued in groups under the
    automatic sort choice) and 100 (graph replays of re-sorted steps), state
    within 1e-5, plus size-independent checks."""
    s = sc.settling()
    assert s.n == 4194304 and s.total_cells == 11193546
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    for chunk, graph in ((1, False), (1, False), (10, False), (88, True)):
        sim.step_graph(chunk) if graph else sim.step(chunk)
        ref.step(chunk)
        g, o = sim.download(), ref.state()
        compare(g, o)
        h = g["hash"]
        assert np.all(h[1:] >= h[:-1])                                  # sorted
        assert np.array_equal(np.sort(g["id"]), np.arange(s.n))       # permutation
        assert np.array_equal(g["loc"][g["id"]], np.arange(s.n))      # loc inverts id
    sim.close()


def wide_pack(side=256, squeeze=0.5, seed=3):
    """The settling lattice squeezed so that neighbours overlap (about four
    discs per cell), with random velocities and spins, on wide grid rows."""
    s = sc.settling(side=side)
    rng = np.random.default_rng(seed)
    s.pos[:] = (s.pos * np.float32(squeeze)).astype(np.float32)
    s.pos[:, 1] += np.float32(0.012)
    s.vel[:] = (rng.normal(size=s.vel.shape) * 0.05).astype(np.float32)
    s.omega[:] = (rng.normal(size=s.n) * 20.0).astype(np.float32)
    return s


def test_dense_wide_pack_matches_oracle(gpu_ctx, oracle):
    """65,536 per-particle-radius discs, contacts throughout, 36 candidates a
    disc: the owned step (re-sorted, velocities read through the sorted
    index) equals the oracle bit for bit after each of 3 steps, and the
    fused variant with batched candidate loads equals the split one after
    20."""
    s = wide_pack()
    gpu_ctx.load_scene(s)
    gpu_ctx.set_step_variant(False, False)
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    for _ in range(3):
        sim.step(1)
        ref.step(1)
        g = sim.download()
        compare(g, ref.state())
    ref.close()
    assert np.count_nonzero(g["pressure"]) > s.n // 4   # contacts throughout
    ctx2 = api.Context(0)
    ctx2.load_scene(s)
    ctx2.set_step_variant(True, False, True)
    fused = api.Simulation(ctx2, s)
    fused.step(3)
    sim.step_graph(17)
    fused.step_graph(17)
    da, db = sim.download(), fused.download()
    for k in ("hash", "index", "cellStart", "id", "pos", "vel", "theta", "omega", "acc", "alpha", "pressure"):
        assert bits_equal(da[k], db[k]), k
    sim.close()
    fused.close()
    ctx2.close()


@pytest.mark.parametrize("mult", [1.0, 4.0])
def test_contact_lists_on_dense_packs(gpu_ctx, oracle, mult):
    """The contact-list candidate mode on the dense pack (36 candidates a
    disc, contacts throughout), split and fused, equals the oracle after one
    step and the plain mode after 20; with cells four times wider (rows of
    about 400 candidates) it walks rows over 64 candidates one at a time."""
    from demso_amd import tuning

    s = tuning.with_cell_multiplier(wide_pack(side=128), mult)
    ref = oracle.Sim(s, "f32")
    ref.step(1)
    ctxs, sims = [], []
    for variant in ({}, {"contact_lists": True}, {"contact_lists": True, "fused": True}):
        ctx = api.Context(0)
        ctx.load_scene(s)
        ctx.set_step_variant(**variant)
        sim = api.Simulation(ctx, s)
        sim.step(1)
        compare(sim.download(), ref.state())
        sim.step_graph(19)
        ctxs.append(ctx)
        sims.append(sim)
    ref.close()
    base = sims[0].download()
    for sim in sims[1:]:
        d = sim.download()
        for k in ("hash", "index", "cellStart", "id", "pos", "vel", "theta", "omega", "acc", "alpha", "pressure"):
            assert bits_equal(base[k], d[k]), k
    for sim, ctx in zip(sims, ctxs):
        sim.close()
        ctx.close()


@pytest.mark.parametrize("n", [1, 2, 63, 65, 4097])
def test_tiny_and_ragged_sims_match_oracle(gpu_ctx, oracle, n):
    """Owned simulations of 1, 2 and ragged particle counts (partial waves,
    partial re-sort tiles): 300 steps of discs falling onto the floor and
    each other, eager and graph-replayed, bit for bit with the oracle."""
    rng = np.random.default_rng(n)
    side = int(np.ceil(np.sqrt(n)))
    base = compressed_pack(cols=max(side, 2), rows=max(side, 2), seed=n)
    keep = np.arange(n)
    s = base
    for k in ("pos", "vel", "theta", "omega", "type"):
        setattr(s, k, np.ascontiguousarray(getattr(base, k)[keep]))
    if base.radius is not None:
        s.radius = np.ascontiguousarray(base.radius[keep])
    s.id = np.arange(n, dtype=np.uint32)
    s.loc = s.id.copy()
    s.pos[:, 1] += np.float32(0.05) * rng.random(n).astype(np.float32)
    for graph in (False, True):
        gpu_ctx.load_scene(s)
        sim = api.Simulation(gpu_ctx, s)
        sim.set_sort_mode(2)
        ref = oracle.Sim(s, "f32")
        for chunk in (1, 1, 98, 200):
            sim.step_graph(chunk) if graph else sim.step(chunk)
            ref.step(chunk)
        compare(sim.download(), ref.state())
        ref.close()
        sim.close()
