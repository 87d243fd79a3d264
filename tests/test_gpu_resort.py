"""The re-sort of a state already in the order of its previous keys
(only the particles that changed cell are sorted and merged back) against
the radix sort and the oracle: identical keys, permutation, cell ranges and
state, bit for bit, for few movers (sorted within one wave), hundreds of
movers (bitonic stages across waves through LDS), many movers (the chunked
radix path of the movers workgroup) and ensembles."""
import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from helpers import bits_equal, compressed_pack

pytestmark = pytest.mark.gpu

FIELDS = ("hash", "index", "cellStart", "cellEnd", "id", "type", "loc", "pos", "vel", "theta", "omega", "acc",
          "alpha", "pressure")


def same(a, b):
    # cellEnd too where cellStart is EMPTY: both keep the value the cell's
    # last non-empty step wrote, as reorderDataAndFindCellStartD does
    for k in FIELDS:
        assert bits_equal(a[k], b[k]), k


def fast_column(n_side=128, speed=30.0, seed=5, dt=1e-4):
    """A sparse lattice (3 cm pitch, 1.1 cm cells) of soft discs with random
    velocities of tens of m/s: with dt = 1e-4 s about half of them change
    cell every step."""
    s = sc.column_collapse(seed=seed, cols=n_side, rows=n_side, box=(n_side / 16.0, n_side / 32.0), pitch=0.03,
                           dt=dt, kn=1e4, kt=2e3)
    rng = np.random.default_rng(seed)
    s.vel[:] = (rng.normal(size=s.vel.shape) * speed).astype(np.float32)
    return s


CASES = {
    "pack_few_movers": (lambda: compressed_pack(seed=3), 30),
    "fast_many_movers": (lambda: fast_column(), 12),
    "hundreds_of_movers": (lambda: fast_column(speed=3.0), 12),
    "ensemble_7": (lambda: sc.ensemble(instances=7, counts=[3000, 700, 5200, 90, 1640, 4100, 2222], seed=6), 25),
    "pile_wrap": (lambda: sc.triangle_pile(), 40),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("copy", [False, True], ids=["inplace", "copy"])
def test_resort_equals_radix(gpu_ctx, name, graph, copy):
    """Re-sorted steps (in place, or reordered into the other buffer) equal
    radix-sorted ones."""
    make, steps = CASES[name]
    s = make()
    gpu_ctx.set_step_variant(copy=copy)
    a = api.Simulation(gpu_ctx, s)
    b = api.Simulation(gpu_ctx, s)
    a.set_sort_mode(1)  # radix sort every step
    b.set_sort_mode(2)  # re-sort whenever possible
    movers = []
    try:
        for chunk in (1, 2, steps):
            for sim in (a, b):
                (sim.step_graph if graph else sim.step)(chunk)
            movers.append(b.mover_count())
            print(f"{name} graph={graph} copy={copy}: movers after {chunk} more steps: {movers[-1]}", flush=True)
            same(a.download(), b.download())
    finally:
        gpu_ctx.set_step_variant()
        a.close()
        b.close()
    if name == "fast_many_movers":
        assert max(movers) > 4096  # the chunked radix path of the movers workgroup ran
    if name == "hundreds_of_movers":
        assert 64 < max(movers) <= 4096  # the LDS bitonic stages across waves ran


def test_mover_count_is_cell_changes(gpu_ctx):
    """The mover count of a re-sorted step is the number of particles whose
    cell key differs from the previous step's; the in-place re-sort lists
    those and every particle whose slot changed."""
    s = fast_column(n_side=64, speed=3.0)
    sim = api.Simulation(gpu_ctx, s)
    sim.set_sort_mode(2)
    sim.step(2)
    for _ in range(3):
        d0 = sim.download()
        sim.step(1)
        d1 = sim.download()
        k0 = np.empty(s.n, np.uint32)
        k1 = np.empty(s.n, np.uint32)
        k0[d0["id"]] = d0["hash"]
        k1[d1["id"]] = d1["hash"]
        s0 = np.empty(s.n, np.int64)
        s1 = np.empty(s.n, np.int64)
        s0[d0["id"]] = np.arange(s.n)
        s1[d1["id"]] = np.arange(s.n)
        changed = int((k0 != k1).sum())
        listed = int(((k0 != k1) | (s0 != s1)).sum())
        assert sim.mover_count() == changed
        assert sim.resort_counts() == {"movers": changed, "listed": listed}
        assert changed > 0 and listed > changed
    sim.close()


@pytest.mark.parametrize("graph", [False, True])
def test_auto_sort_under_bursts_equals_radix(gpu_ctx, graph):
    """The automatic choice on a state where most particles change slot
    every step (it alternates radix-sorted runs with re-sort trials) gives
    the radix-sorted steps' state."""
    s = fast_column()
    a = api.Simulation(gpu_ctx, s)
    b = api.Simulation(gpu_ctx, s)
    a.set_sort_mode(1)
    try:
        listed = []
        for chunk in (1, 1, 2, 3, 8, 16):
            for sim in (a, b):
                (sim.step_graph if graph else sim.step)(chunk)
            listed.append(b.resort_counts()["listed"])
            same(a.download(), b.download())
        print(f"graph={graph}: listed after each chunk {listed}", flush=True)
        assert max(listed) > s.n // 4  # the radix-sorted runs were taken
    finally:
        a.close()
        b.close()


def test_resort_matches_oracle_long(gpu_ctx, oracle):
    """Default (automatic) sort choice over 200 steps of a column collapse
    against the oracle's stable sort."""
    s = sc.column_collapse(cols=48, rows=48)
    sim = api.Simulation(gpu_ctx, s)
    ref = oracle.Sim(s, "f32")
    sim.step_graph(200)
    ref.step(200)
    g, o = sim.download(), ref.state()
    for k in ("hash", "index", "cellStart", "id", "pos", "vel", "acc", "pressure"):
        assert bits_equal(g[k], o[k]), k
    sim.close()
