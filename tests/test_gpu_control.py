"""The controlled particle (DemSo's USE_BIG_PARTICLE build) on the GPU
against the oracle, bit for bit: the stage call with the reference's collide
signature (per-particle reaction vectors, block sums, totals), and whole
steps of the owned simulation (pair term, reaction sum, kinematic move) in
every step variant, eager and replayed from graphs.

Each test uses its own context: the controlled particle is context state.
"""
import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from helpers import bits_equal, dev, host
from test_oracle_control import sorted_state, tree_blocks, two_type_scene

pytestmark = pytest.mark.gpu


def layer(n, seed):
    rng = np.random.default_rng(seed)
    side = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    pos = np.stack([gx.ravel(), gy.ravel()], 1)[:n].astype(np.float64) * 0.26 + (0.3, 0.3)
    pos += rng.normal(scale=0.01, size=pos.shape)
    s = two_type_scene(pos, np.zeros(n), grid=(32, 32), box=(12.0, 12.0), attract=5.0)
    s.vel[:] = rng.normal(scale=0.5, size=s.vel.shape).astype(np.float32)
    s.omega[:] = rng.normal(scale=3.0, size=s.n).astype(np.float32)
    return s, pos.mean(0)


@pytest.mark.parametrize("n", [1, 300, 1500])
def test_collide_controlled_stage(oracle, n):
    s, mid = layer(n, n)
    if n == 1:
        mid = s.pos[0].astype(np.float64) + (0.55, 0.1)
    system, cs, ce, st = sorted_state(oracle, s)
    cstate = (float(mid[0]), float(mid[1]), 0.2, -3.0, 0.0, 1.0)
    ctl = oracle.control(cstate[:2], 1, vel=cstate[2:4], omega=cstate[5])
    acc_o, alpha_o, pres_o, (rx_o, ry_o, rm_o), red_o = oracle.collide_controlled(system, st, cs, ce, ctl)
    assert np.count_nonzero(rx_o) >= 1

    ctx = api.Context(0)
    ctx.load_scene(s)
    blocks = -(-n // 512)
    acc, alpha, pres = dev(np.zeros((n, 2), np.float32)), dev(np.zeros(n, np.float32)), dev(np.zeros(n, np.float32))
    rx, ry, rm = (dev(np.zeros(n, np.float32)) for _ in range(3))
    bx, by, bm = (dev(np.zeros(blocks, np.float32)) for _ in range(3))
    f, m = api.collide_controlled(ctx, dev(st["pos"]), dev(st["vel"]), acc, dev(st["omega"]), alpha, dev(st["type"]),
                                  dev(cs), dev(ce), n, s.total_cells, cstate[:2], cstate[2:4], cstate[4], cstate[5], 1,
                                  rx, bx, ry, by, rm, bm, pres)
    ctx.close()
    for got, want, k in ((acc, acc_o, "acc"), (alpha, alpha_o, "alpha"), (pres, pres_o, "pressure"),
                         (rx, rx_o, "react_x"), (ry, ry_o, "react_y"), (rm, rm_o, "react_m")):
        assert bits_equal(host(got), want), k
    assert bits_equal(np.array([f[0], f[1], m], np.float32), red_o.astype(np.float32))
    # block sums: the tree of each block of 512
    for got, v in ((bx, rx_o), (by, ry_o), (bm, rm_o)):
        assert bits_equal(host(got), tree_blocks(v))


def controlled_pile():
    s = sc.triangle_pile(controlled=True)
    return s, s.notes["controlled"]


def oracle_run(oracle, s, c, steps):
    ref = oracle.Sim(s, "f32")
    ref.set_control(oracle.control(c["pos"], c["type"], move=c["move"], wrap_below_y=c["wrap_below_y"],
                                   wrap_above_x=c["wrap_above_x"], reset_y=c["reset_y"], reset_x=c["reset_x"]))
    ref.step(steps)
    o = ref.state()
    cs, red = ref.control()
    ref.close()
    order = np.argsort(o["id"])
    return {k: np.ascontiguousarray(v[order]) for k, v in o.items() if v is not None and v.shape[:1] == (s.n,)}, cs, red


STEPS = 1400  # the controlled particle reaches the pile after ~1100 steps


@pytest.fixture(scope="module")
def pile_reference(oracle):
    s, c = controlled_pile()
    return oracle_run(oracle, s, c, STEPS)


@pytest.mark.parametrize("variant", [(False, False, 1), (True, False, 1), (False, False, 2), (True, True, 2),
                                     (False, True, 0), (False, True, 2, True)])
def test_owned_sim_with_controlled_particle(pile_reference, variant):
    """pilha.dsml's pile with its MOVE particle driven through it: state,
    controlled particle and reaction bit for bit with the oracle after 1400
    steps, for split and fused collide + integrate, radix sort and re-sort,
    eager steps and graph replays, 32-bit and 64-bit addressing in
    collide."""
    fused, graph, sort_mode = variant[:3]
    wide = len(variant) > 3 and variant[3]
    o, cs, red = pile_reference
    s, c = controlled_pile()
    ctx = api.Context(0)
    ctx.load_scene(s)
    ctx.set_step_variant(fused, False, wide=wide)
    ctx.set_controlled(pos=c["pos"], type=c["type"], move=c["move"], wrap_below_y=c["wrap_below_y"],
                       reset_y=c["reset_y"], reset_x=c["reset_x"])
    sim = api.Simulation(ctx, s)
    sim.set_sort_mode(sort_mode)
    print(f"controlled pile: fused={fused} graph={graph} sort_mode={sort_mode}", flush=True)
    done = 0
    for chunk in (1, 99, STEPS - 100):
        sim.step_graph(chunk) if graph else sim.step(chunk)
        done += chunk
    g = sim.download()
    state, reaction = ctx.controlled()
    sim.close()
    ctx.close()
    order = np.argsort(g["id"])
    for k in ("pos", "vel", "theta", "omega", "acc", "alpha", "pressure"):
        assert bits_equal(np.ascontiguousarray(g[k][order]), o[k]), k
    assert state["pos"].tolist() == list(cs.pos) and cs.pos[1] < 3.0
    assert bits_equal(reaction, red.astype(np.float32)) and np.abs(red).max() > 0


def test_controlled_particle_wraps_and_is_refused_by_the_adjoint(oracle):
    """After 2200 steps the particle has wrapped from y < -1 to 10.5 and come
    down again; the adjoint refuses a context with a controlled particle."""
    s, c = controlled_pile()
    o, cs, red = oracle_run(oracle, s, c, 2200)
    ctx = api.Context(0)
    ctx.load_scene(s)
    ctx.set_controlled(pos=c["pos"], type=c["type"], move=c["move"])
    sim = api.Simulation(ctx, s)
    sim.step_graph(2200)
    g = sim.download()
    state, reaction = ctx.controlled()
    order = np.argsort(g["id"])
    assert bits_equal(np.ascontiguousarray(g["pos"][order]), o["pos"])
    assert state["pos"].tolist() == list(cs.pos) and 9.0 < cs.pos[1] < 10.5
    with pytest.raises(api.DemError):
        api.Adjoint(sim, 100)
    # switching it off restores the plain step
    ctx.set_controlled(enabled=False, wrap_above_x=0.0)
    sim.step(1)
    sim.close()
    ctx.close()
