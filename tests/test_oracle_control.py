"""The oracle's controlled particle (DemSo's USE_BIG_PARTICLE build), pinned
by properties the reference fixes:

* the pair with the controlled particle is the ordinary pair law: a particle
  next to it gets the force it would get from a regular particle of the same
  type at the same place (src/particles_kernel.cuh:525-531 call the same
  collideSpheres / attractSpheres as the neighbour loop);
* the reaction per particle is minus that pair force, and its moment is the
  particle's pair moment scaled by -rC/rA (src/particles_kernel.cuh:537-541),
  which by the pair law equals minus the moment a regular particle in the
  controlled particle's place would receive;
* the total reaction is reduceVecD's order: a pairwise tree per block of 512,
  block sums added in order (src/particles_kernel.cuh:552-578,
  src/functions.cu:475-513);
* the move after each step is pos.y += -0.005 in double, with the wraps of
  src/main.cu:510-514.
"""
from __future__ import annotations

import numpy as np
import pytest

from demso_amd.scene import Instance, Material, Scene

F32 = np.float32


def two_type_scene(pos, types, vel=None, omega=None, box=(10.0, 10.0), grid=(16, 16), attract=0.0):
    pos = np.asarray(pos, np.float32).reshape(-1, 2)
    n = pos.shape[0]

    def mat(r, m, kn, kt, mu, a):
        inertia = float(F32(2) * F32(m) * F32(r) * F32(r) / F32(5))
        return Material(radius=r, mass=m, inertia=inertia, kn=kn, kt=kt, cn=0.2, bdamp=0.5, mu=mu, attract=a)

    mats = [mat(0.125, 1e-2, 1000.0, 1000.0, 0.6, attract), mat(0.5, 100.0, 5e4, 5e4, 0.4, 0.0)]
    return Scene(
        name="control", pos=pos,
        vel=np.zeros((n, 2), np.float32) if vel is None else np.asarray(vel, np.float32).reshape(-1, 2),
        theta=np.zeros(n, np.float32),
        omega=np.zeros(n, np.float32) if omega is None else np.asarray(omega, np.float32),
        id=np.arange(n, dtype=np.uint32), type=np.asarray(types, np.uint32), loc=np.arange(n, dtype=np.uint32),
        radius=None, materials=mats, instances=[Instance(box=box, extent=box, grid=grid)], dt=1e-3,
        gravity=(0.0, -9.81), bkn=1000.0, bkt=100.0, flags=0)


def sorted_state(oracle, s):
    system = oracle.System(s)
    h, i = oracle.calc_hash(system, s.pos)
    h, i = oracle.sort_particles(h, i)
    old = dict(pos=s.pos, vel=s.vel, theta=s.theta, omega=s.omega, id=s.id, type=s.type, radius=s.radius)
    cs, ce, st = oracle.reorder(h, i, old, s.total_cells)
    return system, cs, ce, st


@pytest.mark.parametrize("attract", [0.0, 10.0])
def test_control_pair_is_the_neighbour_pair(oracle, attract):
    """Particle A beside the controlled particle C gets bit for bit the
    acceleration, spin and pressure it gets from a regular particle of C's
    type in C's place; the reaction is minus the force on that regular
    particle and minus its moment (to FP32 rounding of the two evaluations)."""
    pa, pc = (5.0, 5.0), (5.0 + 0.55, 5.0 + 0.1)
    va, vc = (0.3, -0.2), (0.0, -5.0)
    wa, wc = 2.0, 0.0
    pair = two_type_scene([pa, pc], [0, 1], vel=[va, vc], omega=[wa, wc], attract=attract)
    system, cs, ce, st = sorted_state(oracle, pair)
    acc2, alpha2, pres2 = oracle.collide(system, st, cs, ce)
    a = int(np.flatnonzero(st["id"] == 0)[0])
    b = int(np.flatnonzero(st["id"] == 1)[0])

    alone = two_type_scene([pa], [0], vel=[va], omega=[wa], attract=attract)
    system1, cs1, ce1, st1 = sorted_state(oracle, alone)
    ctl = oracle.control(pc, 1, vel=vc, omega=wc)
    acc1, alpha1, pres1, (rx, ry, rm), red = oracle.collide_controlled(system1, st1, cs1, ce1, ctl)
    assert np.abs(acc1[0]).max() > 0 and alpha1[0] != 0
    assert acc1[0].view(np.uint32).tolist() == acc2[a].view(np.uint32).tolist()
    assert alpha1[0] == alpha2[a] and pres1[0] == pres2[a]

    mats = pair.materials
    force_b = acc2[b].astype(np.float64) * mats[1].mass
    moment_b = float(alpha2[b]) * mats[1].inertia
    np.testing.assert_allclose([rx[0], ry[0]], force_b, rtol=1e-5, atol=1e-6 * np.abs(force_b).max())
    assert rm[0] == pytest.approx(-moment_b, rel=1e-5)
    # one particle: the total is its reaction
    assert red.tolist() == [rx[0], ry[0], rm[0]]


def tree_blocks(v):
    """reduceVecD's per-block sums in float32, written out independently of
    the oracle: blocks of 512 (zero padded), halving pairwise sums."""
    v = np.asarray(v, np.float32)
    blocks = -(-v.size // 512)
    c = np.zeros(blocks * 512, np.float32)
    c[:v.size] = v
    c = c.reshape(blocks, 512)
    h = 256
    while h:
        c[:, :h] = c[:, :h] + c[:, h:2 * h]
        h //= 2
    return c[:, 0].copy()


def tree_total(v):
    """The block sums added in order to a total that starts at 0
    (src/functions.cu:505-513)."""
    total = F32(0)
    for b in tree_blocks(v):
        total = F32(total + b)
    return total


@pytest.mark.parametrize("n", [300, 512, 1500])
def test_reaction_total_is_reduce_vec_order(oracle, n):
    """The oracle's total equals the tree order and, for a sample where it
    matters, not the plain running sum."""
    rng = np.random.default_rng(n)
    # a dense layer right under the controlled particle
    side = int(np.ceil(np.sqrt(n)))
    gx, gy = np.meshgrid(np.arange(side), np.arange(side))
    pos = np.stack([gx.ravel(), gy.ravel()], 1)[:n].astype(np.float64) * 0.26 + (0.3, 0.3)
    pos += rng.normal(scale=0.01, size=pos.shape)
    s = two_type_scene(pos, np.zeros(n), grid=(32, 32), box=(12.0, 12.0))
    s.vel[:] = rng.normal(scale=0.5, size=s.vel.shape).astype(np.float32)
    system, cs, ce, st = sorted_state(oracle, s)
    # a big controlled disc over the middle of the layer
    mid = pos.mean(0)
    ctl = oracle.control((float(mid[0]), float(mid[1])), 1, vel=(0.2, -3.0), omega=1.0)
    _, _, _, (rx, ry, rm), red = oracle.collide_controlled(system, st, cs, ce, ctl)
    assert np.count_nonzero(rx) > 3
    for k, v in enumerate((rx, ry, rm)):
        assert red[k] == tree_total(v), k


def test_reaction_reduction_differs_from_running_sum():
    """The tree order matters in float32 (so the order test above can fail)."""
    v = np.full(1024, 1.0, np.float32)
    v[0] = 2.0 ** 25
    running = F32(0)
    for x in v:
        running = F32(running + x)
    assert tree_total(v) != running


def test_move_and_wraps(oracle):
    """pos.y += -0.005 (double), then y < -1 -> reset_y and
    x > Lx + 0.5 -> reset_x, after each step (src/main.cu:510-514)."""
    s = two_type_scene([(1.0, 8.0)], [0])
    sim = oracle.Sim(s, "f32")
    y0 = F32(-0.996)
    ctl = oracle.control((10.4995, float(y0)), 1, move=(0.001, -0.005), wrap_below_y=-1.0, wrap_above_x=10.5,
                         reset_y=10.5, reset_x=-0.5)
    sim.set_control(ctl)
    sim.step(1)
    c, _ = sim.control()
    # -0.996 - 0.005 = -1.001 < -1 -> 10.5; 10.4995 + 0.001 = 10.5005 > 10.5 -> -0.5
    assert c.pos[1] == 10.5 and c.pos[0] == -0.5
    sim.step(1)
    c, _ = sim.control()
    assert c.pos[1] == F32(np.float64(F32(10.5)) - 0.005)
    assert c.pos[0] == F32(np.float64(F32(-0.5)) + 0.001)
    sim.close()
    # DemSo's run: 9.5 - 0.005 per step stays above -1 for 2100 steps
    s2 = two_type_scene([(1.0, 8.0)], [0])
    sim = oracle.Sim(s2, "f32")
    sim.set_control(oracle.control((5.0, 9.5), 1, move=(0.0, -0.005), wrap_below_y=-1.0, wrap_above_x=10.5,
                                   reset_y=10.5, reset_x=-0.5))
    y = F32(9.5)
    wrapped_at = None
    for k in range(1, 2201):
        y = F32(np.float64(y) - 0.005)
        if y < -1:
            y = F32(10.5)
            wrapped_at = wrapped_at or k
    sim.step(2200)
    c, _ = sim.control()
    assert c.pos[1] == y and 2095 < wrapped_at < 2105
    sim.close()
