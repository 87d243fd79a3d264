"""Pins of the CPU oracle by facts it does not share code with: library
routines (numpy's stable sort, searchsorted), brute force, closed-form
physics (free fall under symplectic Euler, undamped/damped two-body contact,
slip-to-roll on the floor) and conservation laws.  The reference ships no
golden vectors or tests for this path, so these are the pins.
"""
import math

import numpy as np
import pytest

from demso_amd import scene as sc
from demso_amd.scene import Instance, Material, Scene

DEM_GRID_CLAMP, DEM_OPEN_TOP = 1, 2


def one_type_scene(pos, vel=None, omega=None, radius=0.1, mass=0.1, kn=1e3, kt=1e3, cn=0.0, mu=0.0,
                   bdamp=0.5, attract=0.0, box=(10.0, 10.0), grid=(16, 16), gravity=(0.0, 0.0), dt=1e-3,
                   bkn=1e3, bkt=1e3, flags=0, radius_array=None, areal=0.0):
    pos = np.asarray(pos, np.float32).reshape(-1, 2)
    n = pos.shape[0]
    inertia = float(np.float32(2) * np.float32(mass) * np.float32(radius) * np.float32(radius) / np.float32(5))
    m = Material(radius=radius, mass=mass, inertia=inertia, kn=kn, kt=kt, cn=cn, bdamp=bdamp, mu=mu,
                 attract=attract, areal=areal)
    ext = (box[0], box[1])
    return Scene(
        name="kat", pos=pos,
        vel=np.zeros((n, 2), np.float32) if vel is None else np.asarray(vel, np.float32).reshape(-1, 2),
        theta=np.zeros(n, np.float32),
        omega=np.zeros(n, np.float32) if omega is None else np.asarray(omega, np.float32),
        id=np.arange(n, dtype=np.uint32), type=np.zeros(n, np.uint32), loc=np.arange(n, dtype=np.uint32),
        radius=None if radius_array is None else np.asarray(radius_array, np.float32),
        materials=[m], instances=[Instance(box=box, extent=ext, grid=grid)], dt=dt, gravity=gravity,
        bkn=bkn, bkt=bkt, flags=flags)


# ---- hashing -----------------------------------------------------------------

def test_hash_cell_centres_and_edges(oracle):
    """Cell centres land in their own cell; 0 is cell 0; out-of-box positions
    clamp to the edge cells (clamp mode) or wrap mod G (wrap mode)."""
    G, L = 16, 8.0  # cell size 0.5, exact in binary
    xs = np.array([(i + 0.5) * 0.5 for i in range(G)] + [0.0, 7.999, -0.25, 8.0, 9.3, -3.0], np.float32)
    pos = np.stack([xs, np.full_like(xs, 0.25)], 1)
    for flags, expect_tail in ((DEM_GRID_CLAMP, [0, 15, 0, 15, 15, 0]), (0, [0, 15, 15, 0, 2, 10])):
        s = one_type_scene(pos, box=(L, L), grid=(G, G), flags=flags)
        h, idx = oracle.calc_hash(oracle.System(s), s.pos)
        assert list(h[:G]) == list(range(G))
        assert list(h[G:]) == expect_tail
        assert list(idx) == list(range(len(xs)))


def test_hash_row_major_and_instances(oracle):
    """key = cell_base + gy*Gx + gx, with the instance found from the type."""
    ens = sc.ensemble(instances=3, counts=[50, 70, 30], seed=5)
    system = oracle.System(ens)
    h, _ = oracle.calc_hash(system, ens.pos, ens.type)
    for k, (a, b) in enumerate(zip(ens.segments[:-1], ens.segments[1:])):
        g = ens.instances[k]
        cell = g.extent[0] / g.grid[0]
        gx = np.clip(np.floor(ens.pos[a:b, 0].astype(np.float64) / cell), 0, g.grid[0] - 1)
        gy = np.clip(np.floor(ens.pos[a:b, 1].astype(np.float64) / cell), 0, g.grid[1] - 1)
        exact = g.cell_base + gy * g.grid[0] + gx
        # float32 rounding may move a disc sitting within an ulp of a cell
        # boundary; everything else is the exact cell.
        assert np.mean(h[a:b] == exact) > 0.99
        assert np.all(h[a:b] >= g.cell_base) and np.all(h[a:b] < g.cell_base + g.cells)


# ---- sorting -----------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 2, 17, 4096, 50001])
def test_sort_is_numpy_stable_sort(oracle, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, max(1, n // 7), size=n, dtype=np.uint32)  # heavy ties
    vals = rng.permutation(n).astype(np.uint32)
    h, v = oracle.sort_particles(keys, vals)
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(h, keys[order])
    assert np.array_equal(v, vals[order])


# ---- reorder / cellStart -------------------------------------------------------

def test_reorder_matches_linear_scan(oracle):
    s = sc.column_collapse(cols=20, rows=20)
    system = oracle.System(s)
    h, i = oracle.calc_hash(system, s.pos)
    h, i = oracle.sort_particles(h, i)
    old = dict(pos=s.pos, vel=s.vel + 1.0, theta=s.theta + 2.0, omega=s.omega + 3.0,
               id=s.id[::-1].copy(), type=s.type, radius=s.radius)
    cs, ce, srt = oracle.reorder(h, i, old, s.total_cells)
    cells = np.arange(s.total_cells)
    first = np.searchsorted(h, cells, side="left")
    last = np.searchsorted(h, cells, side="right")
    occupied = last > first
    assert np.array_equal(cs[occupied], first[occupied])
    assert np.array_equal(ce[occupied], last[occupied])
    assert np.all(cs[~occupied] == 0xFFFFFFFF)
    for k in ("pos", "vel", "theta", "omega", "id", "type", "radius"):
        assert np.array_equal(srt[k], old[k][i])
    assert np.array_equal(srt["loc"][srt["id"]], np.arange(s.n))


# ---- forces ------------------------------------------------------------------

def all_pairs_contacts(pos, radius):
    d = pos[None, :, :].astype(np.float64) - pos[:, None, :].astype(np.float64)
    dist = np.sqrt((d ** 2).sum(-1))
    R = radius[:, None] + radius[None, :]
    c = dist < R
    np.fill_diagonal(c, False)
    return c


def test_neighbour_search_finds_every_contact(oracle):
    """The 3x3 cell search finds exactly the contacts of an all-pairs sweep:
    a particle's force is non-zero iff an all-pairs sweep says it touches."""
    s = sc.column_collapse(cols=24, rows=24, pitch=0.0098, jitter=0.0012)  # many overlaps
    system = oracle.System(s)
    h, i = oracle.calc_hash(system, s.pos)
    h, i = oracle.sort_particles(h, i)
    old = dict(pos=s.pos, vel=s.vel, theta=s.theta, omega=s.omega, id=s.id, type=s.type, radius=s.radius)
    cs, ce, st = oracle.reorder(h, i, old, s.total_cells)
    acc, alpha, pres = oracle.collide(system, st, cs, ce)
    touching = all_pairs_contacts(st["pos"], st["radius"]).any(1)
    assert touching.sum() > 100
    assert np.array_equal(np.linalg.norm(acc, axis=1) > 0, touching)
    assert np.array_equal(pres > 0, touching)


def test_newton_third_law(oracle):
    """Pair forces are antisymmetric, so internal forces sum to ~zero
    (no gravity, no walls reached)."""
    s = sc.column_collapse(cols=24, rows=24, pitch=0.0098, jitter=0.0012)
    s.pos[:, 0] += 0.2
    s.pos[:, 1] += 0.05
    s.vel[:] = np.random.default_rng(1).normal(size=s.vel.shape).astype(np.float32) * 0.1
    s.omega[:] = np.random.default_rng(2).normal(size=s.n).astype(np.float32) * 5
    system = oracle.System(s)
    h, i = oracle.calc_hash(system, s.pos)
    h, i = oracle.sort_particles(h, i)
    old = dict(pos=s.pos, vel=s.vel, theta=s.theta, omega=s.omega, id=s.id, type=s.type, radius=s.radius)
    cs, ce, st = oracle.reorder(h, i, old, s.total_cells)
    acc, _, _ = oracle.collide(system, st, cs, ce)
    m = (np.float64(s.materials[0].areal) * st["radius"] * st["radius"])
    F = acc.astype(np.float64) * m[:, None]
    assert np.abs(F.sum(0)).max() <= 1e-5 * np.abs(F).sum()


def run_steps(oracle, s, steps, prec="f64"):
    sim = oracle.Sim(s, prec)
    sim.step(steps)
    st = sim.state()
    sim.close()
    order = np.argsort(st["id"])
    return {k: (v[order] if v is not None and v.shape[:1] == (s.n,) else v) for k, v in st.items()}


def test_free_fall_symplectic_euler_closed_form(oracle):
    """v_k = v0 + k g dt and x_k = x0 + dt * sum_{j=1..k} v_j (velocity first)."""
    dt, g, k = 1e-3, -9.81, 50
    s = one_type_scene([[5.0, 8.0]], vel=[[0.3, 1.0]], gravity=(0.0, g), dt=dt)
    st = run_steps(oracle, s, k)
    # parameters are FP32 values even in the FP64 run
    dt, g, vx = float(np.float32(dt)), float(np.float32(g)), float(np.float32(0.3))
    v = 1.0 + np.arange(1, k + 1) * g * dt
    assert st["vel"][0, 1] == pytest.approx(v[-1], rel=1e-12)
    assert st["pos"][0, 1] == pytest.approx(8.0 + dt * v.sum(), rel=1e-12)
    assert st["pos"][0, 0] == pytest.approx(5.0 + k * dt * vx, rel=1e-12)


def test_spin_and_theta_wrap(oracle):
    dt, w, k = 1e-3, 40.0, 400  # 16 rad: wraps twice
    s = one_type_scene([[5.0, 5.0]], omega=[w], dt=dt)
    st = run_steps(oracle, s, k)
    dt = float(np.float32(dt))
    assert st["theta"][0] == pytest.approx(math.fmod(k * w * dt, 2 * math.pi), abs=1e-9)
    assert 0 <= st["theta"][0] < 2 * math.pi


def test_side_wall_reflect(oracle):
    """Side walls clamp to r and reverse the normal velocity scaled by the
    boundary damping."""
    s = one_type_scene([[9.895, 5.0]], vel=[[2.0, 0.0]], bdamp=0.5, dt=1e-2)
    st = run_steps(oracle, s, 1)
    assert st["pos"][0, 0] == pytest.approx(10.0 - float(np.float32(0.1)), rel=1e-12)
    assert st["vel"][0, 0] == pytest.approx(-1.0, rel=1e-12)


def two_body(kn, cn, v0, dt, steps, mass=0.1, r=0.1):
    return one_type_scene([[4.8, 5.0], [5.0 + 1e-4, 5.0]], vel=[[v0, 0.0], [0.0, 0.0]], radius=r, mass=mass,
                          kn=kn, cn=cn, mu=0.0, dt=dt)


def test_two_body_undamped_contact(oracle):
    """Undamped head-on contact (colisao-elastica.dsml, damping 0): the pair
    spring is kn/2 (series), m_eff = m/2, so contact lasts pi*sqrt(m/kn) and
    the velocities are exchanged."""
    kn, m, v0, dt = 1000.0, 0.1, 2.0, 1e-5
    t_contact = math.pi * math.sqrt(m / kn)
    steps = int(0.25 / dt)
    s = two_body(kn, 0.0, v0, dt, steps)
    sim = oracle.Sim(s, "f64")
    in_contact, was = [], False
    for k in range(steps // 100):
        sim.step(100)
        st = sim.state()
        o = np.argsort(st["id"])
        p = st["pos"][o]
        touching = abs(p[1, 0] - p[0, 0]) < 0.2
        in_contact.append(touching)
    st = sim.state()
    o = np.argsort(st["id"])
    v = st["vel"][o]
    assert v[0, 0] == pytest.approx(0.0, abs=2e-3 * v0)
    assert v[1, 0] == pytest.approx(v0, rel=2e-3)
    dur = sum(in_contact) * 100 * dt
    assert dur == pytest.approx(t_contact, abs=2 * 100 * dt)


@pytest.mark.parametrize("e", [0.5, 0.8])
def test_two_body_restitution(oracle, e):
    """Approach-only dashpot: the per-type damping from
    damping_for_restitution gives the target restitution."""
    kn, m, v0, dt = 1e5, 0.1, 1.0, 1e-6
    cn = sc.damping_for_restitution(e, kn, m)
    s = two_body(kn, cn, v0, dt, 0)
    st = run_steps(oracle, s, 20000)
    v = st["vel"]
    e_meas = (v[1, 0] - v[0, 0]) / v0
    assert e_meas == pytest.approx(e, abs=5e-3)
    assert v[0, 0] + v[1, 0] == pytest.approx(v0, rel=1e-9)  # momentum


def test_slip_to_roll_on_floor(oracle):
    """rolamento-sobre-plano.dsml: a disc sliding with backspin on the floor
    (v0=1, w0=-7.5, r=0.1, mu=0.4, I=2/5 m r^2) ends rolling at
    v = v0 - (v0 + w0 r)/3.5 with w = -v/r."""
    s = one_type_scene([[2.0, 0.1]], vel=[[1.0, 0.0]], omega=[-7.5], radius=0.1, mass=0.1, kn=1e5, kt=1e7,
                       cn=0.0, mu=0.4, bdamp=0.0, gravity=(0.0, -9.81), dt=0.25e-4, bkn=1e5, bkt=1e7)
    st = run_steps(oracle, s, 8000)  # 0.2 s, as the scene's simduration
    v_roll = 1.0 - (1.0 - 0.75) / 3.5
    assert st["vel"][0, 0] == pytest.approx(v_roll, abs=2e-3)
    assert st["omega"][0] == pytest.approx(-v_roll / 0.1, abs=2e-2)


def test_attraction_magnitude(oracle):
    """attractSpheres: inside 2.4 r each side pulls with a (2.4 r - d)."""
    r, a, d = 0.1, 10.0, 0.22
    s = one_type_scene([[5.0, 5.0], [5.0 + d, 5.0]], radius=r, attract=a, mass=1.0, kn=0.0, cn=0.0)
    system = oracle.System(s)
    h, i = oracle.calc_hash(system, s.pos)
    h, i = oracle.sort_particles(h, i)
    old = dict(pos=s.pos, vel=s.vel, theta=s.theta, omega=s.omega, id=s.id, type=s.type)
    cs, ce, st = oracle.reorder(h, i, old, s.total_cells)
    acc, _, _ = oracle.collide(system, st, cs, ce)
    # d > 2r: no contact, only attraction; both sides pull => 2*a*(2.4r-d)
    d32 = float(np.float32(5.0 + d) - np.float32(5.0))
    reach = float(np.float32(float(np.float32(2 * r)) * 1.2))
    expect = 2 * a * (reach - d32)
    k0 = int(np.where(st["id"] == 0)[0][0])
    assert acc[k0, 0] == pytest.approx(expect, rel=1e-5)
    assert acc[1 - k0, 0] == pytest.approx(-expect, rel=1e-5)


def test_fp32_tracks_fp64(oracle):
    """FP32 oracle vs the FP64 run of the same restatement over 200 steps of
    the column collapse: the trajectory error stays small."""
    s = sc.column_collapse(cols=16, rows=16)
    a = run_steps(oracle, s, 200, "f32")
    b = run_steps(oracle, s, 200, "f64")
    err = np.abs(a["pos"] - b["pos"]).max()
    assert err < 1e-5
