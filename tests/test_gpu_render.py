"""Images on the GPU (dem_plot_particles, dem_sim_plot,
dem_update_pressure_scale) against the oracle's sequential restatement of
plotParticles, byte for byte: overlapping discs, rotation lines, pressure
colours, the controlled particle, discs partly outside the image, per-particle
radii, and the owned simulation's image after steps."""
import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from helpers import compressed_pack, dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rctx():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    ctx = api.Context(0)
    yield ctx
    ctx.close()


def colored(s):
    pal = [(255.0, 64.0, 0.0), (30.0, 200.0, 90.0), (0.0, 128.0, 255.0)]
    for k, m in enumerate(s.materials):
        m.color = pal[k % 3]
    return s


def gpu_image(ctx, s, rp, theta, pressure, control=None, radius=None):
    import torch

    img = torch.zeros((rp["image_dim"][1], rp["image_dim"][0], 4), dtype=torch.uint8, device="cuda")
    api.plotParticles(ctx, img, dev(s.pos), dev(theta), dev(s.type), s.n, rp, dev(pressure),
                      radius=None if radius is None else dev(radius), control=control)
    return img.cpu().numpy()


def check(a, b):
    diff = np.argwhere(np.any(a != b, axis=-1))
    assert diff.size == 0, f"{len(diff)} pixels differ, first at {diff[:3].tolist()}"


def stepped(oracle, s, steps):
    ref = oracle.Sim(s, "f32")
    ref.step(steps)
    st = ref.state()
    ref.close()
    return st


@pytest.mark.parametrize("rot,by_p,bg", [(True, False, 0), (False, True, 0), (True, True, 255)])
def test_pile_images(rctx, oracle, rot, by_p, bg):
    s = colored(sc.triangle_pile(controlled=True))
    st = stepped(oracle, s, 400)
    rctx.load_scene(s)
    pmax = oracle.pressure_scale(st["pressure"])
    rp = api.render_params(s, image_height=800, controlled_type=1, bg_color=bg, view_rotations=rot,
                           color_by_pressure=by_p, max_pressure=pmax)
    print(f"pile image rot={rot} by_pressure={by_p} bg={bg}", flush=True)
    ctl = oracle.control((5.0, 3.2), 1)
    s2 = s.copy()
    s2.pos, s2.type = st["pos"], st["type"]
    got = gpu_image(rctx, s2, rp, st["theta"], st["pressure"], control={"pos": (5.0, 3.2), "type": 1})
    want = oracle.plot(s, rp, st["pos"], st["theta"], st["type"], st["pressure"], ctl=ctl)
    check(got, want)
    assert (got[..., 3] > 0).sum() > 10000


def test_dense_pack_and_outside_positions(rctx, oracle):
    """A compressed pack (every disc overlaps its neighbours) plus discs left
    of, below, right of and above the image (unsigned wrap, last-pixel clamp)."""
    s = colored(compressed_pack(cols=48, rows=48))
    extra = np.array([[-0.01, 0.2], [0.5, -0.004], [1.2, 0.3], [0.3, 0.7], [-3.0, -3.0], [np.nan, 0.1]], np.float32)
    s.pos = np.concatenate([s.pos, extra])
    n = s.pos.shape[0]
    s.type = np.concatenate([s.type, np.zeros(len(extra), np.uint32)])
    theta = np.random.default_rng(0).uniform(-1.0, 7.0, n).astype(np.float32)
    theta[-1] = np.inf
    s.theta = theta
    pressure = np.random.default_rng(1).uniform(0, 2, n).astype(np.float32)
    s.vel = np.zeros((n, 2), np.float32)
    s.omega = np.zeros(n, np.float32)
    s.id = np.arange(n, dtype=np.uint32)
    s.loc = s.id.copy()
    radius = None
    if s.radius is not None:
        radius = np.concatenate([s.radius, np.full(len(extra), 0.02, np.float32)])
        s.radius = radius
    rctx.load_scene(s)
    for rot, by_p in ((True, False), (True, True)):
        rp = api.render_params(s, image_height=512, view_rotations=rot, color_by_pressure=by_p, max_pressure=1.5)
        got = gpu_image(rctx, s, rp, theta, pressure, radius=radius)
        want = oracle.plot(s, rp, s.pos, theta, s.type, pressure, radius=radius)
        check(got, want)
        assert (got[..., 3] > 0).sum() > 20000


def test_per_particle_radius(rctx, oracle):
    s = colored(sc.settling(side=48))
    rctx.load_scene(s)
    rp = api.render_params(s, image_height=600)
    got = gpu_image(rctx, s, rp, s.theta, np.zeros(s.n, np.float32), radius=s.radius)
    want = oracle.plot(s, rp, s.pos, s.theta, s.type, np.zeros(s.n, np.float32), radius=s.radius)
    check(got, want)


def test_sim_plot_after_steps(oracle):
    """dem_sim_plot of an owned simulation with its controlled particle after
    1,300 steps equals the oracle's image of the oracle's state."""
    s = colored(sc.triangle_pile(controlled=True))
    c = s.notes["controlled"]
    ctx = api.Context(0)
    ctx.load_scene(s)
    ctx.set_controlled(pos=c["pos"], type=c["type"], move=c["move"])
    sim = api.Simulation(ctx, s)
    sim.step_graph(1300)
    ref = oracle.Sim(s, "f32")
    ref.set_control(oracle.control(c["pos"], c["type"], move=c["move"], wrap_below_y=c["wrap_below_y"],
                                   wrap_above_x=c["wrap_above_x"], reset_y=c["reset_y"], reset_x=c["reset_x"]))
    ref.step(1300)
    st = ref.state()
    cs, _ = ref.control()
    ref.close()
    pmax = api.updatePressureScale(ctx, dev(st["pressure"]), s.n)
    assert pmax == oracle.pressure_scale(st["pressure"])
    for by_p in (False, True):
        rp = api.render_params(s, image_height=800, controlled_type=1, color_by_pressure=by_p, max_pressure=pmax)
        got = sim.plot(rp)
        want = oracle.plot(s, rp, st["pos"], st["theta"], st["type"], st["pressure"], ctl=cs)
        check(got, want)
    sim.close()
    ctx.close()


@pytest.mark.parametrize("vals", [[0.0, 3.0, 2.5], [0.0, 0.0], [-1.0, float("nan"), 0.5], [-2.0, -1.0]])
def test_update_pressure_scale(rctx, oracle, vals):
    v = np.asarray(vals, np.float32)
    assert api.updatePressureScale(rctx, dev(v), v.size) == oracle.pressure_scale(v)


def test_update_pressure_scale_large(rctx, oracle):
    v = np.random.default_rng(5).exponential(size=3_000_001).astype(np.float32)
    assert api.updatePressureScale(rctx, dev(v), v.size) == float(v.max())
