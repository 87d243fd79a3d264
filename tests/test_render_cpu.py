"""The oracle's image output (render_oracle.hpp), pinned by what the
reference's plot arithmetic fixes: disc coverage is the integer lattice
x^2 + y^2 < r^2 around the truncated centre, the centre pixel carries the
full colour, rotation lines run along theta, a later disc covers an earlier
one, the colour bar's values, the controlled particle on top, the unsigned
wrap of pixels left of the image, and the pressure scale."""
import ctypes
import math
import subprocess

import numpy as np
import pytest

from demso_amd import api
from demso_amd import scene as sc
from demso_amd.scene import Instance, Material, Scene

ROOT = __import__("pathlib").Path(__file__).resolve().parents[1]


def scene_of(pos, radius=(0.5,), colors=((255.0, 0.0, 0.0),), types=None, box=(10.0, 10.0)):
    pos = np.asarray(pos, np.float32).reshape(-1, 2)
    n = pos.shape[0]
    mats = [Material(radius=r, mass=1.0, inertia=1.0, color=c) for r, c in zip(radius, colors)]
    t = np.zeros(n, np.uint32) if types is None else np.asarray(types, np.uint32)
    return Scene(name="img", pos=pos, vel=np.zeros((n, 2), np.float32), theta=np.zeros(n, np.float32),
                 omega=np.zeros(n, np.float32), id=np.arange(n, dtype=np.uint32), type=t,
                 loc=np.arange(n, dtype=np.uint32), radius=None, materials=mats,
                 instances=[Instance(box=box, extent=box, grid=(16, 16))], dt=1e-3)


def draw(oracle, s, rp, theta=None, pressure=None, ctl=None):
    th = s.theta if theta is None else np.asarray(theta, np.float32)
    pr = np.zeros(s.n, np.float32) if pressure is None else np.asarray(pressure, np.float32)
    return oracle.plot(s, rp, s.pos, th, s.type, pr, ctl=ctl)


def test_disc_coverage_and_shading(oracle):
    s = scene_of([(5.0, 5.0)], radius=(0.5,))
    rp = api.render_params(s, image_height=100, view_rotations=False)
    assert rp["image_dim"] == [100, 100]
    img = draw(oracle, s, rp)
    pr = np.float32(100) / np.float32(10) * np.float32(0.5)  # 5 px
    cx = cy = 50
    want = np.zeros((100, 100), bool)
    for x in range(-5, 5):
        for y in range(-5, 5):
            if x * x + y * y < pr * pr:
                want[cy + y, cx + x] = True
    assert np.array_equal(img[..., 3] > 0, want)
    assert img[cy, cx].tolist() == [255, 0, 0, 255]          # full colour at the centre
    assert img[cy, cx + 4, 0] == int(255 * math.sqrt((25 - 16) / 25))   # shaded towards the rim
    assert not img[~want].any()                                # black background elsewhere


def test_later_disc_covers_earlier(oracle):
    s = scene_of([(5.0, 5.0), (5.3, 5.0)], radius=(0.5, 0.5), colors=((255.0, 0, 0), (0, 255.0, 0)), types=[0, 1])
    rp = api.render_params(s, image_height=100, view_rotations=False)
    img = draw(oracle, s, rp)
    assert img[50, 52].tolist()[:3] == [0, int(255 * math.sqrt((25 - 1) / 25)), 0]   # overlap: particle 1
    assert img[50, 46, 0] > 0 and img[50, 46, 1] == 0                                 # only particle 0
    s2 = scene_of(s.pos[::-1].copy(), radius=(0.5, 0.5), colors=((255.0, 0, 0), (0, 255.0, 0)), types=[1, 0])
    img2 = draw(oracle, s2, rp)
    assert img2[50, 52, 0] > 0 and img2[50, 52, 1] == 0                                # order decides


@pytest.mark.parametrize("theta,step", [(0.0, (1, 0)), (math.pi / 2, (0, 1)), (math.pi, (-1, 0)),
                                        (3 * math.pi / 2, (0, -1))])
def test_rotation_line(oracle, theta, step):
    s = scene_of([(5.0, 5.0)], radius=(1.0,))
    rp = api.render_params(s, image_height=100)
    img = draw(oracle, s, rp, theta=[theta])
    for r in range(10):
        # the series give cos/sin 0 and +-1 exactly on the axes
        x, y = 50 + r * step[0], 50 + r * step[1]
        assert img[y, x].tolist() == [0, 0, 0, 0], (r, x, y)
    assert img[50 + 10 * step[1], 50 + 10 * step[0], 3] == 0  # the disc ends there anyway


def test_colorbar_values(oracle):
    """Pressure colouring: zero pressure is blue 127, the maximum dark red
    127, half the maximum green 255 with red and blue 127 (colorbar,
    src/particles_kernel.cuh:629)."""
    s = scene_of([(2.0, 5.0), (5.0, 5.0), (8.0, 5.0)], radius=(0.5,))
    rp = api.render_params(s, image_height=100, view_rotations=False, color_by_pressure=True, max_pressure=4.0)
    img = draw(oracle, s, rp, pressure=[0.0, 2.0, 4.0])
    assert img[50, 20].tolist() == [0, 0, 127, 255]
    assert img[50, 50].tolist() == [127, 255, 127, 255]
    assert img[50, 80].tolist() == [127, 0, 0, 255]


def test_controlled_particle_on_top_and_wraps_left(oracle):
    s = scene_of([(5.0, 5.0), (0.0, 2.0)], radius=(0.5, 1.0), colors=((255.0, 0, 0), (0, 0, 255.0)), types=[0, 0])
    ctl = oracle.control((5.2, 5.0), 1)
    rp = api.render_params(s, image_height=100, view_rotations=False, controlled_type=1)
    assert rp["control_dim"] == [20, 20]
    img = draw(oracle, s, rp, ctl=ctl)
    assert img[50, 52].tolist() == [0, 0, 255, 255]     # the controlled disc's centre, over particle 0
    # the disc at x = 0 reaches x = -1 .. -5: unsigned wrap onto the previous row's right end
    assert img[20 - 1, 99, 0] > 0 and img[20, 0, 0] > 0


def test_pressure_scale(oracle):
    assert oracle.pressure_scale([0.0, 3.0, 2.5]) == 3.0
    assert oracle.pressure_scale([0.0, 0.0]) == 1.0
    assert oracle.pressure_scale([-1.0, float("nan"), 0.5]) == 0.5
    assert oracle.pressure_scale([]) == 1.0


def test_render_params_match_demso(tmp_path):
    """imageDIMx = Lx/Ly * imageheight in float and the controlled square
    (src/main.cu:172-173, 209-215)."""
    s = sc.triangle_pile(controlled=True)
    rp = api.render_params(s, image_height=800, controlled_type=1)
    assert rp["image_dim"] == [800, 800] and rp["control_dim"] == [80, 80]
    s2 = sc.settling(side=16)
    rp2 = api.render_params(s2, image_height=333)
    lx, ly = (np.float32(v) for v in s2.instances[0].box)
    assert rp2["image_dim"][0] == int(lx / ly * np.float32(333))


def test_render_struct_layout(tmp_path):
    from demso_amd import _lib

    src = tmp_path / "r.c"
    src.write_text('#include <stdio.h>\n#include "demso.h"\nint main(void){ printf("%zu\\n", '
                   'sizeof(dem_render_params)); return 0; }\n')
    exe = tmp_path / "r"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    assert int(subprocess.run([str(exe)], capture_output=True, text=True).stdout) == ctypes.sizeof(_lib.RenderProps)
