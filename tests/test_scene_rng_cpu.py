"""The scene generators' random draws and block layouts on the host.

The rectangle jitter replaces curand_normal (time(NULL)-seeded in the
reference, src/main.cu:291) by Box-Muller normals built from correctly
rounded operations only, so the host driver (C++), numpy and the gfx950
kernels agree bit for bit.  Pinned here against the mathematics (numpy's
log and cos to a few ulp, the normal's moments) and the C++ host driver
against numpy over tens of thousands of draws.
"""
import math

import numpy as np
import pytest

from demso_amd import dumps
from demso_amd import scene as sc
from dsml_scenes import BIG_BLOCKS_COUNTS, big_blocks, run_driver, write


def test_series_cos_matches_libm():
    u = np.concatenate([sc.uniform(3, 11, 400_000),
                        [0.0, 2.0 ** -53, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1 - 2.0 ** -53]])
    got = sc._series_cos2pi(u)
    assert np.abs(got - np.cos(2.0 * np.pi * u)).max() < 2e-15
    # exact at the reduction points
    assert sc._series_cos2pi(np.array([0.0, 0.25, 0.5, 0.75])).tolist() == [1.0, 0.0, -1.0, 0.0]


def test_series_log_matches_libm():
    x = np.concatenate([1.0 - sc.uniform(5, 12, 400_000), [1.0, 0.5, 2.0 ** -53, 0.7071067811865476,
                                                            0.7071067811865475]])
    got = sc._series_log(x)
    ref = np.log(x)
    assert np.all(np.abs(got - ref) <= 4 * np.spacing(np.maximum(np.abs(ref), 1e-300)))
    assert sc._series_log(np.array([1.0]))[0] == 0.0
    assert sc._series_log(np.array([2.0 ** -53]))[0] == pytest.approx(-53 * math.log(2), rel=1e-16)


def test_jitter_normal_moments():
    g = sc.jitter_normal(1, 1000, 2_000_000)
    assert abs(g.mean()) < 3e-3 and abs(g.std() - 1.0) < 3e-3
    # fourth moment 3 and symmetric tails
    assert abs((g ** 4).mean() - 3.0) < 0.03
    assert abs((g > 2).mean() - (g < -2).mean()) < 1e-3
    # the draws differ per stream and per seed
    assert not np.array_equal(g[:100], sc.jitter_normal(1, 1001, 100))
    assert not np.array_equal(g[:100], sc.jitter_normal(2, 1000, 100))


def test_host_driver_blocks_equal_numpy(tmp_path):
    """The C++ host generator (--dump-initial) equals scene.rectangle_layout
    and scene.triangle_layout bit for bit over 103,000 particles."""
    scene = write(tmp_path, "big.dsml", big_blocks())
    out = tmp_path / "init.bin"
    r = run_driver("--seed", 11, "--dump-initial", out, scene)
    assert r.returncode == 0, r.stderr
    d = dumps.read_dump(out)
    ns, nr, nr2, nt = BIG_BLOCKS_COUNTS
    assert d["n"] == ns + nr + nr2 + nt
    at = ns
    pos, types = sc.rectangle_layout((0.5, 0.5), (1.6, 1.2), (300, 200), [0, 1], seed=11, stream=1000)
    assert np.array_equal(d["pos"][at:at + nr], pos) and np.array_equal(d["type"][at:at + nr], types)
    at += nr
    pos, types = sc.rectangle_layout((2.0, 1.5), (3.5, 3.5), (37, 53), [2, 0, 1], seed=11, stream=1002)
    assert np.array_equal(d["pos"][at:at + nr2], pos) and np.array_equal(d["type"][at:at + nr2], types)
    at += nr2
    f32 = np.float32
    space = f32(1.0) / f32(399)
    height = f32(float(space) * math.sqrt(3.0) / 2.0)
    pos, types = sc.triangle_layout((3.0, 0.1), 400, space, height, [1, 0])
    assert np.array_equal(d["pos"][at:], pos) and np.array_equal(d["type"][at:], types)
