"""The scene generators on the device (createRectangles, createTriangles,
createSingleParticles, src/functions.cu:198-311) against their numpy
restatements in demso_amd.scene, bit for bit, at full-workload sizes, and
the driver's device-created initial state against its host one."""
import math

import numpy as np
import pytest

from demso_amd import api, dumps
from demso_amd import scene as sc
from dsml_scenes import big_blocks, run_driver, write
from helpers import dev, host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gen_ctx():
    """A context without parameters: the generators then accept any type
    index (with parameters they check them against the material table)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    ctx = api.Context(0)
    yield ctx
    ctx.close()


def zeros(n, cols=None, dt=np.float32):
    return dev(np.zeros((n, cols) if cols else n, dt))


@pytest.mark.parametrize("side,seed,stream", [((1500, 1000), 3, None), ((2048, 2048), 12345678901, 1002),
                                              ((7, 1), 1, 1000)])
def test_create_rectangles(gen_ctx, side, seed, stream):
    n = side[0] * side[1]
    start, end = (0.25, 0.5), (2.25, 1.5)
    types = [3, 1, 4]
    pos, ID, loc, typ = zeros(n, 2), zeros(n, dt=np.uint32), zeros(n, dt=np.uint32), zeros(n, dt=np.uint32)
    print(f"createRectangles {side} seed={seed} stream={stream}", flush=True)
    api.createRectangles(gen_ctx, pos, ID, loc, typ, start, (end[0] - start[0], end[1] - start[1]), side, 17, types,
                         seed, stream)
    want_pos, want_type = sc.rectangle_layout(start, end, side, types, seed=seed,
                                              stream=1000 if stream is None else stream)
    with np.errstate(invalid="ignore"):
        assert np.array_equal(host(pos).view(np.uint32), want_pos.view(np.uint32))
    assert np.array_equal(host(typ, np.uint32), want_type)
    ids = np.arange(17, 17 + n, dtype=np.uint32)
    assert np.array_equal(host(ID, np.uint32), ids) and np.array_equal(host(loc, np.uint32), ids)


@pytest.mark.parametrize("N,count", [(2000, None), (16, None), (300, 1000), (1, 1)])
def test_create_triangles(gen_ctx, N, count):
    f32 = np.float32
    full = N * (N + 1) // 2
    count = full if count is None else count
    space = f32(3.75) / f32(max(N - 1, 1))
    height = f32(float(space) * math.sqrt(3.0) / 2.0)
    pos, ID, loc, typ = zeros(count, 2), zeros(count, dt=np.uint32), zeros(count, dt=np.uint32), zeros(count,
                                                                                                       dt=np.uint32)
    print(f"createTriangles N={N} count={count}", flush=True)
    api.createTriangles(gen_ctx, pos, ID, loc, typ, (5.0, 0.125), N, [0, 1], float(space), float(height), 5, count)
    want_pos, want_type = sc.triangle_layout((5.0, 0.125), N, space, height, [0, 1], count=count)
    assert np.array_equal(host(pos).view(np.uint32), want_pos.view(np.uint32))
    assert np.array_equal(host(typ, np.uint32), want_type)
    assert np.array_equal(host(ID, np.uint32), np.arange(5, 5 + count, dtype=np.uint32))
    if N == 16:
        assert np.array_equal(host(pos), sc.triangle_pile().pos)


def test_create_triangles_rejects_too_many(gen_ctx):
    buf = zeros(64, 2)
    ids = zeros(64, dt=np.uint32)
    with pytest.raises(api.DemError):
        api.createTriangles(gen_ctx, buf, ids, ids, ids, (0.0, 0.0), 4, [0], 0.1, 0.1, 0, 11)  # 4 rows hold 10


def test_create_single_particles(gen_ctx):
    rng = np.random.default_rng(0)
    n = 1000
    p = rng.normal(size=(n, 2)).astype(np.float32)
    v = rng.normal(size=(n, 2)).astype(np.float32)
    th = rng.normal(size=n).astype(np.float32)
    w = rng.normal(size=n).astype(np.float32)
    t = rng.integers(0, 1, size=n).astype(np.uint32)
    for src in ("host", "device"):
        bufs = [zeros(n, 2), zeros(n, 2), zeros(n), zeros(n), zeros(n, dt=np.uint32), zeros(n, dt=np.uint32),
                zeros(n, dt=np.uint32)]
        args = (p, v, th, w, t) if src == "host" else tuple(dev(a) for a in (p, v, th, w, t))
        api.createSingleParticles(gen_ctx, *bufs, *args, n, 100)
        for got, want in zip(bufs[:4], (p, v, th, w)):
            assert np.array_equal(host(got), want)
        assert np.array_equal(host(bufs[6], np.uint32), t)
        assert np.array_equal(host(bufs[4], np.uint32), np.arange(100, 100 + n, dtype=np.uint32))


@pytest.mark.parametrize("loop", ["stages", "graph"])
def test_driver_creates_on_device_as_on_host(tmp_path, loop):
    """The driver's state created on the device (both loops, dumped before the
    first step) equals its host preparation (--dump-initial) bit for bit."""
    scene = write(tmp_path, "big.dsml", big_blocks())
    r = run_driver("--seed", 11, "--dump-initial", tmp_path / "host.bin", scene)
    assert r.returncode == 0, r.stderr
    print(f"driver: big blocks --loop {loop} --steps 0", flush=True)
    r = run_driver("--seed", 11, "--loop", loop, "--steps", 0, "--dump-final", "--dump-prefix", tmp_path / "dev_",
                   "--output", "", scene, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    a, b = dumps.read_dump(tmp_path / "host.bin"), dumps.read_dump(tmp_path / "dev_0.bin")
    for k in ("pos", "vel", "theta", "omega", "type"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
