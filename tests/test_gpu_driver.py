"""The headless driver on the GPU: a .dsml scene stepped through the
reference's loop (the five wrappers with their reference signatures) and
through graph replays, against each other and against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from dsml_scenes import MIXED, MIXED_BOX, python_scene, run_driver, write
from demso_amd import dumps

pytestmark = pytest.mark.gpu

STEPS = 100


def run_loop(tmp_path, loop: str):
    scene = write(tmp_path, "mixed.dsml", MIXED)
    out = tmp_path / f"out_{loop}.csv"
    prefix = tmp_path / f"{loop}_"
    print(f"driver: --loop {loop} --steps {STEPS}", flush=True)
    r = run_driver("--loop", loop, "--steps", STEPS, "--output", out, "--dump-every", 50, "--dump-prefix", prefix,
                   "--seed", 3, scene, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return out, prefix


@pytest.fixture(scope="module")
def runs(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("driver")
    scene = write(tmp, "mixed.dsml", MIXED)
    r = run_driver("--seed", 3, "--dump-initial", tmp / "initial.bin", scene)
    assert r.returncode == 0, r.stderr
    return tmp, {loop: run_loop(tmp, loop) for loop in ("stages", "graph")}


def test_stage_loop_equals_graph_loop(runs):
    tmp, res = runs
    for step in (50, 100):
        a = dumps.read_dump(f"{res['stages'][1]}{step}.bin")
        b = dumps.read_dump(f"{res['graph'][1]}{step}.bin")
        assert a["step"] == b["step"] == step
        for k in ("pos", "vel", "theta", "omega", "type", "acc", "alpha", "pressure"):
            assert np.array_equal(a[k], b[k]), k
    assert res["stages"][0].read_text() == res["graph"][0].read_text()


def test_driver_matches_oracle(runs, oracle):
    """The driver's state after 50 and 100 steps against the oracle run from
    the driver's own initial state (wrapping grid, attraction on one type)."""
    tmp, res = runs
    init = dumps.read_dump(tmp / "initial.bin")
    scene = python_scene(init, ["a", "b", "big"], MIXED_BOX, dt=1e-3, bkn=1000.0, bkt=100.0)
    sim = oracle.Sim(scene, "f32")
    done = 0
    for step in (50, 100):
        sim.step(step - done)
        done = step
        o = sim.state()
        ref = {}
        order = np.argsort(o["id"])
        for k in ("pos", "vel", "theta", "omega", "acc", "alpha", "pressure"):
            ref[k] = o[k][order]
        g = dumps.read_dump(f"{res['stages'][1]}{step}.bin")
        for k, v in ref.items():
            scale = max(float(np.abs(v).max()), 1e-30)
            err = float(np.abs(g[k].astype(np.float64) - v).max()) / scale
            print(f"step {step} {k}: max rel err {err:.2e}")
            assert err <= 1e-5, k
    sim.close()


def test_output_file(runs):
    """One line per step: the elapsed time (timeStep * step in float), then
    the followed particles (the two follow="true" singles, then <follow>
    IDs 5 and 40) with the state of the step's end."""
    tmp, res = runs
    times, rows = dumps.read_output(res["stages"][0])
    assert rows.shape == (STEPS, 4, 12)
    assert np.allclose(times, np.float32(1e-3) * np.arange(1, STEPS + 1, dtype=np.float32), atol=1e-6)
    assert rows[:, :, 0].astype(int).tolist() == [[0, 1, 5, 40]] * STEPS
    final = dumps.read_dump(f"{res['stages'][1]}{STEPS}.bin")
    for j, pid in enumerate((0, 1, 5, 40)):
        r = rows[-1, j]
        assert int(r[1]) == int(final["type"][pid])
        exp = np.concatenate([final["pos"][pid], final["vel"][pid], final["acc"][pid],
                              [final["theta"][pid], final["omega"][pid], final["alpha"][pid],
                               final["pressure"][pid]]])
        assert np.allclose(r[2:], exp, rtol=0, atol=5e-7 + 1e-6 * np.abs(exp)), (pid, r[2:], exp)


@pytest.mark.parametrize("loop", ["stages", "graph"])
def test_reference_pile_scene_matches_oracle(tmp_path, oracle, loop):
    """DemSo's own pilha.dsml as its default build runs it (controlled
    particle: 136 discs on the 16 x 16 wrapping grid, attraction 10, the MOVE
    disc driven down through the pile) for 1400 steps, through the
    reference's stage loop and through graph replays, against the oracle's
    run of the same scene (scene.triangle_pile(controlled=True)): state, the
    controlled particle and the last step's reaction on it, bit for bit."""
    from pathlib import Path

    from demso_amd import scene as sc
    from test_gpu_control import oracle_run

    steps = 1400
    pile = Path(__file__).resolve().parent / "golden" / "dsml" / "pilha.dsml"
    print(f"driver: pilha.dsml --controlled --loop {loop} --steps {steps}", flush=True)
    r = run_driver("--controlled", "--loop", loop, "--steps", steps, "--output", tmp_path / "out.csv", "--dump-final",
                   "--dump-prefix", tmp_path / "pile_", pile, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    g = dumps.read_dump(tmp_path / f"pile_{steps}.bin")
    s = sc.triangle_pile(controlled=True)
    o, cs, red = oracle_run(oracle, s, s.notes["controlled"], steps)
    for k in ("pos", "vel", "theta", "omega", "acc", "alpha", "pressure"):
        assert np.array_equal(g[k].view(np.uint32), o[k].view(np.uint32)), k
    line = [l for l in r.stdout.splitlines() if l.startswith("controlled particle:")]
    assert len(line) == 1, r.stdout
    w = line[0].split()
    x, y, fx, fy, m = (float.fromhex(v) for v in (w[3], w[4], w[6], w[7], w[8]))
    assert [x, y] == list(cs.pos) and cs.pos[1] < 3.0
    assert [fx, fy, m] == [float(np.float32(v)) for v in red] and max(map(abs, red)) > 0
    # pilha.dsml follows IDs 500 and 2200 of 136 discs: reported, nothing written
    assert "followed particle 500 does not exist" in r.stderr
    assert (tmp_path / "out.csv").read_text() == ""


def read_ppm(path):
    data = path.read_bytes()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255"
    w, h = map(int, parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)


@pytest.mark.parametrize("loop", ["stages", "graph"])
def test_driver_frames_match_oracle(tmp_path, oracle, loop):
    """--image-every: pilha.dsml's frames (800 x 800, controlled particle,
    pressure colours) from both loops equal the oracle's image of its own
    run, rows flipped to top-first."""
    from pathlib import Path

    from demso_amd import api
    from demso_amd import scene as sc

    steps = 1300
    pile = Path(__file__).resolve().parent / "golden" / "dsml" / "pilha.dsml"
    print(f"driver frames: pilha.dsml --loop {loop}", flush=True)
    r = run_driver("--controlled", "--loop", loop, "--steps", steps, "--output", "", "--image-every", 650,
                   "--color-by-pressure", "--image-prefix", tmp_path / "f_", pile, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_ppm(tmp_path / f"f_{steps}.ppm")
    assert (tmp_path / "f_650.ppm").exists()
    s = sc.triangle_pile(controlled=True)
    s.materials[0].color = (255.0, 255.0, 255.0)   # "ffffff"
    s.materials[1].color = (58.0, 153.0, 51.0)     # "3a9933"
    # the frame is drawn from the state in its sorted order
    ref = oracle.Sim(s, "f32")
    c = s.notes["controlled"]
    ref.set_control(oracle.control(c["pos"], c["type"], move=c["move"], wrap_below_y=c["wrap_below_y"],
                                   wrap_above_x=c["wrap_above_x"], reset_y=c["reset_y"], reset_x=c["reset_x"]))
    ref.step(steps)
    st = ref.state()
    cs, _ = ref.control()
    ref.close()
    rp = api.render_params(s, image_height=800, controlled_type=1, color_by_pressure=True,
                           max_pressure=oracle.pressure_scale(st["pressure"]))
    want = oracle.plot(s, rp, st["pos"], st["theta"], st["type"], st["pressure"], ctl=cs)
    assert np.array_equal(got, want[::-1, :, :3])
