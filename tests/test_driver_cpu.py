"""The .dsml reader and PrepareSim of the headless driver (no GPU needed:
--print-config and --dump-initial stop before any device call)."""
from __future__ import annotations

import math
from pathlib import Path

import numpy as np
import pytest

from dsml_scenes import MIXED, MIXED_BOX, MIXED_COUNTS, TYPES, run_driver, write
from demso_amd import dumps
from demso_amd import scene as sc

EXAMPLES = Path(__file__).resolve().parent / "golden" / "dsml"


def initial(tmp_path, text, *extra):
    scene = write(tmp_path, "scene.dsml", text)
    out = tmp_path / "initial.bin"
    r = run_driver(*extra, "--dump-initial", out, scene)
    assert r.returncode == 0, r.stdout + r.stderr
    return dumps.read_dump(out), r.stdout


@pytest.mark.skipif(not EXAMPLES.exists(), reason="DemSo example scenes not present")
@pytest.mark.parametrize("name,plain,controlled", [
    ("default", None, None),
    ("default2", None, None),
    ("pilha", None, None),
    ("singleparticle", None, "Nonexistant particle id: 'MOVE'"),
    ("colisao-elastica", None, "Controlled particle tag missing"),
    ("rolamento-sobre-plano", None, "Controlled particle tag missing"),
])
def test_reference_examples_load(name, plain, controlled):
    """Every example scene of the reference loads as the reference's parser
    reads it; with the controlled-particle build (the reference's default)
    the scenes without a usable <controlled> tag fail with its messages."""
    path = EXAMPLES / f"{name}.dsml"
    for flags, expect in (((), plain), (("--controlled",), controlled)):
        r = run_driver(*flags, "--print-config", path)
        if expect is None:
            assert r.returncode == 0, r.stderr
            assert "-- Properties" in r.stdout
        else:
            assert r.returncode == 1
            assert f"Parsing Error: {expect}" in r.stderr


@pytest.mark.skipif(not EXAMPLES.exists(), reason="DemSo example scenes not present")
def test_reference_default_scene_contents():
    r = run_driver("--print-config", EXAMPLES / "default.dsml")
    out = r.stdout
    assert "timeStep: 0.001" in out and "simDuration: 7.5" in out
    assert "numParticleTypes: 6" in out
    assert "numRectangles: 1" in out and "num: (40, 40)" in out
    assert "types: [0] = branca [3] = vermelha" in out
    assert "numTriangles: 1" in out and "types: [4] = amarela [1] = azul" in out
    assert "numSingleParticles: 1" in out and "omega: 3.1415" in out
    assert "color: (58, 153, 51)" in out  # 3a9933


@pytest.mark.skipif(not EXAMPLES.exists(), reason="DemSo example scenes not present")
def test_reference_pile_matches_python_scene(tmp_path):
    """pilha.dsml read as the controlled-particle build: the 136-disc pile of
    createTrianglesD on the 16 x 16 wrapping grid (the MOVE type's radius is
    left out of the grid sizing) equals scene.triangle_pile bit for bit."""
    out = tmp_path / "pile.bin"
    r = run_driver("--controlled", "--dump-initial", out, EXAMPLES / "pilha.dsml")
    assert r.returncode == 0, r.stderr
    assert "grid 16 x 16" in r.stdout
    d = dumps.read_dump(out)
    ref = sc.triangle_pile()
    assert d["n"] == ref.n == 136
    assert np.array_equal(d["pos"], ref.pos)
    assert np.array_equal(d["type"], np.zeros(136, np.uint32))
    # without the controlled-particle build the MOVE radius (0.5) sizes the grid
    r = run_driver("--dump-initial", out, EXAMPLES / "pilha.dsml")
    assert "grid 4 x 4" in r.stdout


def test_mixed_scene_layout(tmp_path):
    """Singles first, then the rectangle, then the triangle (src/main.cu:249-324),
    IDs in that order, the types of each block by (x+y) mod k and (C+L) mod k."""
    d, out = initial(tmp_path, MIXED, "--seed", "7")
    ns, nr, nt = MIXED_COUNTS
    assert d["n"] == ns + nr + nt
    assert "grid 8 x 8" in out
    # singles: big at (2, 2.5) with v (1.5, -0.5), w 3; 'a' at (0.3, 3) with theta 0.5
    assert d["pos"][0].tolist() == [2.0, 2.5] and d["vel"][0].tolist() == [1.5, -0.5]
    assert d["omega"][0] == np.float32(3.0) and d["theta"][1] == np.float32(0.5)
    assert d["type"][:2].tolist() == [2, 0]
    # rectangle: 10 x 6, types a,b alternating with x+y
    f32 = np.float32
    x = np.arange(nr) % 10
    y = np.arange(nr) // 10
    assert d["type"][ns:ns + nr].tolist() == [[0, 1][(i + j) % 2] for i, j in zip(x, y)]
    dx = f32(f32(1.6) - f32(0.5)) / f32(9)
    dy = f32(f32(1.2) - f32(0.5)) / f32(5)
    nx = sc.jitter_normal(7, 1000, nr).astype(np.float32)
    ny = sc.jitter_normal(7, 1001, nr).astype(np.float32)
    ex = f32(0.5) + dx * (x.astype(f32) + (nx - f32(0.5)) / f32(100))
    ey = f32(0.5) + dy * (y.astype(f32) + (ny - f32(0.5)) / f32(100))
    got = d["pos"][ns:ns + nr]
    assert np.array_equal(got[:, 0], ex) and np.array_equal(got[:, 1], ey)
    pos, types = sc.rectangle_layout((0.5, 0.5), (1.6, 1.2), (10, 6), [0, 1], seed=7, stream=1000)
    assert np.array_equal(pos, got) and np.array_equal(types, d["type"][ns:ns + nr])
    # triangle: 10 rows, base centred on (3, 0.1), rows 1/(N-1)*sqrt(3)/2 apart
    tri = d["pos"][ns + nr:]
    ref = sc.triangle_pile(rows=10, start=(3.0, 0.1), width=1.0)
    assert np.array_equal(tri, ref.pos)
    L = np.round((tri[:, 1] - f32(0.1)) / (f32(1.0 / 9) * f32(math.sqrt(3) / 2))).astype(int)
    C = np.round((tri[:, 0] - f32(3.0)) / f32(1.0 / 9) + (9 - L) / 2).astype(int) // 1
    # (C + L) mod 2 picks b, a; recover C from x = 3 + (L + 2C - 9) * space / 2
    C = np.round(((tri[:, 0] - 3.0) / (1.0 / 9) * 2 - L + 9) / 2).astype(int)
    assert d["type"][ns + nr:].tolist() == [[1, 0][(c + l) % 2] for c, l in zip(C, L)]
    assert np.array_equal(d["acc"], np.zeros((d["n"], 2), np.float32))


@pytest.mark.parametrize("box,rmax", [((10.0, 10.0), 0.125), ((10.0, 10.0), 0.04), ((4.0, 4.0), 0.1),
                                      ((1.0, 0.5), 0.0055), ((2.2, 1.1), 0.01), ((5.0, 5.0), 0.2)])
def test_grid_sizing_matches_python(tmp_path, box, rmax):
    text = MIXED.replace("<dimensions><x>4</x><y>4</y></dimensions>",
                         f"<dimensions><x>{box[0]}</x><y>{box[1]}</y></dimensions>")
    # one type of radius rmax, one free disc inside the box
    text = text.replace("<radius>0.1</radius>", f"<radius>{rmax}</radius>")
    text = text.replace("<radius>0.05</radius>", f"<radius>{rmax}</radius>")
    text = text.replace("<radius>0.04</radius>", f"<radius>{rmax}</radius>")
    text = text.replace("<follow>40</follow>", "")
    d, out = initial(tmp_path, text)
    gx, gy = sc.reference_grid_size(box, rmax)
    assert f"grid {gx} x {gy}" in out


ERRORS = [
    ("<timestep>1e-3</timestep>", "", "Simulation timestep not specified"),
    ('<stiffness dir="shear">100</stiffness>', '<stiffness dir="diagonal">100</stiffness>',
     "Unrecognized stiffness direction"),
    ('<stiffness dir="normal">1000</stiffness>\n    <stiffness dir="shear">100</stiffness>',
     '<stiffness>1000</stiffness>', "Stiffness direction not specified"),
    ('<damping type="boundary">0.5</damping>', '<damping type="wall">0.5</damping>', "Unrecognized damping type"),
    ("<friction>0.4</friction>", "<friction>0.4</friction><colour>1</colour>",
     "Unrecognized tag inside <particletype>"),
    ('<rectangle particletype="a,b">', '<rectangle particletype="a,zz">', "Nonexistant particle id: 'zz'"),
    ('<rectangle particletype="a,b">', '<rectangle particletype="a,">', "Nonexistant particle id: ''"),
    ("<start><x>0.5</x><y>0.5</y></start>", "", "<start> tag not found inside rectangle"),
    ("<num><x>10</x><y>6</y></num>", "", "<num> tag not found inside rectangle"),
    ("<width>1.0</width>", "", "<width> tag not found inside triangle"),
    ("<pos><x>3.0</x><y>0.1</y></pos>", "", "<pos> tag not found inside triangle"),
    ("<pos><x>2.0</x><y>2.5</y></pos>", "", "Missing <pos> tag on particle definition"),
    ("</particles>", "</particle>", "XML error at line"),
    ("<environment>", "<environment x=1>", "XML error at line"),
]


@pytest.mark.parametrize("old,new,message", ERRORS)
def test_parse_errors(tmp_path, old, new, message):
    assert old in MIXED
    scene = write(tmp_path, "bad.dsml", MIXED.replace(old, new, 1))
    r = run_driver("--print-config", scene)
    assert r.returncode == 1
    assert r.stderr.startswith("Parsing Error: ") and message in r.stderr, r.stderr


def test_no_types_and_missing_sections(tmp_path):
    a = MIXED.index("<properties>") + len("<properties>")
    b = MIXED.index("</properties>")
    r = run_driver("--print-config", write(tmp_path, "a.dsml", MIXED[:a] + MIXED[b:]))
    assert "Parsing Error: No particle types defined, aborting." in r.stderr
    r = run_driver("--print-config", write(tmp_path, "b.dsml", MIXED.replace("<environment>", "<environmentX>")
                                           .replace("</environment>", "</environmentX>")))
    assert "Parsing Error: <environment> section missing" in r.stderr


def test_default_types_and_entities(tmp_path):
    """A block without particletype= takes the first type (with the
    reference's note on stdout); attribute values decode entities."""
    text = MIXED.replace('<rectangle particletype="a,b">', "<rectangle>")
    text = text.replace('name="type a"', 'name="type &lt;a&gt; &#233;"')
    r = run_driver("--print-config", write(tmp_path, "d.dsml", text))
    assert r.returncode == 0, r.stderr
    assert "Particle type not specified for rectangle, defaulting to first one: a" in r.stdout
    assert "name: type <a> é" in r.stdout


def test_missing_followed_particle_is_skipped(tmp_path):
    """A <follow> ID beyond the particle count (pilha.dsml follows 500 and
    2200 of 136 discs; the reference reads past its arrays) is reported and
    left out of the output."""
    scene = write(tmp_path, "f.dsml", MIXED.replace("<follow>40</follow>", "<follow>117</follow>"))
    r = run_driver("--dump-initial", tmp_path / "x.bin", scene)
    assert r.returncode == 0
    assert "warning: followed particle 117 does not exist (117 particles)" in r.stderr


def test_steps_from_simduration_or_flag(tmp_path):
    """--steps is required when the scene has no <simduration>; the check
    runs before any device call."""
    text = MIXED.replace("<simduration>0.1</simduration>", "")
    r = run_driver("--output", "", write(tmp_path, "s.dsml", text))
    assert r.returncode == 2 and "pass --steps" in r.stderr
