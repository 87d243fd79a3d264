"""Small .dsml scenes written by the tests, and the matching Python scenes."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "demso_amd" / "demso_run"

TYPES = {
    # id: (mass, radius, kn, kt, damping, boundary damping, friction, attraction)
    "a": (1e-2, 0.05, 1000.0, 1000.0, 0.2, 0.5, 0.4, 0.0),
    "b": (8e-3, 0.04, 1500.0, 800.0, 0.3, 0.6, 0.5, 5.0),
    "big": (5e-2, 0.1, 2000.0, 1000.0, 0.2, 0.5, 0.3, 0.0),
}


def particletype(tid: str, color: str = "ffffff") -> str:
    m, r, kn, kt, c, bd, mu, at = TYPES[tid]
    return f"""    <particletype id="{tid}" name="type {tid}" color="{color}">
      <mass>{m!r}</mass>
      <radius>{r!r}</radius>
      <stiffness dir="normal">{kn!r}</stiffness>
      <stiffness dir="shear">{kt!r}</stiffness>
      <damping>{c!r}</damping>
      <damping type="boundary">{bd!r}</damping>
      <friction>{mu!r}</friction>
      <attraction>{at!r}</attraction>
    </particletype>
"""


MIXED = """<?xml version="1.0"?>
<simulation xmlns="http://fem.unicamp.br/ns/dsml">
  <parameters>
    <timestep>1e-3</timestep>
    <simduration>0.1</simduration>
    <follow>5</follow>
    <follow>40</follow>
  </parameters>
  <environment>
    <dimensions><x>4</x><y>4</y></dimensions>
    <gravity><y>-9.81</y></gravity>
    <stiffness dir="normal">1000</stiffness>
    <stiffness dir="shear">100</stiffness>
  </environment>
  <properties>
""" + particletype("a", "ff0000") + particletype("b", "00ff00") + particletype("big", "0000ff") + """  </properties>
  <particles>
    <!-- two free discs, followed -->
    <particle follow="true" particletype="big">
      <pos><x>2.0</x><y>2.5</y></pos>
      <vel><x>1.5</x><y>-0.5</y></vel>
      <omega>3.0</omega>
    </particle>
    <particle follow="true" particletype="a">
      <pos><x>0.3</x><y>3.0</y></pos>
      <theta>0.5</theta>
    </particle>
    <rectangle particletype="a,b">
      <start><x>0.5</x><y>0.5</y></start>
      <end><x>1.6</x><y>1.2</y></end>
      <num><x>10</x><y>6</y></num>
    </rectangle>
    <triangle particletype="b,a">
      <pos><x>3.0</x><y>0.1</y></pos>
      <width>1.0</width>
      <num>10</num>
    </triangle>
  </particles>
</simulation>
"""

MIXED_BOX = (4.0, 4.0)
MIXED_COUNTS = (2, 60, 55)  # singles, rectangle, triangle


def write(tmp_path: Path, name: str, text: str) -> Path:
    p = tmp_path / name
    p.write_text(text)
    return p


def run_driver(*args, cwd=None, timeout=120) -> subprocess.CompletedProcess:
    if not DRIVER.exists():
        raise FileNotFoundError(f"{DRIVER} is not built (demso_amd.build.build_driver)")
    return subprocess.run([str(DRIVER), *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=timeout)


def python_scene(dump: dict, type_ids, box, dt: float, bkn: float, bkt: float, gravity=(0.0, -9.81)):
    """The Python scene equivalent of a prepared .dsml scene: the driver's
    initial state (from its --dump-initial file) and the type table."""
    from demso_amd import scene as sc

    f32 = np.float32
    mats = []
    for tid in type_ids:
        m, r, kn, kt, c, bd, mu, at = TYPES[tid]
        inertia = float(f32(2) * f32(m) * f32(r) * f32(r) / f32(5))
        mats.append(sc.Material(radius=r, mass=m, inertia=inertia, kn=kn, kt=kt, cn=c, bdamp=bd, mu=mu, attract=at))
    rmax = max(TYPES[t][1] for t in type_ids)
    grid = sc.reference_grid_size(box, rmax)
    n = dump["n"]
    ids = np.arange(n, dtype=np.uint32)
    return sc.Scene(name="dsml", pos=np.ascontiguousarray(dump["pos"]), vel=np.ascontiguousarray(dump["vel"]),
                    theta=np.ascontiguousarray(dump["theta"]), omega=np.ascontiguousarray(dump["omega"]), id=ids,
                    type=np.ascontiguousarray(dump["type"]), loc=ids.copy(), radius=None, materials=mats,
                    instances=[sc.Instance(box=box, extent=box, grid=grid)], dt=dt, gravity=gravity, bkn=bkn,
                    bkt=bkt, flags=0)


def big_blocks(rect=(300, 200), rows=400) -> str:
    """MIXED with a large rectangle, a second rectangle (draw streams 1002 and
    1003) and a large triangle: tens of thousands of generator draws."""
    text = MIXED.replace("<num><x>10</x><y>6</y></num>", f"<num><x>{rect[0]}</x><y>{rect[1]}</y></num>")
    text = text.replace("<num>10</num>", f"<num>{rows}</num>")
    second = """    <rectangle particletype="big,a,b">
      <start><x>2.0</x><y>1.5</y></start>
      <end><x>3.5</x><y>3.5</y></end>
      <num><x>37</x><y>53</y></num>
    </rectangle>
"""
    return text.replace("    <triangle", second + "    <triangle", 1)


BIG_BLOCKS_COUNTS = (2, 300 * 200, 37 * 53, 400 * 401 // 2)
