"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The CPU restatement of the reference kernels (dem_oracle.hpp).  Imported by
tests, by the smoke check and by the benchmark's CPU-baseline leg, never by
the product package.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_void_p
from pathlib import Path
from typing import Optional

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class OrcMaterial(ctypes.Structure):
    _fields_ = [(f, c_float) for f in ("radius", "mass", "inertia", "kn", "kt", "cn", "bdamp", "mu", "attract",
                                       "areal")] + [("inst", c_uint32), ("pad", c_uint32)]


class OrcInstance(ctypes.Structure):
    _fields_ = [("box", c_float * 2), ("extent", c_float * 2), ("grid", c_uint32 * 2), ("cell_base", c_uint32),
                ("pad", c_uint32)]


class OrcSystem(ctypes.Structure):
    _fields_ = [("dt", c_float), ("g", c_float * 2), ("bkn", c_float), ("bkt", c_float), ("flags", c_uint32),
                ("num_types", c_uint32), ("num_instances", c_uint32), ("mat", c_void_p), ("inst", c_void_p)]


P = c_void_p
_lib = None


def build() -> Path:
    res = subprocess.run(["make", "-C", str(HERE), "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    lib = ctypes.CDLL(str(LIB_PATH))
    proto = {
        "orc_abi_version": (c_int, []),
        "orc_set_threads": (c_int, [c_int]),
        "orc_sizeof_system": (c_int, []),
        "orc_sizeof_material": (c_int, []),
        "orc_sizeof_instance": (c_int, []),
        "orc_calc_hash": (c_int, [POINTER(OrcSystem), P, P, P, P, c_uint32]),
        "orc_sort_particles": (c_int, [P, P, c_uint32]),
        "orc_reorder": (c_int, [P] * 19 + [c_uint32, c_uint32]),
        "orc_collide": (c_int, [POINTER(OrcSystem)] + [P] * 9 + [c_uint32, P]),
        "orc_integrate": (c_int, [POINTER(OrcSystem)] + [P] * 9 + [c_uint32]),
        "orc_sim_create_f32": (c_void_p, [POINTER(OrcSystem), c_uint32, c_uint32, c_uint32, c_uint32, c_int]),
        "orc_sim_create_f64": (c_void_p, [POINTER(OrcSystem), c_uint32, c_uint32, c_uint32, c_uint32, c_int]),
        "orc_sim_destroy_f32": (None, [c_void_p]),
        "orc_sim_destroy_f64": (None, [c_void_p]),
        "orc_sim_set_state_f32": (c_int, [c_void_p] + [P] * 8),
        "orc_sim_set_state_f64": (c_int, [c_void_p] + [P] * 8),
        "orc_sim_step_f32": (c_double, [c_void_p, c_uint32]),
        "orc_sim_step_f64": (c_double, [c_void_p, c_uint32]),
        "orc_sim_set_control_f32": (c_int, [c_void_p, c_void_p]),
        "orc_sim_set_control_f64": (c_int, [c_void_p, c_void_p]),
        "orc_sim_get_control_f32": (c_int, [c_void_p, c_void_p, P]),
        "orc_sim_get_control_f64": (c_int, [c_void_p, c_void_p, P]),
        "orc_collide_controlled": (c_int, [c_void_p] + [P] * 9 + [c_uint32, P, c_void_p, P, P, P, P]),
        "orc_sim_get_state_f32": (c_int, [c_void_p] + [P] * 15),
        "orc_sim_get_state_f64": (c_int, [c_void_p] + [P] * 15),
        "orc_adj_create": (c_void_p, [POINTER(OrcSystem), c_uint32, c_uint32, c_uint32, c_uint32, c_int]),
        "orc_adj_destroy": (None, [c_void_p]),
        "orc_adj_set_state": (c_int, [c_void_p] + [P] * 8),
        "orc_adj_forward": (c_int, [c_void_p, c_uint32]),
        "orc_adj_final_state": (c_int, [c_void_p, P, P]),
        "orc_adj_backward": (c_int, [c_void_p] + [P] * 5),
        "orc_plot_particles": (None, [P] * 5 + [c_uint32, P, c_uint32, P, c_void_p, P]),
        "orc_pressure_scale": (ctypes.c_float, [P, c_uint32]),
    }
    for name, (res, args) in proto.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    assert lib.orc_sizeof_system() == ctypes.sizeof(OrcSystem)
    assert lib.orc_sizeof_material() == ctypes.sizeof(OrcMaterial)
    assert lib.orc_sizeof_instance() == ctypes.sizeof(OrcInstance)
    _lib = lib
    return lib


def set_threads(t: int) -> int:
    return load().orc_set_threads(t)


def _p(a: Optional[np.ndarray]) -> Optional[int]:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class System:
    """orc::System built from a demso_amd.scene.Scene (keeps its tables alive)."""

    def __init__(self, scene):
        self.mat = (OrcMaterial * len(scene.materials))()
        for k, m in enumerate(scene.materials):
            r = self.mat[k]
            r.radius, r.mass, r.inertia, r.kn, r.kt, r.cn = m.radius, m.mass, m.inertia, m.kn, m.kt, m.cn
            r.bdamp, r.mu, r.attract, r.areal, r.inst = m.bdamp, m.mu, m.attract, m.areal, m.instance
        self.inst = (OrcInstance * len(scene.instances))()
        for k, g in enumerate(scene.instances):
            self.inst[k].box[:] = g.box
            self.inst[k].extent[:] = g.extent
            self.inst[k].grid[:] = g.grid
            self.inst[k].cell_base = g.cell_base
        s = OrcSystem()
        s.dt = scene.dt
        s.g[:] = scene.gravity
        s.bkn, s.bkt, s.flags = scene.bkn, scene.bkt, scene.flags
        s.num_types, s.num_instances = len(scene.materials), len(scene.instances)
        s.mat = ctypes.cast(self.mat, c_void_p).value
        s.inst = ctypes.cast(self.inst, c_void_p).value
        self.sys = s
        self.total_cells = scene.total_cells
        self.multi = len(scene.instances) > 1


# ---- single stages (FP32) -------------------------------------------------------

def calc_hash(system: System, pos: np.ndarray, type: Optional[np.ndarray] = None):
    lib = load()
    n = pos.shape[0]
    h = np.empty(n, np.uint32)
    i = np.empty(n, np.uint32)
    lib.orc_calc_hash(ctypes.byref(system.sys), _p(pos), _p(type), _p(h), _p(i), n)
    return h, i


def sort_particles(hash_: np.ndarray, index: np.ndarray):
    h, i = hash_.copy(), index.copy()
    load().orc_sort_particles(_p(h), _p(i), h.shape[0])
    return h, i


def reorder(hash_, index, old: dict, num_cells: int, cell_end_init: Optional[np.ndarray] = None):
    """Returns (cellStart, cellEnd, sorted) with sorted a dict of arrays."""
    lib = load()
    n = hash_.shape[0]
    cs = np.empty(num_cells, np.uint32)
    ce = np.zeros(num_cells, np.uint32) if cell_end_init is None else cell_end_init.copy()
    rad = old.get("radius")
    s = {
        "pos": np.empty((n, 2), np.float32), "vel": np.empty((n, 2), np.float32),
        "theta": np.empty(n, np.float32), "omega": np.empty(n, np.float32), "id": np.empty(n, np.uint32),
        "loc": np.zeros(n, np.uint32), "type": np.empty(n, np.uint32),
        "radius": None if rad is None else np.empty(n, np.float32),
    }
    lib.orc_reorder(_p(cs), _p(ce), _p(s["pos"]), _p(s["vel"]), _p(s["theta"]), _p(s["omega"]), _p(s["id"]),
                    _p(s["loc"]), _p(s["type"]), _p(s["radius"]), _p(hash_), _p(index), _p(old["pos"]),
                    _p(old["vel"]), _p(old["theta"]), _p(old["omega"]), _p(old["id"]), _p(old["type"]), _p(rad),
                    n, num_cells)
    return cs, ce, s


def collide(system: System, st: dict, cell_start, cell_end):
    lib = load()
    n = st["pos"].shape[0]
    acc = np.empty((n, 2), np.float32)
    alpha = np.empty(n, np.float32)
    pres = np.empty(n, np.float32)
    lib.orc_collide(ctypes.byref(system.sys), _p(st["pos"]), _p(st["vel"]), _p(acc), _p(st["omega"]), _p(alpha),
                    _p(st["type"]), _p(st.get("radius")), _p(cell_start), _p(cell_end), n, _p(pres))
    return acc, alpha, pres


def integrate(system: System, st: dict, acc, alpha, pressure):
    """Returns a new state dict (pos, vel, theta, omega) and the updated pressure."""
    lib = load()
    n = st["pos"].shape[0]
    out = {k: st[k].copy() for k in ("pos", "vel", "theta", "omega")}
    pres = pressure.copy()
    lib.orc_integrate(ctypes.byref(system.sys), _p(out["pos"]), _p(out["vel"]), _p(acc), _p(out["theta"]),
                      _p(out["omega"]), _p(alpha), _p(st["type"]), _p(st.get("radius")), _p(pres), n)
    return out, pres


class Control(ctypes.Structure):
    """orc::Control: the controlled particle (pos, vel, theta, omega, type,
    enabled, per-step move in double, wrap thresholds in double, resets)."""

    _fields_ = [
        ("pos", c_float * 2), ("vel", c_float * 2), ("theta", c_float), ("omega", c_float),
        ("type", c_uint32), ("enabled", c_uint32), ("move", ctypes.c_double * 2),
        ("wrap_below_y", ctypes.c_double), ("wrap_above_x", ctypes.c_double),
        ("reset_y", c_float), ("reset_x", c_float),
    ]


def control(pos, type_, vel=(0.0, 0.0), omega=0.0, theta=0.0, move=(0.0, 0.0), wrap_below_y=-1e300,
            wrap_above_x=1e300, reset_y=0.0, reset_x=0.0) -> Control:
    c = Control()
    c.pos[:] = pos
    c.vel[:] = vel
    c.theta, c.omega, c.type, c.enabled = theta, omega, type_, 1
    c.move[:] = move
    c.wrap_below_y, c.wrap_above_x, c.reset_y, c.reset_x = wrap_below_y, wrap_above_x, reset_y, reset_x
    return c


def collide_controlled(system: System, st: dict, cs, ce, ctl: Control):
    """collide with the controlled particle: (acc, alpha, pressure,
    (react_x, react_y, react_m), reaction[3])."""
    lib = load()
    n = st["pos"].shape[0]
    acc = np.zeros((n, 2), np.float32)
    alpha = np.zeros(n, np.float32)
    pres = np.zeros(n, np.float32)
    rx, ry, rm = (np.zeros(n, np.float32) for _ in range(3))
    red = np.zeros(3, np.float32)
    lib.orc_collide_controlled(ctypes.byref(system.sys), _p(st["pos"]), _p(st["vel"]), _p(acc), _p(st["omega"]),
                               _p(alpha), _p(st["type"]), _p(st.get("radius")), _p(cs), _p(ce), n, _p(pres),
                               ctypes.byref(ctl), _p(rx), _p(ry), _p(rm), _p(red))
    return acc, alpha, pres, (rx, ry, rm), red


# ---- owned simulation ----------------------------------------------------------------

class Sim:
    """Whole steps on the CPU (precision 'f32' or 'f64')."""

    def __init__(self, scene, precision: str = "f32"):
        assert precision in ("f32", "f64")
        self.lib = load()
        self.prec = precision
        self.system = System(scene)
        self.n = scene.n
        self.cells = scene.total_cells
        self.with_radius = scene.radius is not None
        create = getattr(self.lib, f"orc_sim_create_{precision}")
        self.h = create(ctypes.byref(self.system.sys), len(scene.materials), len(scene.instances), self.n,
                        self.cells, int(self.with_radius))
        if not self.h:
            raise MemoryError("oracle sim allocation failed")
        rc = getattr(self.lib, f"orc_sim_set_state_{precision}")(
            self.h, _p(scene.pos), _p(scene.vel), _p(scene.theta), _p(scene.omega), _p(scene.id), _p(scene.type),
            _p(scene.loc), _p(scene.radius))
        if rc != 0:
            raise ValueError("oracle sim: radius array missing")

    def close(self):
        if self.h:
            getattr(self.lib, f"orc_sim_destroy_{self.prec}")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, steps: int = 1) -> float:
        """Runs `steps` steps; returns the wall time in seconds."""
        return float(getattr(self.lib, f"orc_sim_step_{self.prec}")(self.h, steps))

    def set_control(self, ctl: Control) -> None:
        getattr(self.lib, f"orc_sim_set_control_{self.prec}")(self.h, ctypes.byref(ctl))

    def control(self):
        """(Control, reaction[3] float64) after the steps run so far."""
        c = Control()
        r = np.zeros(3, np.float64)
        getattr(self.lib, f"orc_sim_get_control_{self.prec}")(self.h, ctypes.byref(c), _p(r))
        return c, r

    def state(self) -> dict:
        n, c = self.n, self.cells
        f = np.float32 if self.prec == "f32" else np.float64
        out = {
            "pos": np.empty((n, 2), f), "vel": np.empty((n, 2), f), "theta": np.empty(n, f),
            "omega": np.empty(n, f), "id": np.empty(n, np.uint32), "type": np.empty(n, np.uint32),
            "loc": np.empty(n, np.uint32), "radius": np.empty(n, f) if self.with_radius else None,
            "acc": np.empty((n, 2), f), "alpha": np.empty(n, f), "pressure": np.empty(n, f),
            "hash": np.empty(n, np.uint32), "index": np.empty(n, np.uint32), "cellStart": np.empty(c, np.uint32),
            "cellEnd": np.empty(c, np.uint32),
        }
        g = lambda k: _p(out.get(k))
        getattr(self.lib, f"orc_sim_get_state_{self.prec}")(
            self.h, g("pos"), g("vel"), g("theta"), g("omega"), g("id"), g("type"), g("loc"), g("radius"),
            g("acc"), g("alpha"), g("pressure"), g("hash"), g("index"), g("cellStart"), g("cellEnd"))
        return out


class AdjointOracle:
    """FP64 forward + hand-written adjoint on the CPU (the reference for the
    HIP adjoint).  Gradients are by particle id; params per material in the
    order (kn, kt, cn, mu)."""

    def __init__(self, scene):
        self.lib = load()
        self.system = System(scene)
        self.n = scene.n
        self.nt = len(scene.materials)
        self.h = self.lib.orc_adj_create(ctypes.byref(self.system.sys), self.nt, len(scene.instances), self.n,
                                         scene.total_cells, int(scene.radius is not None))
        self.lib.orc_adj_set_state(self.h, _p(scene.pos), _p(scene.vel), _p(scene.theta), _p(scene.omega),
                                   _p(scene.id), _p(scene.type), _p(scene.loc), _p(scene.radius))

    def close(self):
        if self.h:
            self.lib.orc_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, steps: int):
        self.lib.orc_adj_forward(self.h, steps)
        pos = np.empty((self.n, 2), np.float64)
        vel = np.empty((self.n, 2), np.float64)
        self.lib.orc_adj_final_state(self.h, _p(pos), _p(vel))
        return pos, vel

    def backward(self, gpos: np.ndarray, gvel: Optional[np.ndarray] = None):
        gpos = np.ascontiguousarray(gpos, np.float64)
        gv = None if gvel is None else np.ascontiguousarray(gvel, np.float64)
        gp0 = np.empty((self.n, 2), np.float64)
        gv0 = np.empty((self.n, 2), np.float64)
        gpar = np.empty((self.nt, 4), np.float64)
        self.lib.orc_adj_backward(self.h, _p(gpos), _p(gv), _p(gp0), _p(gv0), _p(gpar))
        return gp0, gv0, gpar


# ---- images ---------------------------------------------------------------------------

class RenderParams(ctypes.Structure):
    """orc::RenderParams (render_oracle.hpp)."""

    _fields_ = [("dimx", ctypes.c_int32), ("dimy", ctypes.c_int32), ("cdimx", ctypes.c_int32),
                ("cdimy", ctypes.c_int32), ("bg", ctypes.c_int32), ("rotations", c_uint32),
                ("by_pressure", c_uint32), ("max_pressure", c_float), ("box", c_float * 2)]


def plot(scene, rp: dict, pos, theta, type_, pressure, radius=None, ctl: Optional["Control"] = None) -> np.ndarray:
    """The image (dimy, dimx, 4) uint8 of plotParticles for arrays in the
    given order; rp: image_dim, control_dim, bg_color, view_rotations,
    color_by_pressure, max_pressure (as demso_amd.api.render_params)."""
    lib = load()
    types = np.array([[m.radius, *m.color] for m in scene.materials], np.float32)
    r = RenderParams()
    r.dimx, r.dimy = rp["image_dim"]
    r.cdimx, r.cdimy = rp["control_dim"]
    r.bg, r.rotations, r.by_pressure = rp["bg_color"], int(rp["view_rotations"]), int(rp["color_by_pressure"])
    r.max_pressure = rp["max_pressure"]
    r.box[:] = scene.instances[0].box
    img = np.empty((r.dimy, r.dimx, 4), np.uint8)
    n = pos.shape[0]
    c = lambda a, dt: None if a is None else np.ascontiguousarray(a, dt)
    arrs = [c(pos, np.float32), c(theta, np.float32), c(type_, np.uint32), c(radius, np.float32),
            c(pressure, np.float32)]
    lib.orc_plot_particles(_p(img), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]), n, _p(types),
                           len(scene.materials), ctypes.byref(r), None if ctl is None else ctypes.byref(ctl),
                           _p(arrs[4]))
    return img


def pressure_scale(p) -> float:
    a = np.ascontiguousarray(p, np.float32)
    return float(load().orc_pressure_scale(_p(a), a.size))
