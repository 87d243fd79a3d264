"""Python mirror of DemSo's step interface (src/functions.cuh:29-89) over the
C ABI of libdemso.so.

The stage functions keep the reference's names, argument order and meaning
(calcHash, sortParticles, reorderDataAndFindCellStart, collide,
integrateSystem, set_gravity, allocateVectors/desAllocateVectors), with a
Context first; arrays are torch tensors on the GPU whose memory is handed to
the HIP kernels as raw device pointers.  Errors raise DemError with the
library's message instead of being ignored.

Simulation owns the double-buffered state and runs whole steps (eagerly or
by hipGraph replay), like SimLooping (src/main.cu:371-565) without the
renderer.
"""
from __future__ import annotations

import ctypes
import weakref
from ctypes import byref, c_float, c_uint32, c_void_p
from typing import Optional

import numpy as np

from . import _lib
from ._lib import DemError, InstanceProps, ParticleProps, SystemProps, check
from .scene import Scene

__all__ = [
    "Context",
    "Simulation",
    "DeviceState",
    "calcHash",
    "sortParticles",
    "reorderDataAndFindCellStart",
    "collide",
    "integrateSystem",
    "set_gravity",
    "allocateVectors",
    "desAllocateVectors",
    "DemError",
]


def _ptr(t) -> Optional[int]:
    """Device address of a tensor (None passes NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("arrays must be GPU tensors")
    if not t.is_contiguous():
        raise ValueError("arrays must be contiguous")
    return t.data_ptr()


def _np_ptr(a: Optional[np.ndarray]) -> Optional[int]:
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("host arrays must be C-contiguous")
    return a.ctypes.data


def system_props(scene: Scene) -> SystemProps:
    inst0 = scene.instances[0]
    sp = SystemProps()
    sp.num_particles = scene.n
    sp.cube_dimension[:] = inst0.box
    sp.grid_size[:] = inst0.grid
    sp.num_cells = inst0.grid[0] * inst0.grid[1]
    sp.time_step = scene.dt
    sp.gravity[:] = scene.gravity
    sp.boundary_normal_stiffness = scene.bkn
    sp.boundary_shear_stiffness = scene.bkt
    sp.grid_extent[:] = inst0.extent
    sp.flags = scene.flags
    return sp


def particle_props(scene: Scene, single_instance: bool) -> ctypes.Array:
    arr = (ParticleProps * len(scene.materials))()
    for k, m in enumerate(scene.materials):
        p = arr[k]
        p.radius, p.mass, p.inertia = m.radius, m.mass, m.inertia
        p.normal_stiffness, p.shear_stiffness, p.normal_damping = m.kn, m.kt, m.cn
        p.boundary_damping, p.friction_coefficient, p.attract_coefficient = m.bdamp, m.mu, m.attract
        p.color_r, p.color_g, p.color_b = (float(v) for v in m.color)
        p.areal_density = m.areal
        p.instance = 0 if single_instance else m.instance
    return arr


def instance_props(scene: Scene) -> ctypes.Array:
    arr = (InstanceProps * len(scene.instances))()
    for k, g in enumerate(scene.instances):
        arr[k].cube_dimension[:] = g.box
        arr[k].grid_extent[:] = g.extent
        arr[k].grid_size[:] = g.grid
        arr[k].cell_base = g.cell_base
    return arr


class Context:
    """dem_context: device, stream and the parameter tables."""

    def __init__(self, device: int = 0, stream: Optional[int] = None):
        self.lib = _lib.load()
        h = c_void_p()
        check(self.lib.dem_create(byref(h), device, stream), None)
        self.handle = h.value
        self.device = device
        self.box = None  # (Lx, Ly) of the last scene loaded
        # simulations made on this context: destroyed before it, so that no
        # sim outlives the C context it points to
        self._sims = weakref.WeakSet()

    def close(self) -> None:
        if self.handle:
            for sim in list(self._sims):
                sim.close()
            self.lib.dem_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, code: int) -> None:
        check(code, self.handle)

    def set_params(self, sys: SystemProps, types: ctypes.Array) -> None:
        self._check(self.lib.dem_set_params(self.handle, byref(sys), types, len(types)))

    def set_instances(self, inst: ctypes.Array) -> None:
        self._check(self.lib.dem_set_instances(self.handle, inst, len(inst)))

    def set_materials(self, types: ctypes.Array) -> None:
        self._check(self.lib.dem_set_materials(self.handle, types, len(types)))

    def load_scene(self, scene: Scene) -> None:
        """Upload the scene's parameters: one system, or an ensemble (geometry
        first, then the materials that name their instances)."""
        self.set_params(system_props(scene), particle_props(scene, single_instance=True))
        self.box = tuple(scene.instances[0].box)
        if len(scene.instances) > 1:
            self.set_instances(instance_props(scene))
            self.set_materials(particle_props(scene, single_instance=False))

    def set_step_variant(self, fused: bool = False, prefetch: bool = False, batch: bool = False,
                         copy: bool = False, wide: bool = False, contact_lists: bool = False,
                         per_lane: Optional[dict] = None) -> None:
        """Owned simulations: collide + integrate as one kernel or two, and
        how collide loads its candidates: one at a time, one ahead
        (prefetch) or the first of every row at once (batch); copy
        makes re-sorted steps reorder the whole state into the other buffer
        instead of moving only the particles whose slot changes; wide
        makes collide use 64-bit addresses (as states over 2^29 particles
        or 2^30 cells do); contact_lists makes collide mark the candidates that may
        touch before running the contact law over them (contact-dense
        states); per_lane gives the particles a lane (1, 2 or 4) of "hash",
        "reorder" and "integrate".  The results are the same in every
        combination."""
        flags = (int(bool(fused)) | (2 if prefetch else 0) | (4 if batch else 0) | (8 if copy else 0)
                 | (16 if wide else 0) | (32 if contact_lists else 0))
        lg = {1: 0, 2: 1, 4: 2}
        for shift, k in ((8, "hash"), (10, "reorder"), (12, "integrate")):
            per = (per_lane or {}).get(k, 1)
            if per not in lg:
                raise ValueError("particles a lane must be 1, 2 or 4")
            flags |= lg[per] << shift
        self._check(self.lib.dem_set_step_variant(self.handle, flags))

    def set_gravity(self, gx: float, gy: float) -> None:
        self._check(self.lib.dem_set_gravity(self.handle, gx, gy))

    def set_controlled(self, pos=(5.0, 9.5), type: int = 0, vel=(0.0, 0.0), theta: float = 0.0,
                       omega: float = 0.0, move=(0.0, -0.005), wrap_below_y: float = -1.0,
                       wrap_above_x: Optional[float] = None, reset_y: float = 10.5, reset_x: float = -0.5,
                       enabled: bool = True) -> None:
        """The controlled particle of this context's owned simulations
        (src/main.cu:510-514 and src/particles_kernel.cuh:525-543); the
        defaults are DemSo's.  wrap_above_x defaults to Lx + 0.5."""
        if wrap_above_x is None:
            if self.box is None:
                raise ValueError("wrap_above_x needs the box: load a scene first or pass it")
            wrap_above_x = float(np.float32(self.box[0])) + 0.5
        c = _lib.ControlledProps()
        c.pos[:] = [float(pos[0]), float(pos[1])]
        c.vel[:] = [float(vel[0]), float(vel[1])]
        c.theta, c.omega, c.type, c.enabled = float(theta), float(omega), int(type), int(bool(enabled))
        c.move[:] = [float(move[0]), float(move[1])]
        c.wrap_below_y, c.wrap_above_x = float(wrap_below_y), float(wrap_above_x)
        c.reset_y, c.reset_x = float(reset_y), float(reset_x)
        self._check(self.lib.dem_set_controlled(self.handle, byref(c)))

    def controlled(self):
        """(state dict, reaction (fx, fy, moment) as float32) after the steps
        run so far."""
        c = _lib.ControlledProps()
        r = (c_float * 3)()
        self._check(self.lib.dem_get_controlled(self.handle, byref(c), r))
        state = {"pos": np.array(c.pos[:], np.float32), "vel": np.array(c.vel[:], np.float32),
                 "theta": np.float32(c.theta), "omega": np.float32(c.omega), "type": int(c.type),
                 "enabled": bool(c.enabled)}
        return state, np.array(r[:], np.float32)

    def set_contact_params(self, params) -> None:
        """params: (num_types, 4) array of (kn, kt, cn, mu) per material."""
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1, 4))
        self._check(self.lib.dem_set_contact_params(
            self.handle, p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), p.shape[0]))

    def reserve(self, n: int) -> None:
        self._check(self.lib.dem_reserve(self.handle, n))

    def synchronize(self) -> None:
        self._check(self.lib.dem_synchronize(self.handle))


# ---- the five stages, reference names ----------------------------------------

def render_params(scene: Scene, image_height: int = 800, controlled_type: Optional[int] = None, bg_color: int = 0,
                  view_rotations: bool = True, color_by_pressure: bool = False, max_pressure: float = 1.0) -> dict:
    """DemSo's RenderParameters for a scene (src/main.cu:172-177, 209-215):
    imageDIMx = Lx/Ly * imageDIMy and the controlled particle's square of
    2 ceil(pixels per metre * its radius) (at least 2), in FP32."""
    f32 = np.float32
    lx, ly = (f32(v) for v in scene.instances[0].box)
    dimy = int(image_height)
    dimx = int(lx / ly * f32(dimy))
    cdim = [0, 0]
    if controlled_type is not None:
        rc = f32(scene.materials[controlled_type].radius)
        cdim = [max(int(np.ceil(f32(dimx) / lx * rc) * f32(2)), 2), max(int(np.ceil(f32(dimy) / ly * rc) * f32(2)), 2)]
    return {"image_dim": [dimx, dimy], "control_dim": cdim, "bg_color": int(bg_color),
            "view_rotations": bool(view_rotations), "color_by_pressure": bool(color_by_pressure),
            "max_pressure": float(max_pressure)}


def _render_struct(rp: dict):
    r = _lib.RenderProps()
    r.image_dim[:] = rp["image_dim"]
    r.control_dim[:] = rp["control_dim"]
    r.bg_color = rp["bg_color"]
    r.view_rotations = int(rp["view_rotations"])
    r.color_by_pressure = int(rp["color_by_pressure"])
    r.max_pressure = rp["max_pressure"]
    return r


def plotParticles(ctx: Context, image, pos, theta, type, numParticles: int, rp: dict, pressure, radius=None,
                  control: Optional[dict] = None) -> None:
    """plotParticles (src/functions.cu:557-600) into a device uint8 tensor of
    image_dim[1] x image_dim[0] x 4 (RGBA, row 0 at the bottom).  control:
    the controlled particle {"pos", "type"} or None."""
    c = None
    if control is not None:
        c = _lib.ControlledProps()
        c.pos[:] = [float(control["pos"][0]), float(control["pos"][1])]
        c.type, c.enabled = int(control["type"]), 1
    r = _render_struct(rp)
    ctx._check(ctx.lib.dem_plot_particles(ctx.handle, _ptr(image), _ptr(pos), _ptr(theta), _ptr(type), _ptr(radius),
                                          numParticles, byref(r), None if c is None else byref(c), _ptr(pressure)))


def updatePressureScale(ctx: Context, pressure, numParticles: int) -> float:
    """updatePressureScale (src/functions.cu:684-693): the largest pressure,
    1 when none is positive."""
    out = c_float()
    ctx._check(ctx.lib.dem_update_pressure_scale(ctx.handle, _ptr(pressure), numParticles, byref(out)))
    return float(out.value)


def calcHash(ctx: Context, pos, gridParticleIndex, gridParticleHash, numParticles: int, type=None) -> None:
    """calcHash (src/functions.cu:315-327)."""
    if type is None:
        ctx._check(ctx.lib.dem_calc_hash(ctx.handle, _ptr(pos), _ptr(gridParticleIndex),
                                         _ptr(gridParticleHash), numParticles))
    else:
        ctx._check(ctx.lib.dem_calc_hash_ex(ctx.handle, _ptr(pos), _ptr(type), _ptr(gridParticleIndex),
                                            _ptr(gridParticleHash), numParticles))


def sortParticles(ctx: Context, gridParticleHash, gridParticleIndex, numParticles: int) -> None:
    """sortParticles (src/functions.cu:336-341): stable key/index sort, in place."""
    ctx._check(ctx.lib.dem_sort_particles(ctx.handle, _ptr(gridParticleHash), _ptr(gridParticleIndex),
                                          numParticles))


def sortParticles_noalloc(ctx: Context, gridParticleHash, gridParticleIndex, numParticles: int) -> None:
    """sortParticles without device allocation: the library's fixed per-GPU
    scratch (at most DEM_SORT_NOALLOC_MAX keys; calls overlapping on one GPU
    race)."""
    ctx._check(ctx.lib.dem_sort_particles_noalloc(ctx.handle, _ptr(gridParticleHash), _ptr(gridParticleIndex),
                                                  numParticles))


def reorderDataAndFindCellStart(ctx: Context, cellStart, cellEnd, sortedPos, sortedVel, sortedTheta,
                                sortedOmega, sortedID, sortedLoc, sortedType, gridParticleHash,
                                gridParticleIndex, oldPos, oldVel, oldTheta, oldOmega, oldID, oldType,
                                numParticles: int, numCells: int, sortedRadius=None, oldRadius=None) -> None:
    """reorderDataAndFindCellStart (src/functions.cu:347-393)."""
    ctx._check(ctx.lib.dem_reorder_ex(
        ctx.handle, _ptr(cellStart), _ptr(cellEnd), _ptr(sortedPos), _ptr(sortedVel), _ptr(sortedTheta),
        _ptr(sortedOmega), _ptr(sortedID), _ptr(sortedLoc), _ptr(sortedType), _ptr(sortedRadius),
        _ptr(gridParticleHash), _ptr(gridParticleIndex), _ptr(oldPos), _ptr(oldVel), _ptr(oldTheta),
        _ptr(oldOmega), _ptr(oldID), _ptr(oldType), _ptr(oldRadius), numParticles, numCells))


def collide(ctx: Context, oldPos, oldVel, newAcc, oldOmega, newAlpha, oldType, cellStart, cellEnd,
            numParticles: int, numCells: int, pressure, radius=None) -> None:
    """collide (src/functions.cu:398-521, without the controlled particle)."""
    ctx._check(ctx.lib.dem_collide_ex(
        ctx.handle, _ptr(oldPos), _ptr(oldVel), _ptr(newAcc), _ptr(oldOmega), _ptr(newAlpha), _ptr(oldType),
        _ptr(radius), _ptr(cellStart), _ptr(cellEnd), numParticles, numCells, _ptr(pressure)))


def collide_controlled(ctx: Context, oldPos, oldVel, newAcc, oldOmega, newAlpha, oldType, cellStart, cellEnd,
                       numParticles: int, numCells: int, controlPos, controlVel, controlTheta: float,
                       controlOmega: float, controlType: int, controlForceVecX, hCFVx, controlForceVecY, hCFVy,
                       controlMomentVec, hCMV, pressure):
    """collide with the controlled particle (src/functions.cu:398-521,
    USE_BIG_PARTICLE 1, USE_ATOMIC 0); returns (ctrlF (2,), ctrlM) as
    float32."""
    state = (c_float * 6)(float(controlPos[0]), float(controlPos[1]), float(controlVel[0]), float(controlVel[1]),
                          float(controlTheta), float(controlOmega))
    f = (c_float * 2)()
    m = (c_float * 1)()
    ctx._check(ctx.lib.dem_collide_controlled(
        ctx.handle, _ptr(oldPos), _ptr(oldVel), _ptr(newAcc), _ptr(oldOmega), _ptr(newAlpha), _ptr(oldType),
        _ptr(cellStart), _ptr(cellEnd), numParticles, numCells, state, int(controlType), _ptr(controlForceVecX),
        _ptr(hCFVx), _ptr(controlForceVecY), _ptr(hCFVy), _ptr(controlMomentVec), _ptr(hCMV), f, m,
        _ptr(pressure)))
    return np.array(f[:], np.float32), np.float32(m[0])


def createRectangles(ctx: Context, pos, ID, loc, type, start, sideLenght, side, startID: int,
                     particleTypeVec, seed: int, stream: Optional[int] = None) -> None:
    """createRectangles (src/functions.cu:198-245) on the device; the jitter
    draws are the reproducible normals of streams `stream` (x) and stream + 1
    (y), default DEM_RECT_STREAM (1000)."""
    tv = np.ascontiguousarray(particleTypeVec, dtype=np.uint32)
    st = (c_float * 2)(float(start[0]), float(start[1]))
    ln = (c_float * 2)(float(sideLenght[0]), float(sideLenght[1]))
    sd = (c_uint32 * 2)(int(side[0]), int(side[1]))
    tvp = tv.ctypes.data_as(ctypes.POINTER(c_uint32))
    if stream is None:
        ctx._check(ctx.lib.dem_create_rectangles(ctx.handle, _ptr(pos), _ptr(ID), _ptr(loc), _ptr(type), st, ln, sd,
                                                 int(startID), tv.size, tvp, int(seed)))
    else:
        ctx._check(ctx.lib.dem_create_rectangles_ex(ctx.handle, _ptr(pos), _ptr(ID), _ptr(loc), _ptr(type), st, ln,
                                                    sd, int(startID), tv.size, tvp, int(seed), int(stream)))


def createTriangles(ctx: Context, pos, ID, loc, type, start, N: int, particleTypeVec, space: float, height: float,
                    startID: int, numParticles: int) -> None:
    """createTriangles (src/functions.cu:247-281) on the device."""
    tv = np.ascontiguousarray(particleTypeVec, dtype=np.uint32)
    st = (c_float * 2)(float(start[0]), float(start[1]))
    ctx._check(ctx.lib.dem_create_triangles(ctx.handle, _ptr(pos), _ptr(ID), _ptr(loc), _ptr(type), st, int(N),
                                            tv.size, tv.ctypes.data_as(ctypes.POINTER(c_uint32)), float(space),
                                            float(height), int(startID), int(numParticles)))


def createSingleParticles(ctx: Context, pos, vel, theta, omega, ID, loc, type, usrPos, usrVel, usrTheta, usrOmega,
                          usrType, numParticles: int, startID: int) -> None:
    """createSingleParticles (src/functions.cu:283-311): usr* are host numpy
    arrays or device tensors."""
    def src(a, dt):
        if a is None or hasattr(a, "data_ptr"):
            return _ptr(a), None
        keep = np.ascontiguousarray(a, dtype=dt)
        return keep.ctypes.data, keep
    srcs = [src(usrPos, np.float32), src(usrVel, np.float32), src(usrTheta, np.float32), src(usrOmega, np.float32),
            src(usrType, np.uint32)]
    ctx._check(ctx.lib.dem_create_single_particles(
        ctx.handle, _ptr(pos), _ptr(vel), _ptr(theta), _ptr(omega), _ptr(ID), _ptr(loc), _ptr(type),
        *[p for p, _ in srcs], int(numParticles), int(startID)))


def integrateSystem(ctx: Context, pos, vel, acc, theta, omega, alpha, type, pressure, numParticles: int,
                    radius=None) -> None:
    """integrateSystem (src/functions.cu:529-551)."""
    ctx._check(ctx.lib.dem_integrate_ex(
        ctx.handle, _ptr(pos), _ptr(vel), _ptr(acc), _ptr(theta), _ptr(omega), _ptr(alpha), _ptr(type),
        _ptr(radius), _ptr(pressure), numParticles))


def set_gravity(ctx: Context, gravity) -> None:
    """set_gravity (src/functions.cu:663-668)."""
    ctx.set_gravity(float(gravity[0]), float(gravity[1]))


class DeviceState:
    """The per-particle arrays of ParticlesValues (src/main.cuh:58-104) for
    one buffer set, as torch tensors."""

    FIELDS = ("pos", "vel", "theta", "omega", "id", "type", "loc", "radius")

    def __init__(self, n: int, with_radius: bool, device: str = "cuda"):
        import torch

        f32, i32 = torch.float32, torch.int32
        self.n = n
        self.pos = torch.zeros((n, 2), dtype=f32, device=device)
        self.vel = torch.zeros((n, 2), dtype=f32, device=device)
        self.theta = torch.zeros(n, dtype=f32, device=device)
        self.omega = torch.zeros(n, dtype=f32, device=device)
        self.id = torch.zeros(n, dtype=i32, device=device)
        self.type = torch.zeros(n, dtype=i32, device=device)
        self.loc = torch.zeros(n, dtype=i32, device=device)
        self.radius = torch.zeros(n, dtype=f32, device=device) if with_radius else None

    def upload(self, scene: Scene) -> None:
        import torch

        for f in self.FIELDS:
            src = getattr(scene, f)
            dst = getattr(self, f)
            if dst is None:
                continue
            dst.copy_(torch.from_numpy(np.ascontiguousarray(src).view(np.int32 if src.dtype == np.uint32 else src.dtype)))


def allocateVectors(scene: Scene, ctx: Context) -> dict:
    """allocateVectors (src/functions.cu:68-149): device arrays for both
    buffer sets plus acc/alpha/pressure/cell/key arrays, parameters uploaded
    to ctx.  Returns a dict of tensors (freed by desAllocateVectors)."""
    import torch

    n, cells = scene.n, scene.total_cells
    rad = scene.radius is not None
    ctx.load_scene(scene)
    ctx.reserve(n)
    v = {
        "s1": DeviceState(n, rad),
        "s2": DeviceState(n, rad),
        "acc": torch.zeros((n, 2), dtype=torch.float32, device="cuda"),
        "alpha": torch.zeros(n, dtype=torch.float32, device="cuda"),
        "pressure": torch.zeros(n, dtype=torch.float32, device="cuda"),
        "cellStart": torch.zeros(cells, dtype=torch.int32, device="cuda"),
        "cellEnd": torch.zeros(cells, dtype=torch.int32, device="cuda"),
        "gridParticleIndex": torch.zeros(n, dtype=torch.int32, device="cuda"),
        "gridParticleHash": torch.zeros(n, dtype=torch.int32, device="cuda"),
    }
    return v


def desAllocateVectors(values: dict) -> None:
    """desAllocateVectors (src/functions.cu:154-194)."""
    values.clear()


class Simulation:
    """dem_sim: owned double-buffered device state and whole steps."""

    def __init__(self, ctx: Context, scene: Scene):
        self.ctx = ctx
        self.lib = ctx.lib
        self.n = scene.n
        self.cells = scene.total_cells
        self.with_radius = scene.radius is not None
        ctx.load_scene(scene)
        h = c_void_p()
        ctx._check(self.lib.dem_sim_create(ctx.handle, self.n, self.cells, int(self.with_radius), byref(h)))
        self.handle = h.value
        self._adjoints = weakref.WeakSet()  # destroyed before this sim
        ctx._sims.add(self)
        self.upload(scene)

    def close(self) -> None:
        if getattr(self, "handle", None):
            for adj in list(getattr(self, "_adjoints", ())):
                adj.close()
            self.lib.dem_sim_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, scene: Scene) -> None:
        r = scene.radius if self.with_radius else None
        self.ctx._check(self.lib.dem_sim_upload(
            self.handle, _np_ptr(scene.pos), _np_ptr(scene.vel), _np_ptr(scene.theta), _np_ptr(scene.omega),
            _np_ptr(scene.id), _np_ptr(scene.type), _np_ptr(scene.loc), _np_ptr(r)))

    def download(self, extra: bool = True) -> dict:
        n, c = self.n, self.cells
        out = {
            "pos": np.empty((n, 2), np.float32), "vel": np.empty((n, 2), np.float32),
            "theta": np.empty(n, np.float32), "omega": np.empty(n, np.float32),
            "id": np.empty(n, np.uint32), "type": np.empty(n, np.uint32), "loc": np.empty(n, np.uint32),
            "radius": np.empty(n, np.float32) if self.with_radius else None,
        }
        if extra:
            out.update({
                "acc": np.empty((n, 2), np.float32), "alpha": np.empty(n, np.float32),
                "pressure": np.empty(n, np.float32), "hash": np.empty(n, np.uint32),
                "index": np.empty(n, np.uint32), "cellStart": np.empty(c, np.uint32),
                "cellEnd": np.empty(c, np.uint32),
            })
        g = lambda k: _np_ptr(out.get(k))
        self.ctx._check(self.lib.dem_sim_download(
            self.handle, g("pos"), g("vel"), g("theta"), g("omega"), g("id"), g("type"), g("loc"), g("radius"),
            g("acc"), g("alpha"), g("pressure"), g("hash"), g("index"), g("cellStart"), g("cellEnd")))
        return out

    def step(self, steps: int = 1) -> None:
        self.ctx._check(self.lib.dem_sim_step(self.handle, steps))

    def step_graph(self, steps: int) -> None:
        self.ctx._check(self.lib.dem_sim_step_graph(self.handle, steps))

    def stage_times(self, reps: int = 5) -> dict:
        ms = (c_float * 7)()
        self.ctx._check(self.lib.dem_sim_stage_times(self.handle, reps, ms))
        names = ("hash", "sort", "cell_fill", "reorder", "collide", "integrate", "step")
        return {k: float(ms[i]) for i, k in enumerate(names)}

    def autotune(self) -> dict:
        b = (c_uint32 * 5)()
        self.ctx._check(self.lib.dem_sim_autotune(self.handle, b))
        return {"hash": b[0], "reorder": b[1], "collide": b[2], "integrate": b[3], "fused": bool(b[4] & 1),
                "prefetch": bool(b[4] & 2), "batch": bool(b[4] & 4), "copy": bool(b[4] & 8),
                "contact_lists": bool(b[4] & 32),
                "per_lane": {"hash": 1 << ((b[4] >> 8) & 3), "reorder": 1 << ((b[4] >> 10) & 3),
                             "integrate": 1 << ((b[4] >> 12) & 3)}}

    def sort_passes(self) -> int:
        return int(self.lib.dem_sim_sort_passes(self.handle))

    def set_sort_mode(self, mode: int) -> None:
        """0 auto, 1 radix sort every step, 2 re-sort whenever possible."""
        self.ctx._check(self.lib.dem_sim_set_sort_mode(self.handle, mode))

    def mover_count(self) -> int:
        """Particles that changed cell in the last re-sorted step."""
        self.ctx.synchronize()
        v = c_uint32()
        self.ctx._check(self.lib.dem_sim_mover_count(self.handle, byref(v)))
        return int(v.value)

    def resort_counts(self) -> dict:
        """The last re-sort's movers and, in place, the particles it listed
        (moved to another slot or given a new key)."""
        self.ctx.synchronize()
        a, b = c_uint32(), c_uint32()
        self.ctx._check(self.lib.dem_sim_resort_counts(self.handle, byref(a), byref(b)))
        return {"movers": int(a.value), "listed": int(b.value)}

    def get_by_id(self):
        """Current positions and velocities as (n, 2) GPU tensors indexed by particle id."""
        import torch

        pos = torch.empty((self.n, 2), dtype=torch.float32, device="cuda")
        vel = torch.empty((self.n, 2), dtype=torch.float32, device="cuda")
        self.ctx._check(self.lib.dem_sim_get_by_id(self.handle, _ptr(pos), _ptr(vel)))
        return pos, vel

    def set_by_id(self, pos=None, vel=None) -> None:
        """Overwrite positions / velocities from (n, 2) GPU tensors indexed by id."""
        self.ctx._check(self.lib.dem_sim_set_by_id(self.handle, _ptr(pos), _ptr(vel)))

    def pressure_scale(self) -> float:
        """updatePressureScale over the current pressure."""
        out = c_float()
        self.ctx._check(self.lib.dem_sim_pressure_scale(self.handle, byref(out)))
        return float(out.value)

    def plot(self, rp: dict) -> np.ndarray:
        """The image of the current state (dem_sim_plot), (dimy, dimx, 4) uint8."""
        import torch

        img = torch.empty((rp["image_dim"][1], rp["image_dim"][0], 4), dtype=torch.uint8, device="cuda")
        r = _render_struct(rp)
        self.ctx._check(self.lib.dem_sim_plot(self.handle, _ptr(img), byref(r)))
        return img.cpu().numpy()

    def synchronize(self) -> None:
        self.ctx.synchronize()


class Adjoint:
    """dem_adjoint: reverse mode of collide+integrate over a run of steps,
    with checkpoints every `checkpoint_every` steps."""

    def __init__(self, sim: Simulation, checkpoint_every: int = 100):
        self.sim = sim
        self.lib = sim.lib
        h = c_void_p()
        sim.ctx._check(self.lib.dem_adj_create(sim.handle, checkpoint_every, byref(h)))
        self.handle = h.value
        sim._adjoints.add(self)

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.dem_adj_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, steps: int) -> None:
        self.sim.ctx._check(self.lib.dem_adj_forward(self.handle, steps))

    def backward(self, grad_pos=None, grad_vel=None, num_types: int = 1):
        """grad_pos/grad_vel: (n, 2) float32 GPU tensors by particle id.
        Returns (dL/dpos0, dL/dvel0, dL/dparams (num_types, 4) float64)."""
        import torch

        n = self.sim.n
        gp0 = torch.empty((n, 2), dtype=torch.float32, device="cuda")
        gv0 = torch.empty((n, 2), dtype=torch.float32, device="cuda")
        gpar = torch.empty((num_types, 4), dtype=torch.float64, device="cuda")
        self.sim.ctx._check(self.lib.dem_adj_backward(self.handle, _ptr(grad_pos), _ptr(grad_vel), _ptr(gp0),
                                                      _ptr(gv0), _ptr(gpar)))
        return gp0, gv0, gpar
