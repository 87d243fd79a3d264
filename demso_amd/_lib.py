"""ctypes binding of libdemso.so (the C ABI declared in include/demso.h).

The library is the only compute path: if it is missing this module raises,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_uint32, c_void_p, c_char_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libdemso.so"

DEM_OK = 0
DEM_EMPTY_CELL = 0xFFFFFFFF
DEM_GRID_CLAMP = 0x1
DEM_OPEN_TOP = 0x2
STATUS_NAMES = {
    0: "DEM_OK",
    -1: "DEM_ERR_INVALID",
    -2: "DEM_ERR_NO_DEVICE",
    -3: "DEM_ERR_ALLOC",
    -4: "DEM_ERR_LAUNCH",
    -5: "DEM_ERR_STATE",
    -6: "DEM_ERR_NONFINITE",
}


class SystemProps(ctypes.Structure):
    """dem_system_props (SystemProperties, src/main.cuh:108-122, + additions)."""

    _fields_ = [
        ("num_particles", c_uint32),
        ("cube_dimension", c_float * 2),
        ("grid_size", c_uint32 * 2),
        ("num_cells", c_uint32),
        ("time_step", c_float),
        ("gravity", c_float * 2),
        ("boundary_normal_stiffness", c_float),
        ("boundary_shear_stiffness", c_float),
        ("grid_extent", c_float * 2),
        ("flags", c_uint32),
    ]


class ParticleProps(ctypes.Structure):
    """dem_particle_props (ParticleProperties, src/main.cuh:39-54, + additions)."""

    _fields_ = [
        ("radius", c_float),
        ("mass", c_float),
        ("inertia", c_float),
        ("normal_stiffness", c_float),
        ("shear_stiffness", c_float),
        ("normal_damping", c_float),
        ("boundary_damping", c_float),
        ("friction_coefficient", c_float),
        ("attract_coefficient", c_float),
        ("color_r", c_float),
        ("color_g", c_float),
        ("color_b", c_float),
        ("areal_density", c_float),
        ("instance", c_uint32),
    ]


class InstanceProps(ctypes.Structure):
    """dem_instance."""

    _fields_ = [
        ("cube_dimension", c_float * 2),
        ("grid_extent", c_float * 2),
        ("grid_size", c_uint32 * 2),
        ("cell_base", c_uint32),
        ("reserved", c_uint32),
    ]


class ControlledProps(ctypes.Structure):
    """dem_controlled."""

    _fields_ = [
        ("pos", c_float * 2),
        ("vel", c_float * 2),
        ("theta", c_float),
        ("omega", c_float),
        ("type", c_uint32),
        ("enabled", c_uint32),
        ("move", ctypes.c_double * 2),
        ("wrap_below_y", ctypes.c_double),
        ("wrap_above_x", ctypes.c_double),
        ("reset_y", c_float),
        ("reset_x", c_float),
    ]


class RenderProps(ctypes.Structure):
    """dem_render_params."""

    _fields_ = [
        ("image_dim", ctypes.c_int32 * 2),
        ("control_dim", ctypes.c_int32 * 2),
        ("bg_color", ctypes.c_int32),
        ("view_rotations", c_uint32),
        ("color_by_pressure", c_uint32),
        ("max_pressure", c_float),
    ]


P = c_void_p  # every device pointer travels as a plain address

# name -> (restype, argtypes)
PROTOTYPES = {
    "dem_abi_version": (c_int, []),
    "dem_last_error": (c_char_p, [c_void_p]),
    "dem_create": (c_int, [POINTER(c_void_p), c_int, c_void_p]),
    "dem_destroy": (c_int, [c_void_p]),
    "dem_set_stream": (c_int, [c_void_p, c_void_p]),
    "dem_synchronize": (c_int, [c_void_p]),
    "dem_set_params": (c_int, [c_void_p, POINTER(SystemProps), POINTER(ParticleProps), c_uint32]),
    "dem_set_instances": (c_int, [c_void_p, POINTER(InstanceProps), c_uint32]),
    "dem_set_materials": (c_int, [c_void_p, POINTER(ParticleProps), c_uint32]),
    "dem_set_gravity": (c_int, [c_void_p, c_float, c_float]),
    "dem_reserve": (c_int, [c_void_p, c_uint32]),
    "dem_calc_hash": (c_int, [c_void_p, P, P, P, c_uint32]),
    "dem_sort_particles": (c_int, [c_void_p, P, P, c_uint32]),
    "dem_sort_particles_noalloc": (c_int, [c_void_p, P, P, c_uint32]),
    "dem_reorder_data_and_find_cell_start": (
        c_int,
        [c_void_p] + [P] * 17 + [c_uint32, c_uint32],
    ),
    "dem_collide": (c_int, [c_void_p] + [P] * 8 + [c_uint32, c_uint32, P]),
    "dem_integrate_system": (c_int, [c_void_p] + [P] * 8 + [c_uint32]),
    "dem_calc_hash_ex": (c_int, [c_void_p, P, P, P, P, c_uint32]),
    "dem_reorder_ex": (c_int, [c_void_p] + [P] * 19 + [c_uint32, c_uint32]),
    "dem_collide_ex": (c_int, [c_void_p] + [P] * 9 + [c_uint32, c_uint32, P]),
    "dem_integrate_ex": (c_int, [c_void_p] + [P] * 9 + [c_uint32]),
    "dem_sim_create": (c_int, [c_void_p, c_uint32, c_uint32, c_int, POINTER(c_void_p)]),
    "dem_sim_destroy": (c_int, [c_void_p]),
    "dem_sim_upload": (c_int, [c_void_p] + [P] * 8),
    "dem_sim_download": (c_int, [c_void_p] + [P] * 15),
    "dem_sim_step": (c_int, [c_void_p, c_uint32]),
    "dem_sim_step_graph": (c_int, [c_void_p, c_uint32]),
    "dem_sim_stage_times": (c_int, [c_void_p, c_uint32, POINTER(c_float)]),
    "dem_sim_autotune": (c_int, [c_void_p, POINTER(c_uint32)]),
    "dem_sim_sort_passes": (c_int, [c_void_p]),
    "dem_sim_info": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_int)]),
    "dem_sim_set_sort_mode": (c_int, [c_void_p, c_int]),
    "dem_set_step_variant": (c_int, [c_void_p, c_int]),
    "dem_sim_mover_count": (c_int, [c_void_p, POINTER(c_uint32)]),
    "dem_sim_resort_counts": (c_int, [c_void_p, POINTER(c_uint32), POINTER(c_uint32)]),
    "dem_sim_get_by_id": (c_int, [c_void_p, P, P]),
    "dem_sim_set_by_id": (c_int, [c_void_p, P, P]),
    "dem_gather_rows": (c_int, [c_void_p, P, c_uint32] + [P] * 10 + [c_uint32, P]),
    "dem_sim_gather_rows": (c_int, [c_void_p, P, c_uint32, P]),
    "dem_create_rectangles": (
        c_int, [c_void_p] + [P] * 4 + [POINTER(c_float), POINTER(c_float), POINTER(c_uint32), c_uint32, c_uint32,
                                       POINTER(c_uint32), ctypes.c_uint64]),
    "dem_create_rectangles_ex": (
        c_int, [c_void_p] + [P] * 4 + [POINTER(c_float), POINTER(c_float), POINTER(c_uint32), c_uint32, c_uint32,
                                       POINTER(c_uint32), ctypes.c_uint64, ctypes.c_uint64]),
    "dem_create_triangles": (
        c_int, [c_void_p] + [P] * 4 + [POINTER(c_float), c_uint32, c_uint32, POINTER(c_uint32), c_float, c_float,
                                       c_uint32, c_uint32]),
    "dem_create_single_particles": (c_int, [c_void_p] + [P] * 12 + [c_uint32, c_uint32]),
    "dem_plot_particles": (c_int, [c_void_p] + [P] * 5 + [c_uint32, POINTER(RenderProps),
                                                        POINTER(ControlledProps), P]),
    "dem_sim_plot": (c_int, [c_void_p, P, POINTER(RenderProps)]),
    "dem_update_pressure_scale": (c_int, [c_void_p, P, c_uint32, POINTER(c_float)]),
    "dem_sim_pressure_scale": (c_int, [c_void_p, POINTER(c_float)]),
    "dem_set_controlled": (c_int, [c_void_p, POINTER(ControlledProps)]),
    "dem_get_controlled": (c_int, [c_void_p, POINTER(ControlledProps), POINTER(c_float)]),
    "dem_collide_controlled": (
        c_int,
        [c_void_p] + [P] * 8 + [c_uint32, c_uint32, POINTER(c_float), c_uint32] + [P] * 6
        + [POINTER(c_float), POINTER(c_float), P],
    ),
    "dem_set_contact_params": (c_int, [c_void_p, POINTER(ctypes.c_double), c_uint32]),
    "dem_adj_create": (c_int, [c_void_p, c_uint32, POINTER(c_void_p)]),
    "dem_adj_destroy": (c_int, [c_void_p]),
    "dem_adj_forward": (c_int, [c_void_p, c_uint32]),
    "dem_adj_generation": (c_int, [c_void_p, POINTER(ctypes.c_uint64)]),
    "dem_adj_backward": (c_int, [c_void_p, P, P, P, P, P]),
}

_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load libdemso.so and declare every prototype; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m demso_amd.build` "
            "(there is no fallback path)"
        )
    # torch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7) as the ROCm one libdemso.so is linked against; the
    # first one loaded serves the whole process.  Loading torch first keeps
    # torch on the runtime it was built with (torch then finds no GPU if
    # libdemso.so has brought in the other one), and libdemso.so uses it too.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.dem_abi_version() != 1:
        raise ImportError("libdemso.so ABI version mismatch")
    _lib = lib
    return lib


class DemError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


def check(code: int, ctx: int | None) -> None:
    if code != DEM_OK:
        msg = load().dem_last_error(ctx)
        raise DemError(code, msg.decode() if msg else "")
