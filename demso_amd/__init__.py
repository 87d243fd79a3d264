"""MI355X (gfx950) implementation of DemSo's 2D DEM time step.

The compute path is libdemso.so (HIP kernels behind the C ABI of
include/demso.h); this package holds its build script, the ctypes binding
(_lib), the Python mirror of the reference's step interface (api) and the
scene generators (scene).
"""
