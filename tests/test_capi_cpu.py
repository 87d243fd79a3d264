"""The C ABI without a GPU: libdemso.so loads, exports every function that
include/demso.h declares, the Python binding declares exactly those, the
header is plain C, and creating a context without a device fails with a
status code and a message rather than a crash."""
import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from demso_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "demso.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(dem_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_functions():
    names = header_functions()
    assert "dem_calc_hash" in names and "dem_collide" in names and len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in header_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_matches_header():
    assert sorted(_lib.PROTOTYPES) == header_functions()


def test_header_is_c(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "demso.h"\nint main(void){ return dem_abi_version() == DEMSO_ABI_VERSION ? 0 : 1; }\n')
    res = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


def test_struct_layouts_match_header(tmp_path):
    """ctypes struct sizes equal the C compiler's."""
    src = tmp_path / "s.c"
    src.write_text('#include <stdio.h>\n#include "demso.h"\nint main(void){ printf("%zu %zu %zu\\n", '
                   'sizeof(dem_system_props), sizeof(dem_particle_props), sizeof(dem_instance)); '
                   'printf("%zu\\n", sizeof(dem_controlled)); return 0; }\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [ctypes.sizeof(_lib.SystemProps), ctypes.sizeof(_lib.ParticleProps),
                     ctypes.sizeof(_lib.InstanceProps), ctypes.sizeof(_lib.ControlledProps)]


def test_create_without_device_reports_error():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    lib = _lib.load()
    h = ctypes.c_void_p()
    rc = lib.dem_create(ctypes.byref(h), 0, None)
    assert rc == -2 and not h.value
    assert b"device" in lib.dem_last_error(None)
    assert lib.dem_destroy(None) == -1
    assert lib.dem_sim_step(None, 1) == -1


def test_reference_shim_signatures(tmp_path):
    """The C++ shim exposes the reference's wrapper signatures
    (src/functions.cuh:50-72) and links against libdemso.so."""
    src = tmp_path / "shim.cpp"
    src.write_text(r'''
#include "demso_reference_shim.hpp"
struct V2f { float x, y; };
struct V2u { unsigned x, y; };
struct Sys { unsigned numParticles; V2f cubeDimension; V2u gridSize; unsigned numCells; float timeStep;
             V2f gravity; float boundaryNormalStiffness, boundaryShearStiffness; };
struct Part { float radius, mass, inertia, normalStiffness, shearStiffness, normalDamping, boundaryDamping,
              frictionCoefficient, attractCoefficient, colorR, colorG, colorB; };
void (*f1)(float*, unsigned*, unsigned*, unsigned) = calcHash;
void (*f2)(unsigned*, unsigned*, unsigned) = sortParticles;
void (*f3)(unsigned*, unsigned*, float*, float*, float*, float*, unsigned*, unsigned*, unsigned*, unsigned*,
           unsigned*, float*, float*, float*, float*, unsigned*, unsigned*, unsigned, unsigned) = reorderDataAndFindCellStart;
void (*f4)(float*, float*, float*, float*, float*, unsigned*, unsigned*, unsigned*, unsigned, unsigned, float*) = collide;
void (*f5)(float*, float*, float*, float*, float*, float*, unsigned*, float*, unsigned) = integrateSystem;
void (*f6)(float*, float*, float*, float*, float*, unsigned*, unsigned*, unsigned*, unsigned, unsigned, V2f, V2f, float,
           float, unsigned, float*, float*, float*, float*, float*, float*, V2f*, float*, float*) = collide<V2f>;
int main() {
  Sys s{}; Part p[1]{};
  s.numParticles = 1; s.cubeDimension = {1, 1}; s.gridSize = {4, 4}; s.numCells = 16; s.timeStep = 1e-3f;
  int rc = demso_shim::init(s, p, 1);
  return (rc == DEM_OK || rc == DEM_ERR_NO_DEVICE) ? 0 : 1;
}
''')
    exe = tmp_path / "shim"
    lib_dir = ROOT / "demso_amd"
    res = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                          f"-L{lib_dir}", "-ldemso", f"-Wl,-rpath,{lib_dir}"], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True)
    assert run.returncode == 0, run.stderr


def test_torch_extension_loads():
    """The libtorch extension is built in-tree and imports without a GPU."""
    from demso_amd import torch_ext

    mod = torch_ext.load()
    assert hasattr(mod, "trajectory")
