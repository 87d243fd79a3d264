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
                   'sizeof(dem_system_props), sizeof(dem_particle_props), sizeof(dem_instance)); return 0; }\n')
    exe = tmp_path / "s"
    subprocess.run(["gcc", "-std=c99", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    sizes = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True).stdout.split()]
    assert sizes == [ctypes.sizeof(_lib.SystemProps), ctypes.sizeof(_lib.ParticleProps),
                     ctypes.sizeof(_lib.InstanceProps)]


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
