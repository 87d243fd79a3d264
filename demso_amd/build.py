"""Build the gfx950 HIP library (libdemso.so) and the headless driver
(demso_run) in-tree.

Every HIP source is compiled with:
  -ffp-contract=off                       no FMA contraction: the FP32 arithmetic
                                          is the plain IEEE sequence of the
                                          reference kernels
  -target-feature -packed-fp32-ops        no packed FP32 instructions
Division and sqrt stay correctly rounded (the HIP default).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libdemso.so"
ARCH = os.environ.get("DEMSO_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or Path(cand).exists()):
            return cand
    raise RuntimeError("hipcc not found")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build_library(verbose: bool = False) -> Path:
    """Compile all csrc/*.hip into demso_amd/libdemso.so (always rebuilds)."""
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-o", str(tmp), *map(str, sources())]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the GPU-only feature flag.
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


DRIVER = PKG / "demso_run"
DRIVER_SOURCES = ["xml.cpp", "dsml.cpp", "prepare.cpp", "demso_run.cpp"]


def build_driver(verbose: bool = False) -> Path:
    """Compile the headless .dsml driver (host C++ over libdemso.so and the
    HIP runtime) into demso_amd/demso_run."""
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    tmp = DRIVER.with_name("demso_run.tmp")
    # -ffp-contract=off: the rectangle draws (csrc/scene_rng.h) must round as
    # the device generator does
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           f"-I{rocm}/include", f"-I{ROOT / 'include'}", f"-I{CSRC}",
           *[str(PKG / "driver" / f) for f in DRIVER_SOURCES],
           f"-L{PKG}", "-ldemso", f"-L{rocm}/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"driver build failed ({res.returncode}):\n{res.stderr}")
    os.replace(tmp, DRIVER)
    return DRIVER


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    odir = ROOT / "oracle"
    res = subprocess.run(["make", "-C", str(odir), "-B", "liboracle.so"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(res.stdout.strip())
    return odir / "liboracle.so"


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_driver(verbose=True))
    print(build_oracle(verbose=True))
