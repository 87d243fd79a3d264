"""The libtorch extension exposing the adjoint as an autograd Function, and
a calibration helper built on it.

build() compiles demso_amd/csrc/torch_adjoint.cpp against libtorch and
libdemso.so into demso_amd/_torch_build/ (in-tree); load() imports that
module without recompiling.
"""
from __future__ import annotations

import importlib.machinery
import importlib.util
import math
import sys
from pathlib import Path
from typing import Optional, Sequence

import numpy as np

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
BUILD_DIR = PKG / "_torch_build"
NAME = "demso_torch"
_mod = None


def build(verbose: bool = False):
    """Compile the extension (always) and return the loaded module."""
    import torch.utils.cpp_extension as ce

    BUILD_DIR.mkdir(exist_ok=True)
    mod = ce.load(
        name=NAME,
        sources=[str(PKG / "csrc" / "torch_adjoint.cpp")],
        extra_include_paths=[str(ROOT / "include")],
        extra_cflags=["-O2"],
        extra_ldflags=[f"-L{PKG}", "-ldemso", f"-Wl,-rpath,{PKG}"],
        build_directory=str(BUILD_DIR),
        verbose=verbose,
    )
    global _mod
    _mod = mod
    return mod


def load():
    """Import the built extension (fails loudly if it was not built)."""
    global _mod
    if _mod is not None:
        return _mod
    import torch  # noqa: F401  (libtorch must be loaded first)

    so = BUILD_DIR / f"{NAME}.so"
    if not so.exists():
        raise ImportError(f"{so} is not built; run demso_amd.torch_ext.build()")
    loader = importlib.machinery.ExtensionFileLoader(NAME, str(so))
    spec = importlib.util.spec_from_file_location(NAME, str(so), loader=loader)
    mod = importlib.util.module_from_spec(spec)
    loader.exec_module(mod)
    sys.modules[NAME] = mod
    _mod = mod
    return mod


class Trajectory:
    """A differentiable run of `steps` DEM steps from a scene: calling it with
    initial positions (n, 2 by id, float32 GPU, requires_grad allowed) and
    contact parameters ((num_types, 4) float64: kn, kt, cn, mu) returns the
    final positions by id; backward goes through the hand-written adjoint."""

    def __init__(self, ctx, scene, steps: int, checkpoint_every: int = 100):
        from . import api

        self.mod = load()
        self.ctx = ctx
        self.scene = scene
        self.steps = steps
        self.sim = api.Simulation(ctx, scene)
        self.adj = api.Adjoint(self.sim, checkpoint_every=checkpoint_every)

    def __call__(self, pos0, params):
        # restart from the scene's initial state (velocities, spins, order)
        self.sim.upload(self.scene)
        return self.mod.trajectory(pos0, params, self.ctx.handle, self.sim.handle, self.adj.handle, self.steps)

    def close(self):
        self.adj.close()
        self.sim.close()


def contact_params(kn: float, mu: float, e: float, m_ref: float, kt_ratio: float = 0.2):
    """(kn, kt, cn, mu) from stiffness, friction and restitution (kt tied to kn)."""
    from .scene import damping_for_restitution

    return [kn, kt_ratio * kn, damping_for_restitution(e, kn, m_ref), mu]
