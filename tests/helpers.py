"""Shared helpers for the parity tests (host <-> device copies, scenes with
live contacts, comparisons)."""
from __future__ import annotations

import numpy as np

from demso_amd import scene as sc


def dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a.copy()).cuda()


def host(t, like=np.float32) -> np.ndarray:
    a = t.detach().cpu().numpy()
    if like == np.uint32:
        a = a.view(np.uint32)
    return a


def compressed_pack(cols=32, rows=32, seed=7, spin=True) -> sc.Scene:
    """A column-collapse scene squeezed so that most discs overlap, with random
    velocities and spins, so every branch of the pair law runs (damping on
    approach, stick and slip shear)."""
    s = sc.column_collapse(seed=seed, cols=cols, rows=rows, pitch=0.0098, jitter=0.0012)
    rng = np.random.default_rng(seed)
    s.pos[:, 1] += 0.002
    s.vel[:] = (rng.normal(size=s.vel.shape) * 0.2).astype(np.float32)
    if spin:
        s.omega[:] = (rng.normal(size=s.n) * 30.0).astype(np.float32)
    s.theta[:] = (rng.uniform(0, 6.2, size=s.n)).astype(np.float32)
    return s


def mixed_pack(cols=32, rows=32, seed=11) -> sc.Scene:
    """compressed_pack with three materials of different stiffness, damping,
    friction and density assigned at random (per-particle radii kept), so the
    kernels read each particle's material row and every ordered pair of
    materials meets in contact."""
    import dataclasses

    s = compressed_pack(cols=cols, rows=rows, seed=seed)
    base = s.materials[0]
    s.materials = [
        base,
        dataclasses.replace(base, kn=base.kn * 3.0, kt=base.kt * 2.5, cn=base.cn * 0.5, mu=0.2, bdamp=0.3,
                            areal=base.areal * 1.7),
        dataclasses.replace(base, kn=base.kn * 0.4, kt=base.kt * 0.7, cn=base.cn * 1.5, mu=0.0, bdamp=0.8,
                            areal=base.areal * 0.6),
    ]
    s.type = np.random.default_rng(seed).integers(0, 3, s.n).astype(np.uint32)
    return s


def state_dict(s: sc.Scene) -> dict:
    return dict(pos=s.pos, vel=s.vel, theta=s.theta, omega=s.omega, id=s.id, type=s.type, loc=s.loc,
                radius=s.radius)


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max(|b|), per array (the 1e-5 relative bar)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    return float(np.abs(a - b).max() / scale)


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    if a.dtype in (np.float32,):
        return np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return np.array_equal(a, b)
