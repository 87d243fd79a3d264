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
