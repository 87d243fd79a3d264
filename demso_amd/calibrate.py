"""Calibration of contact parameters through the hand-written adjoint.

The loss of a run is the mean squared final position error against a target
run (per instance for an ensemble, summed over instances), differentiated
with respect to each material's (kn, kt, cn, mu) and the initial positions by
dem_adj_backward.  The calibrated variables are (kn, mu, e) per material
with kt = kt_ratio * kn and cn = damping_for_restitution(e, kn, m_ref); their
gradients follow from the chain rule:
  dL/dkn = g_kn + kt_ratio g_kt + g_cn cn/(2 kn)
  dL/dmu = g_mu
  dL/de  = g_cn dcn/de,  dcn/de = 2 sqrt(kn m_ref) / (de/dzeta)
and Adam (Kingma & Ba 2015) updates (log kn, mu, e).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from .scene import RHO, THICKNESS, Scene, damping_for_restitution, restitution_ratio, zeta_for_restitution


def reference_mass(r0: float = 0.005) -> float:
    return RHO * THICKNESS * math.pi * r0 * r0


def dcn_de(e: float, kn: float, m_ref: float) -> float:
    z = zeta_for_restitution(e)
    h = 1e-6
    de_dz = (restitution_ratio(z + h) - restitution_ratio(max(z - h, 0.0))) / (z + h - max(z - h, 0.0))
    return 2.0 * math.sqrt(kn * m_ref) / de_dz


def set_params(scene: Scene, kn, mu, e, kt_ratio: float = 0.2, m_ref: Optional[float] = None) -> Scene:
    """Copy of `scene` with material k given (kn[k], kt_ratio kn[k], cn(e[k]), mu[k])
    (boundary damping is left as it is: it is not a calibrated variable)."""
    m_ref = reference_mass() if m_ref is None else m_ref
    s = scene.copy()
    for k, m in enumerate(s.materials):
        m.kn = float(kn[k])
        m.kt = kt_ratio * float(kn[k])
        m.cn = damping_for_restitution(float(e[k]), float(kn[k]), m_ref)
        m.mu = float(mu[k])
    s.notes = dict(s.notes, restitution=[float(v) for v in e])
    return s


def settle(ctx, scene: Scene, steps: int) -> Scene:
    """Copy of `scene` with its state after `steps` plain steps on the GPU
    (a calibration run starts from a settled packing, where contacts act)."""
    from . import api

    sim = api.Simulation(ctx, scene)
    sim.step_graph(steps)
    st = sim.download(extra=False)
    sim.close()
    out = scene.copy()
    for k in ("pos", "vel", "theta", "omega", "id", "type", "loc", "radius"):
        if st[k] is not None:
            setattr(out, k, np.ascontiguousarray(st[k]))
    return out


def contact_rows(scene: Scene) -> np.ndarray:
    return np.array([[m.kn, m.kt, m.cn, m.mu] for m in scene.materials], dtype=np.float64)


def segments_of(scene: Scene) -> np.ndarray:
    return scene.segments if scene.segments is not None else np.array([0, scene.n])


@dataclass
class RunResult:
    loss: float
    grad_params: np.ndarray  # (num_types, 4): d loss / d(kn, kt, cn, mu)
    grad_pos0: np.ndarray    # (n, 2) by particle id
    final: np.ndarray        # (n, 2) by particle id
    seconds_forward: float
    seconds_backward: float


class Runner:
    """One simulation + adjoint reused across calibration iterations."""

    def __init__(self, ctx, scene: Scene, steps: int, checkpoint_every: int = 100):
        from . import api

        self.ctx = ctx
        self.scene = scene
        self.steps = steps
        self.sim = api.Simulation(ctx, scene)
        self.adj = api.Adjoint(self.sim, checkpoint_every=checkpoint_every)

    def close(self) -> None:
        self.adj.close()
        self.sim.close()

    def final_positions(self, scene: Scene) -> np.ndarray:
        """Forward only (plain step) from `scene`'s state and parameters."""
        self.ctx.set_contact_params(contact_rows(scene))
        self.sim.upload(scene)
        self.sim.step_graph(self.steps)
        return self.sim.get_by_id()[0].cpu().numpy()

    def run(self, scene: Scene, target: np.ndarray) -> RunResult:
        """Loss = sum over instances of the mean squared final position error."""
        import torch

        self.ctx.set_contact_params(contact_rows(scene))
        self.sim.upload(scene)
        t0 = time.perf_counter()
        self.adj.forward(self.steps)
        self.sim.synchronize()
        t1 = time.perf_counter()
        final = self.sim.get_by_id()[0].cpu().numpy()
        d = final.astype(np.float64) - target
        seg = segments_of(scene)
        loss = 0.0
        g = np.zeros_like(d)
        for a, b in zip(seg[:-1], seg[1:]):
            nk = max(int(b - a), 1)
            loss += float((d[a:b] ** 2).sum()) / nk
            g[a:b] = 2.0 * d[a:b] / nk
        gp = torch.from_numpy(g.astype(np.float32)).cuda()
        t2 = time.perf_counter()
        gp0, _, gpar = self.adj.backward(gp, None, num_types=len(scene.materials))
        self.sim.synchronize()
        t3 = time.perf_counter()
        return RunResult(loss, gpar.cpu().numpy(), gp0.cpu().numpy(), final, t1 - t0, t3 - t2)


def calibrate(ctx, scene: Scene, target_scene: Scene, steps: int, iters: int = 50, lr: float = 0.05,
              kt_ratio: float = 0.2, checkpoint_every: int = 100, log=None) -> dict:
    """Adam on (log kn, mu, e) per material, starting from `scene`'s materials,
    towards the final positions of `target_scene` (same discs, other
    parameters).  Returns the history and the recovered parameters."""
    m_ref = reference_mass()
    runner = Runner(ctx, scene, steps, checkpoint_every)
    target = runner.final_positions(target_scene).astype(np.float64)
    T = len(scene.materials)
    kn = np.array([m.kn for m in scene.materials])
    mu = np.array([m.mu for m in scene.materials])
    e = np.array(scene.notes.get("restitution", [0.5] * len(scene.materials)), dtype=np.float64)
    x = np.stack([np.log(kn), mu, e], 1)
    m1 = np.zeros_like(x)
    m2 = np.zeros_like(x)
    b1, b2, eps = 0.9, 0.999, 1e-12
    step_lr = np.array([lr, lr * 0.5, lr * 0.5])
    history = []
    for it in range(1, iters + 1):
        cur = set_params(scene, np.exp(x[:, 0]), x[:, 1], x[:, 2], kt_ratio, m_ref)
        res = runner.run(cur, target)
        g = np.zeros_like(x)
        for k in range(T):
            knk = math.exp(x[k, 0])
            cnk = cur.materials[k].cn
            gkn, gkt, gcn, gmu = res.grad_params[k]
            dkn = gkn + kt_ratio * gkt + gcn * cnk / (2.0 * knk)
            g[k, 0] = knk * dkn
            g[k, 1] = gmu
            g[k, 2] = gcn * dcn_de(float(x[k, 2]), knk, m_ref)
        # Adam, per-variable scale-free steps
        m1 = b1 * m1 + (1 - b1) * g
        m2 = b2 * m2 + (1 - b2) * g * g
        mh = m1 / (1 - b1 ** it)
        vh = m2 / (1 - b2 ** it)
        x = x - step_lr * mh / (np.sqrt(vh) + eps)
        x[:, 1] = np.clip(x[:, 1], 0.05, 1.0)
        x[:, 2] = np.clip(x[:, 2], 0.05, 0.99)
        history.append({"iter": it, "loss": res.loss, "kn": np.exp(x[:, 0]).tolist(), "mu": x[:, 1].tolist(),
                        "e": x[:, 2].tolist(), "t_fwd": res.seconds_forward, "t_bwd": res.seconds_backward})
        if log:
            log(history[-1])
    runner.close()
    return {"history": history, "kn": np.exp(x[:, 0]), "mu": x[:, 1], "e": x[:, 2]}
