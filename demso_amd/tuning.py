"""Autotuning of the step: cell-size multiplier (grid re-built on the host)
and per-kernel block sizes (dem_sim_autotune, timed on the device).

The cell size must stay >= the largest contact distance (2 rmax) so the 3x3
search finds every contact; larger cells trade fewer cells (smaller
cellStart/cellEnd arrays, fewer empty-cell visits) for more candidates per
disc.  A different grid changes the order in which a disc's contacts are
summed, so results differ from another multiplier's at rounding level; the
oracle is run with the same multiplier to check parity.
"""
from __future__ import annotations

import copy
import math
import time
from typing import Iterable, Optional

import numpy as np

from .scene import Instance, Scene


def max_contact_diameter(scene: Scene, k: Optional[int] = None) -> float:
    """2 rmax over all discs (k None) or over instance k's discs."""
    if scene.radius is None:
        radii = [m.radius for m in scene.materials]
        return 2.0 * max(radii)
    if k is None or scene.segments is None:
        return 2.0 * float(scene.radius.max())
    a, b = int(scene.segments[k]), int(scene.segments[k + 1])
    return 2.0 * float(scene.radius[a:b].max()) if b > a else 0.0


def with_cell_multiplier(scene: Scene, mult: float) -> Scene:
    """Copy of `scene` whose grids use cells of mult x (2 rmax) (>= 1), each
    instance keeping its box and covering at least the area its grid
    covered before (the hashing extent, which may reach above an open top);
    keys are re-based instance after instance."""
    if mult < 1.0:
        raise ValueError("the cell multiplier must be >= 1 (cells must hold the contact range)")
    out = copy.copy(scene)
    insts = []
    base = 0
    for k, g in enumerate(scene.instances):
        d = max_contact_diameter(scene, k if len(scene.instances) > 1 else None)
        cell = mult * d
        # cover the old hashing extent, not just the walls: discs above an
        # open top would otherwise all clamp into the top cell row
        span = (max(g.box[0], g.extent[0]), max(g.box[1], g.extent[1]))
        gx = max(1, int(math.ceil(span[0] / cell - 1e-9)))
        gy = max(1, int(math.ceil(span[1] / cell - 1e-9)))
        insts.append(Instance(box=g.box, extent=(gx * cell, gy * cell), grid=(gx, gy), cell_base=base))
        base += gx * gy
    out.instances = insts
    out.notes = dict(scene.notes, cell_multiplier=mult)
    return out


def time_graph_steps(sim, steps: int) -> float:
    """Seconds per step of `steps` graph-replayed steps (after one warm replay)."""
    sim.step_graph(2)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step_graph(steps)
    sim.synchronize()
    return (time.perf_counter() - t0) / steps


def autotune(ctx, scene: Scene, multipliers: Iterable[float] = (1.0, 1.25, 1.5, 2.0), steps: int = 20) -> dict:
    """Pick the cell multiplier and block sizes with the lowest step time.

    Each candidate grid gets its own simulation, block-size tuning and a
    timed graph replay of `steps` steps from the scene's initial state.
    Returns {"multiplier", "blocks", "seconds_per_step", "trials"}; the
    context keeps the winning block sizes.
    """
    from .api import Simulation

    trials = []
    for m in multipliers:
        s_m = with_cell_multiplier(scene, m)
        sim = Simulation(ctx, s_m)
        sim.step(1)
        blocks = sim.autotune()
        sec = time_graph_steps(sim, steps)
        trials.append({"multiplier": m, "cells": s_m.total_cells, "blocks": blocks, "seconds_per_step": sec})
        sim.close()
    best = min(trials, key=lambda t: t["seconds_per_step"])
    # leave the winning block sizes in the context
    s_best = with_cell_multiplier(scene, best["multiplier"])
    sim = Simulation(ctx, s_best)
    sim.step(1)
    sim.autotune()
    sim.close()
    return {"multiplier": best["multiplier"], "blocks": best["blocks"],
            "seconds_per_step": best["seconds_per_step"], "trials": trials}


def instance_scene(scene: Scene, k: int) -> Scene:
    """Instance k of an ensemble as a standalone single-system scene (its
    discs, its material as type 0, its box and grid with keys from 0)."""
    if scene.segments is None:
        raise ValueError("not an ensemble")
    a, b = int(scene.segments[k]), int(scene.segments[k + 1])
    g = scene.instances[k]
    m = copy.copy(scene.materials[k])
    m.instance = 0
    n = b - a
    ids = np.arange(n, dtype=np.uint32)
    return Scene(name=f"{scene.name}[{k}]", pos=scene.pos[a:b].copy(), vel=scene.vel[a:b].copy(),
                 theta=scene.theta[a:b].copy(), omega=scene.omega[a:b].copy(), id=ids,
                 type=np.zeros(n, np.uint32), loc=ids.copy(),
                 radius=None if scene.radius is None else scene.radius[a:b].copy(), materials=[m],
                 instances=[Instance(box=g.box, extent=g.extent, grid=g.grid, cell_base=0)], dt=scene.dt,
                 gravity=scene.gravity, bkn=scene.bkn, bkt=scene.bkt, flags=scene.flags, steps=scene.steps)
