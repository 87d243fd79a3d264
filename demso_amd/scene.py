"""Deterministic scenes: initial particle state + parameter tables.

A Scene carries everything one step needs: the SoA state (float2 pos/vel as
(n,2) float32, theta/omega/radius float32, id/type/loc uint32), the material
rows (one per type), the instance geometry (box, hashing extent, grid, key
base) and the global scalars.  The HIP path (demso_amd.api) and the CPU
oracle both consume the same Scene, so parity runs start from bit-identical
inputs.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014; the output
mix of Vigna's splitmix64.c) in counter form: value i of stream s under seed
k is mix(k*G + s*H + (i+1)*G), G = 0x9E3779B97F4A7C15,
H = 0xD1B54A32D192ED03.  Uniforms take the top 53 bits; normals use
Box-Muller.

Scenes for the workloads:
  column_collapse  4,096 discs, 64x64 lattice at 0.0115 m in a 1.0 x 0.5 m box
  settling         2048x2048 log-normal discs, box 2.2 W x 1.1 H
  ensemble         256 column-collapse instances of 2,000..60,000 discs
  triangle_pile    the reference's pilha.dsml pile (type radii, wrap grid)
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import DEM_GRID_CLAMP, DEM_OPEN_TOP

_G = np.uint64(0x9E3779B97F4A7C15)
_H = np.uint64(0xD1B54A32D192ED03)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)

RHO = 2650.0       # kg/m^3
THICKNESS = 0.01   # m (disc thickness)


def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    with np.errstate(over="ignore"):
        base = np.uint64(seed) * _G + np.uint64(stream) * _H
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * _G)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """U[0,1) doubles from the top 53 bits."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1000003, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * math.pi * u2)


def _series_log(x: np.ndarray) -> np.ndarray:
    """ln x on [2^-53, 1] with correctly rounded operations only: x = f 2^e,
    f in [sqrt(1/2), sqrt(2)), ln f = 2 atanh((f-1)/(f+1)) to t^21, e ln 2 in
    fdlibm's two parts.  The same sequence as the device and host generators,
    so the three agree bit for bit."""
    f, e = np.frexp(x)
    low = f < 0.70710678118654752440
    f = np.where(low, f * 2.0, f)
    e = np.where(low, e - 1, e)
    t = (f - 1.0) / (f + 1.0)
    t2 = t * t
    p = np.full_like(t, 1.0 / 21.0)
    for k in (19.0, 17.0, 15.0, 13.0, 11.0, 9.0, 7.0, 5.0, 3.0):
        p = p * t2 + 1.0 / k
    p = p * t2 + 1.0
    lnf = 2.0 * t * p
    de = e.astype(np.float64)
    return de * 6.93147180369123816490e-01 + (de * 1.90821492927058770002e-10 + lnf)


_SIN = (-1.0 / 121645100408832000.0, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0, 1.0 / 6227020800.0,
        -1.0 / 39916800.0, 1.0 / 362880.0, -1.0 / 5040.0, 1.0 / 120.0, -1.0 / 6.0, 1.0)
_COS = (1.0 / 6402373705728000.0, -1.0 / 20922789888000.0, 1.0 / 87178291200.0, -1.0 / 479001600.0,
        1.0 / 3628800.0, -1.0 / 40320.0, 1.0 / 720.0, -1.0 / 24.0, 1.0 / 2.0)


def _series_cos2pi(u: np.ndarray) -> np.ndarray:
    """cos(2 pi u), u in [0, 1): exact reductions to v in [0, 1/8] turns, then
    the Taylor series of sin or cos at x = 2 pi v (the generators' sequence)."""
    u = np.where(u >= 0.5, 1.0 - u, u)
    neg = u > 0.25
    u = np.where(neg, 0.5 - u, u)
    sign = np.where(neg, -1.0, 1.0)
    use_sin = u > 0.125
    v = np.where(use_sin, 0.25 - u, u)
    x = 6.28318530717958647692 * v
    x2 = x * x
    ps = np.full_like(x, _SIN[0])
    for c in _SIN[1:]:
        ps = ps * x2 + c
    pc = np.full_like(x, _COS[0])
    for c in _COS[1:]:
        pc = pc * x2 + c
    return sign * np.where(use_sin, x * ps, 1.0 - x2 * pc)


def jitter_normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Normal draws 0..n-1 of the scene generators (Box-Muller from streams
    `stream` and stream + 1000003, in correctly rounded operations only)."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1000003, n)
    return np.sqrt(-2.0 * _series_log(1.0 - u1)) * _series_cos2pi(u2)


def rectangle_layout(start, end, num, types, seed: int = 0, stream: int = 1000):
    """Positions and type indices of a DemSo rectangle block
    (createRectanglesD, src/particles_kernel.cuh:43-74): particle
    p = x + y*num[0] at start + (end-start)/(num-1) * ((x, y) + (g - 0.5)/100)
    in FP32, g the normal draws p of streams `stream` (x) and stream + 1 (y);
    type types[(x + y) % len(types)]."""
    f32 = np.float32
    nx_, ny_ = int(num[0]), int(num[1])
    n = nx_ * ny_
    p = np.arange(n)
    x = (p % nx_).astype(np.uint32)
    y = (p // nx_).astype(np.uint32)
    gx = jitter_normal(seed, stream, n).astype(f32)
    gy = jitter_normal(seed, stream + 1, n).astype(f32)
    pos = np.empty((n, 2), f32)
    with np.errstate(divide="ignore", invalid="ignore"):  # a single row or column divides by zero, as DemSo does
        dx = f32(f32(end[0]) - f32(start[0])) / f32(nx_ - 1)
        dy = f32(f32(end[1]) - f32(start[1])) / f32(ny_ - 1)
        pos[:, 0] = f32(start[0]) + dx * (x.astype(f32) + (gx - f32(0.5)) / f32(100))
        pos[:, 1] = f32(start[1]) + dy * (y.astype(f32) + (gy - f32(0.5)) / f32(100))
    tv = np.asarray(types, np.uint32)
    return pos, tv[(x + y) % len(tv)]


def restitution_ratio(zeta: float) -> float:
    """Coefficient of restitution of a linear spring with a dashpot that acts
    only while the discs approach (src/particles_kernel.cuh:281): the
    compression phase is a damped oscillator from overlap 0 at speed v0 to
    zero speed at overlap dmax = (v0/w) exp(-zeta g), the restitution phase
    is undamped, so e = w dmax / v0 = exp(-zeta g(zeta)) with
    g = acos(z)/sqrt(1-z^2) (z<1), 1 (z=1), acosh(z)/sqrt(z^2-1) (z>1)."""
    if zeta <= 0.0:
        return 1.0
    if abs(zeta - 1.0) < 1e-12:
        g = 1.0
    elif zeta < 1.0:
        g = math.acos(zeta) / math.sqrt(1.0 - zeta * zeta)
    else:
        g = math.acosh(zeta) / math.sqrt(zeta * zeta - 1.0)
    return math.exp(-zeta * g)


def zeta_for_restitution(e: float) -> float:
    if not (0.0 < e <= 1.0):
        raise ValueError("restitution must be in (0, 1]")
    if e == 1.0:
        return 0.0
    lo, hi = 0.0, 1.0
    while restitution_ratio(hi) > e:
        hi *= 2.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if restitution_ratio(mid) > e:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def damping_for_restitution(e: float, kn: float, m: float) -> float:
    """Per-type normal damping c giving restitution e between two identical
    discs of mass m.  The reference combines per-type values in series
    (src/particles_kernel.cuh:269-274), so the pair sees k/2, c/2 and
    m_eff = m/2: zeta = (c/2) / (2 sqrt(k/2 * m/2)) = c / (2 sqrt(k m))."""
    return 2.0 * zeta_for_restitution(e) * math.sqrt(kn * m)


@dataclass
class Material:
    radius: float = 0.0
    mass: float = 0.0
    inertia: float = 0.0
    kn: float = 0.0
    kt: float = 0.0
    cn: float = 0.0
    bdamp: float = 0.0
    mu: float = 0.0
    attract: float = 0.0
    areal: float = 0.0
    instance: int = 0
    color: Tuple[float, float, float] = (255.0, 255.0, 255.0)  # colorR/G/B, 0..255


@dataclass
class Instance:
    box: Tuple[float, float]
    extent: Tuple[float, float]
    grid: Tuple[int, int]
    cell_base: int = 0

    @property
    def cells(self) -> int:
        return self.grid[0] * self.grid[1]


@dataclass
class Scene:
    name: str
    pos: np.ndarray
    vel: np.ndarray
    theta: np.ndarray
    omega: np.ndarray
    id: np.ndarray
    type: np.ndarray
    loc: np.ndarray
    radius: Optional[np.ndarray]
    materials: List[Material]
    instances: List[Instance]
    dt: float
    gravity: Tuple[float, float] = (0.0, -9.81)
    bkn: float = 0.0
    bkt: float = 0.0
    flags: int = DEM_GRID_CLAMP
    steps: int = 1
    segments: Optional[np.ndarray] = None  # ensemble: first particle of each instance (+ n)
    notes: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def total_cells(self) -> int:
        return max(i.cell_base + i.cells for i in self.instances)

    def copy(self) -> "Scene":
        import copy as _copy

        return _copy.deepcopy(self)


def _state(pos: np.ndarray, radius: Optional[np.ndarray], types: np.ndarray) -> dict:
    n = pos.shape[0]
    ids = np.arange(n, dtype=np.uint32)
    return dict(
        pos=np.ascontiguousarray(pos, dtype=np.float32),
        vel=np.zeros((n, 2), np.float32),
        theta=np.zeros(n, np.float32),
        omega=np.zeros(n, np.float32),
        id=ids,
        type=np.ascontiguousarray(types, dtype=np.uint32),
        loc=ids.copy(),
        radius=None if radius is None else np.ascontiguousarray(radius, dtype=np.float32),
    )


def disc_material(kn: float, kt: float, e: float, mu: float, r0: float, instance: int = 0) -> Material:
    """Material row for the soil scenes: per-particle radii, m = rho*t*pi*r^2,
    damping from restitution at the mean radius r0, walls reflect with e."""
    areal = RHO * THICKNESS * math.pi
    m0 = areal * r0 * r0
    return Material(kn=kn, kt=kt, cn=damping_for_restitution(e, kn, m0), bdamp=e, mu=mu,
                    areal=areal, instance=instance)


def _grid_for(box: Tuple[float, float], cell: float) -> Tuple[Tuple[int, int], Tuple[float, float]]:
    gx = int(math.ceil(box[0] / cell - 1e-9))
    gy = int(math.ceil(box[1] / cell - 1e-9))
    return (gx, gy), (gx * cell, gy * cell)


def _lattice(seed: int, stream: int, cols: int, rows: int, n: int, pitch: float, jitter: float,
             origin: Tuple[float, float]) -> np.ndarray:
    k = np.arange(n)
    cx = (k % cols).astype(np.float64)
    cy = (k // cols).astype(np.float64)
    jx = (uniform(seed, stream, n) * 2.0 - 1.0) * jitter
    jy = (uniform(seed, stream + 1, n) * 2.0 - 1.0) * jitter
    x = origin[0] + (cx + 0.5) * pitch + jx
    y = origin[1] + (cy + 0.5) * pitch + jy
    return np.stack([x, y], axis=1)


def column_collapse(seed: int = 0, cols: int = 64, rows: int = 64, n: Optional[int] = None,
                    box: Tuple[float, float] = (1.0, 0.5), pitch: float = 0.0115,
                    jitter: float = 0.0005, rmin: float = 0.0045, rmax: float = 0.0055,
                    cell: float = 0.011, kn: float = 1e5, kt: float = 2e4, e: float = 0.5,
                    mu: float = 0.5, dt: float = 1e-5, steps: int = 10000) -> Scene:
    """Plain DEM step workload: a jittered square lattice against the left wall
    of a 1.0 m x 0.5 m box; radii U[0.0045, 0.0055]; 91 x 46 cells of 0.011 m.
    The 64-row lattice is 0.74 m tall, taller than the box, so the top is
    open (no top wall) and rows above the grid clamp into its top row."""
    n = cols * rows if n is None else n
    pos = _lattice(seed, 0, cols, rows, n, pitch, jitter, (0.0, 0.0))
    radius = rmin + (rmax - rmin) * uniform(seed, 2, n)
    grid, extent = _grid_for(box, cell)
    mat = disc_material(kn, kt, e, mu, 0.5 * (rmin + rmax))
    st = _state(pos, radius, np.zeros(n, np.uint32))
    return Scene(name="column_collapse", materials=[mat], dt=dt, bkn=kn, bkt=kt,
                 instances=[Instance(box=box, extent=extent, grid=grid)],
                 flags=DEM_GRID_CLAMP | DEM_OPEN_TOP, steps=steps, **st)


def calibration_column(seed: int = 3, cols: int = 256, rows: int = 256, kn: float = 1e5, kt: float = 2e4,
                       e: float = 0.5, mu: float = 0.5, dt: float = 1e-5, steps: int = 2000,
                       wall_restitution: float = 0.5) -> Scene:
    """Column collapse with config-0 distributions at any lattice size (the
    adjoint calibration workload: 256 x 256 = 65,536 discs): the config-0 box
    scaled by cols/64, cells of 0.011 m covering the box and the lattice
    above its open top, one material (kt, cn from e), boundary damping
    fixed so it is not confounded with the calibrated restitution."""
    pitch, jitter = 0.0115, 0.0005
    n = cols * rows
    box = (1.0 * cols / 64.0, 0.5 * cols / 64.0)
    pos = _lattice(seed, 0, cols, rows, n, pitch, jitter, (0.0, 0.0))
    radius = 0.0045 + 0.001 * uniform(seed, 2, n)
    grid, extent = _grid_for((box[0], max(box[1], rows * pitch + jitter)), 0.011)
    mat = disc_material(kn, kt, e, mu, 0.005)
    mat.bdamp = wall_restitution
    st = _state(pos, radius, np.zeros(n, np.uint32))
    return Scene(name="calibration_column", materials=[mat], dt=dt, bkn=kn, bkt=kt,
                 instances=[Instance(box=box, extent=extent, grid=grid)],
                 flags=DEM_GRID_CLAMP | DEM_OPEN_TOP, steps=steps, notes={"restitution": [e]}, **st)


def settling(seed: int = 1, side: int = 2048, r0: float = 0.005, sigma: float = 0.2,
             jitter: float = 0.0005, kn: float = 1e5, kt: float = 2e4, e: float = 0.5,
             mu: float = 0.5, dt: float = 1e-5, steps: int = 1000, pitch: Optional[float] = None) -> Scene:
    """Plain DEM step at scale: side x side jittered lattice, radii
    r0*exp(sigma*z) clipped to [r0/2, 2 r0]; lattice pitch 1.05 x the largest
    diameter (1.05 * 2 * 2 r0), so no disc overlaps at t=0; box 2.2 x the
    lattice width and 1.1 x its height; cells of 2*rmax.  For side=2048 this
    is 4,194,304 discs on 4731 x 2366 = 11,193,546 cells.  Another `pitch`
    packs the same discs closer (settling_dense: the mean diameter)."""
    n = side * side
    rmax = 2.0 * r0
    pitch = 1.05 * 2.0 * rmax if pitch is None else pitch
    width = side * pitch
    box = (2.2 * width, 1.1 * width)
    radius = np.clip(r0 * np.exp(sigma * normal(seed, 2, n)), 0.5 * r0, rmax)
    pos = _lattice(seed, 0, side, side, n, pitch, jitter, (0.0, 0.0))
    grid, extent = _grid_for(box, 2.0 * rmax)
    mat = disc_material(kn, kt, e, mu, r0)
    st = _state(pos, radius, np.zeros(n, np.uint32))
    return Scene(name="settling", materials=[mat], dt=dt, bkn=kn, bkt=kt,
                 instances=[Instance(box=box, extent=extent, grid=grid)],
                 flags=DEM_GRID_CLAMP, steps=steps, **st)


def settling_dense(seed: int = 1, side: int = 2048, r0: float = 0.005, **kw) -> Scene:
    """The settling workload's discs (same radii and law) on a lattice at the
    mean diameter 2*r0 instead of 1.05 x the largest: neighbours whose radii
    add up to more than the pitch overlap (about half the lattice pairs), so
    every step runs the contact law for about two contacts a disc and a disc
    has about 4 others per cell (36 candidates).  For side=2048: 4,194,304
    discs on 2253 x 1127 cells of 2 rmax."""
    s = settling(seed=seed, side=side, r0=r0, pitch=2.0 * r0, **kw)
    s.name = "settling_dense"
    return s


def ensemble(seed: int = 2, instances: int = 256, nmin: int = 2000, nmax: int = 60000,
             kn_range: Tuple[float, float] = (1e4, 1e6), mu_range: Tuple[float, float] = (0.2, 0.8),
             e_range: Tuple[float, float] = (0.3, 0.9), pitch: float = 0.0115, jitter: float = 0.0005,
             rmin: float = 0.0045, rmax: float = 0.0055, cell: float = 0.011, dt: float = 1e-5,
             steps: int = 2000, counts: Optional[Sequence[int]] = None) -> Scene:
    """Batched ensemble: independent column-collapse instances, each with its
    own particle count, box (config-0 box scaled by the lattice side / 64),
    grid (cells of 0.011 m covering the box and the lattice above its open
    top) and material (kn log-uniform, kt = 0.2 kn, mu and restitution
    uniform).  Instance k's particles are the contiguous index range
    segments[k]..segments[k+1]-1 and its type is k."""
    if counts is None:
        u = uniform(seed, 10, instances)
        counts = (nmin + np.floor(u * (nmax - nmin + 1))).astype(np.int64)
    counts = np.asarray(counts, dtype=np.int64)
    uk, umu, ue = uniform(seed, 11, instances), uniform(seed, 12, instances), uniform(seed, 13, instances)
    r0 = 0.5 * (rmin + rmax)
    pos_parts, rad_parts, type_parts, mats, insts, es = [], [], [], [], [], []
    base = 0
    for k, nk in enumerate(counts.tolist()):
        side = int(math.ceil(math.sqrt(nk)))
        rows = int(math.ceil(nk / side))
        scale = side / 64.0
        box = (1.0 * scale, 0.5 * scale)
        # the open-top lattice stands above the box: the hashing grid covers it
        # (cells, not walls) so the upper rows do not crowd the top cell row
        grid, extent = _grid_for((box[0], max(box[1], rows * pitch + jitter)), cell)
        kn = math.exp(math.log(kn_range[0]) + uk[k] * (math.log(kn_range[1]) - math.log(kn_range[0])))
        mu = mu_range[0] + umu[k] * (mu_range[1] - mu_range[0])
        e = e_range[0] + ue[k] * (e_range[1] - e_range[0])
        mats.append(disc_material(kn, 0.2 * kn, e, mu, r0, instance=k))
        es.append(e)
        insts.append(Instance(box=box, extent=extent, grid=grid, cell_base=base))
        base += grid[0] * grid[1]
        pos_parts.append(_lattice(seed, 100 + 4 * k, side, rows, nk, pitch, jitter, (0.0, 0.0)))
        rad_parts.append(rmin + (rmax - rmin) * uniform(seed, 102 + 4 * k, nk))
        type_parts.append(np.full(nk, k, np.uint32))
    st = _state(np.concatenate(pos_parts), np.concatenate(rad_parts), np.concatenate(type_parts))
    seg = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    # one global wall stiffness for all instances (SystemProperties holds one):
    # stiff enough that the series combination with each instance's kt is ~kt
    return Scene(name="ensemble", materials=mats, instances=insts, dt=dt, bkn=1e6, bkt=1e6,
                 flags=DEM_GRID_CLAMP | DEM_OPEN_TOP, steps=steps, segments=seg,
                 notes={"restitution": es}, **st)


def reference_grid_size(box: Tuple[float, float], max_radius: float) -> Tuple[int, int]:
    """Power-of-two grid sizing of PrepareSim (src/main.cu:218-236), including
    its use of the box width for the y cell size (main.cu:231)."""
    f32 = np.float32

    def axis(length: float) -> int:
        grid = int(f32(length) / (f32(4.0) * f32(max_radius)))
        temp = int(math.log2(grid))
        cell = f32(box[0]) / f32(1 << temp)
        if cell / f32(2.0) <= f32(1.2) * f32(max_radius):
            temp -= 1
        elif cell / f32(2.0) >= f32(3.0) * f32(max_radius):
            temp += 1
        return 1 << temp

    return axis(box[0]), axis(box[1])


def triangle_layout(start, N: int, space, height, types, count: Optional[int] = None):
    """Positions and type indices of a DemSo triangle block
    (createTrianglesD, src/particles_kernel.cuh:76-111) in its FP32
    arithmetic: particle `index` sits in row L (from the base) and column C,
    L1 = int((1 + 2 sqrtf(N^2 + N - 2(index+1) + 1)) / 2), L = N - L1,
    s = (L1 + N + N^2 - L1^2) / 2, C = index - s + L1, at
    (start.x + (L + 2C - (N-1)) space / 2, start.y + L height); type
    types[(C + L) % len(types)]."""
    f32 = np.float32
    count = N * (N + 1) // 2 if count is None else count
    index = np.arange(count, dtype=np.int64)
    radicand = (N * N + N - 2 * (index + 1) + 1).astype(f32)
    L1 = ((f32(1) + f32(2) * np.sqrt(radicand)) / f32(2)).astype(np.int64)
    L = N - L1
    sidx = (L1 + N + N * N - L1 * L1) // 2
    C = index - sidx + L1
    A = N - 1
    pos = np.empty((count, 2), f32)
    pos[:, 0] = f32(start[0]) + (L + 2 * C - A).astype(f32) * f32(space) / f32(2)
    pos[:, 1] = f32(start[1]) + L.astype(f32) * f32(height)
    tv = np.asarray(types, np.uint32)
    return pos, tv[(C + L) % len(tv)]


def triangle_pile(rows: int = 16, start: Tuple[float, float] = (5.0, 0.125), width: float = 3.75,
                  radius: float = 0.125, mass: float = 1e-2, kn: float = 1000.0, kt: float = 1000.0,
                  cn: float = 0.2, bdamp: float = 0.5, mu: float = 0.6, attract: float = 10.0,
                  box: Tuple[float, float] = (10.0, 10.0), bkn: float = 1000.0, bkt: float = 100.0,
                  dt: float = 1e-3, steps: int = 1000, controlled: bool = False) -> Scene:
    """The pile of exemplos/pilha.dsml: one particle type with table
    radius/mass (inertia 2/5 m r^2, src/main.cu:193), the triangle layout of
    createTrianglesD (src/particles_kernel.cuh:76-111) and the wrapping
    power-of-two grid of PrepareSim.  controlled: add the scene's "MOVE"
    type (material row 1) and note the controlled particle that DemSo's
    default build drives with it (start (5, 9.5), 0.005 down per step,
    src/main.cu:141-149, 510-514); the grid ignores that type's radius, as
    PrepareSim does."""
    f32 = np.float32
    N = rows
    space = f32(width) / f32(N - 1)
    height = f32(float(space) * math.sqrt(3.0) / 2.0)
    pos, _ = triangle_layout(start, N, space, height, [0])
    count = pos.shape[0]
    inertia = float(f32(2) * f32(mass) * f32(radius) * f32(radius) / f32(5))
    mat = Material(radius=radius, mass=mass, inertia=inertia, kn=kn, kt=kt, cn=cn, bdamp=bdamp, mu=mu,
                   attract=attract)
    grid = reference_grid_size(box, radius)
    st = _state(pos, None, np.zeros(count, np.uint32))
    mats = [mat]
    notes = {}
    if controlled:
        mc, rc = 100.0, 0.5
        mats.append(Material(radius=rc, mass=mc, inertia=float(f32(2) * f32(mc) * f32(rc) * f32(rc) / f32(5)),
                             kn=1e20, kt=1e20, cn=0.2, bdamp=0.5, mu=0.4, attract=0.0))
        notes["controlled"] = dict(pos=(5.0, 9.5), type=1, move=(0.0, -0.005), wrap_below_y=-1.0,
                                   wrap_above_x=float(f32(box[0])) + 0.5, reset_y=10.5, reset_x=-0.5)
    return Scene(name="triangle_pile", materials=mats, dt=dt, bkn=bkn, bkt=bkt,
                 instances=[Instance(box=box, extent=box, grid=grid)], flags=0, steps=steps, notes=notes, **st)
