"""Readers for the headless driver's files (demso_amd/demso_run).

State dump: "DEMSODMP", u32 version (1), u32 n, u64 step, f64 time, then in
particle-ID order pos f32[n][2], vel f32[n][2], theta f32[n], omega f32[n],
type u32[n], acc f32[n][2], alpha f32[n], pressure f32[n].

Output file (writeOutputFile's format, src/functions.cu:622-660): one line
per step, the elapsed time, then 12 values per followed particle (id, type,
pos.x, pos.y, vel.x, vel.y, acc.x, acc.y, theta, omega, alpha, pressure).
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

MAGIC = b"DEMSODMP"


def read_dump(path) -> dict:
    data = Path(path).read_bytes()
    if data[:8] != MAGIC:
        raise ValueError(f"{path}: not a state dump")
    version, n, step, time = struct.unpack_from("<IIQd", data, 8)
    if version != 1:
        raise ValueError(f"{path}: dump version {version}")
    off = 8 + struct.calcsize("<IIQd")
    out = {"n": n, "step": step, "time": time}
    layout = [("pos", np.float32, (n, 2)), ("vel", np.float32, (n, 2)), ("theta", np.float32, (n,)),
              ("omega", np.float32, (n,)), ("type", np.uint32, (n,)), ("acc", np.float32, (n, 2)),
              ("alpha", np.float32, (n,)), ("pressure", np.float32, (n,))]
    for name, dt, shape in layout:
        count = int(np.prod(shape))
        out[name] = np.frombuffer(data, dtype=np.dtype(dt).newbyteorder("<"), count=count, offset=off).reshape(shape)
        off += 4 * count
    if off != len(data):
        raise ValueError(f"{path}: {len(data) - off} trailing bytes")
    return out


def read_output(path) -> tuple[np.ndarray, np.ndarray]:
    """(times (steps,), rows (steps, followed, 12) float64)."""
    times, rows = [], []
    for line in Path(path).read_text().splitlines():
        vals = [float(v) for v in line.split(",")]
        times.append(vals[0])
        rows.append(np.array(vals[1:]).reshape(-1, 12))
    return np.array(times), np.array(rows)
