"""Seeded inputs of the parity tests and benchmarks: the dam-break tanks of the fluid step.

Dam break (the reference's headless dam-break workload): one block of water, x < 0.4 W,
y < 0.6 H, on a 2 x 2 jittered lattice per cell (fill_rect, maps.rs:468-490), in the solid
tank frame (scene.rs:8, WALL = 3 cells), zero initial velocity, fixed dt 1/240 s.
"""
import numpy as np

from sprout_amd import _native as N
from sprout_amd import scene as sc


def dam_params(cell_size=0.005, viscosity=0.0, pressure_tol=1e-6, fixed_hz=240.0):
    p = N.default_params()
    p.cell_size = cell_size
    p.fixed_hz = fixed_hz
    p.pressure_tol = pressure_tol
    p.mat[N.ROW_WATER].viscosity = viscosity
    return p


def settle(parts, params):
    """Particle::settle_spawn (params.rs:70-88) for ids 0..n, isothermal rows."""
    q = parts.copy()
    q["id"] = np.arange(len(q), dtype=np.uint32)
    rows = (q["kind"] & 255).astype(np.int64) - 1
    for row in np.unique(rows):
        m = params.mat[int(row)]
        sel = rows == row
        t = q["t"][sel]
        q["t"][sel] = np.where(t == N.SPAWN_TEMP, np.float32(m.spawn_temp), t)
        sv = q["s"][sel]
        load = np.float32(min(max(m.load, 0.0), 1.0)) if m.diluent >= 0 else np.float32(1.0)
        q["s"][sel] = np.where(sv == N.SPAWN_S, load, sv)
    return q


def dam_break(grid, frac=(0.4, 0.6), **kw):
    """(params, solid field, settled particles, rows) of a dam break on `grid`."""
    params = dam_params(**kw)
    scene, parts = sc.dam_break(grid, frac=frac)
    field = scene.rasterize()
    parts = settle(parts, params)
    return params, field, parts, [N.ROW_WATER]
