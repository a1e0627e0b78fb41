"""Pins of the oracle's evaporation restatement (particles.comp:771-807, the volatility of
MatProps, the ledger's counts of mod.rs:3400-3404), on the CPU:

  * only particles in a cell with an open-air neighbour (air, not solid) can leave;
  * each leaves with probability rate x dt, rate = volatility x 2^((T - 20) / 10) (a
    concentrate's by its concentration): over a pool's surface the share that leaves per
    substep is that probability, within the binomial spread;
  * the ledger counts every particle that left, per row, and a reset clears the counts;
  * a liquid that fumes is refused (its puffs are not carried).
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim

ROW = cases.ROW_VOLATILE


def make(grid, built):
    params, field, parts, rows = built
    W, H = grid
    g, mats = N.derive_grid_params(params, W, H, rows)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return o, g, mats


def counts(o):
    return o.field("LEDGER", np.uint32, N.LEDGER_WORDS)[16:16 + N.MAX_MATERIALS]


def test_tables_caps_and_fume():
    p = cases.volatile_pool((64, 48))[0]
    g, m = N.derive_grid_params(p, 64, 48, [ROW])
    assert g.caps & N.CAP_VOLATILE and m[ROW].rx[3] == 2.0
    g, _ = N.derive_grid_params(p, 64, 48, [N.ROW_WATER])
    assert not g.caps & N.CAP_VOLATILE
    assert not g.caps & N.CAP_PUFFS
    # a liquid that fumes leaves puffs of mist: the tank may hold puffs
    p.mat[ROW].fume = 0.5
    g, m = N.derive_grid_params(p, 64, 48, [ROW])
    assert g.caps & N.CAP_PUFFS and m[ROW].fume == 0.5


@pytest.mark.parametrize("vol", [3.0, 12.0])
def test_surface_particles_leave_at_the_rate(vol):
    grid = (96, 48)
    W, H = grid
    o, g, m = make(grid, cases.volatile_pool(grid, volatility=vol))
    o.begin_batch(2 / 240, 2)
    o.substep(0)
    o.substep(1)  # a grid and a gathered substep: the pool has its levels
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    cell = o.field("CELL", np.float32, (W * H, 4))
    c0 = counts(o)[ROW]
    # which particles can leave: a neighbour of their (clamped) cell is open air
    live = before["kind"] != 0
    qx = np.clip(before["pos"][:, 0].astype(int), 1, W - 2)
    qy = np.clip(before["pos"][:, 1].astype(int), 1, H - 2)
    qi = qy * W + qx
    level = cell[:, 3]
    field = cases.volatile_pool(grid)[1]
    sol = np.asarray(field.solid, np.float32).reshape(-1, 4)[:, 3]
    near = np.zeros(len(before), bool)
    for d in (-1, 1, -W, W):
        n = qi + d
        near |= (level[n] <= 0.0) & (sol[n] > 0.5)
    cand = live & near
    assert cand.sum() > 150 and (live & ~near).sum() > 500
    o.begin_batch(1 / 240, 1)
    o.substep(1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    gone = live & (after["kind"] == 0)
    # nothing below the surface left
    assert not np.any(gone & ~cand)
    n, k = int(cand.sum()), int(gone.sum())
    q = vol / 240.0
    assert abs(k - n * q) <= 4.0 * np.sqrt(n * q * (1 - q)) + 1, (k, n * q)
    assert counts(o)[ROW] - c0 == k
    # a reset clears the counts
    o.reset(before[live])
    assert counts(o)[ROW] == 0


def test_pool_evaporates_over_time():
    grid = (64, 48)
    o, g, m = make(grid, cases.volatile_pool(grid, volatility=6.0))
    n0 = o.n_slots()
    steps = 48
    o.begin_batch(steps / 240, steps)
    for k in range(steps):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    left = int((p["kind"] == 0).sum())
    assert left > 20 and left == counts(o)[ROW]
    assert left < 0.2 * n0  # only the surface evaporates
