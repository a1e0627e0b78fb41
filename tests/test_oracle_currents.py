"""Pins of the oracle's currents and drains (particles.comp:578-592, 809-829; Current,
params.rs:1113-1136), on the CPU, one particle pass over an exact lattice of water with the
grid velocities set after grid_prep:

* a current drives the liquid in its zone along its direction: the velocity along it relaxes
  towards the current's speed as 1 - exp(-rate (1 - r^2)^2 dt) in one pass (r the distance
  from the centre over the radius), the motion across it is left alone, liquid already faster
  along it is not braked, and nothing outside the zone changes;
* a drain removes each particle in its zone with probability -rate (1 - 3/4 r^2) dt: the count
  removed in one pass is that sum within four binomial standard deviations, and none outside.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd import scene as sc
from oracle.binding import OracleSim
from test_oracle_drift import lattice_particles, W, H

CENTRE, RADIUS = (24.0, 20.0), 7.0


def current_pass(currents, grid_vel):
    params = N.default_params()
    params.fixed_hz = 240.0
    params.cell_size = 0.01
    params.crowd_jitter = 0.0
    rows = [N.ROW_WATER]
    parts = cases.settle(lattice_particles([N.ROW_WATER] * 4), params)
    g, mats = N.derive_grid_params(params, W, H, rows, currents=currents)
    assert g.ext[3] == len(currents) and bool(g.caps & N.CAP_CURRENTS) == (len(currents) > 0)
    o = OracleSim((W, H), W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(sc.Scene((W, H)).rasterize())
    o.reset(parts)
    o.begin_batch(1 / 240, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    vel = o.field("VEL", np.float32, (H * W, 4)).copy()
    vel[:, 0] = grid_vel[0]
    vel[:, 1] = grid_vel[1]
    o.put_field("VEL", vel)
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    dt = float(o.field("STEP", N.STEP_DTYPE, 1)[0]["dt"])
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    return before, after, dt, params


def deep(p):
    """Particles whose 3 x 3 gather lies inside the block (the gathered velocity is the grid's)."""
    x, y = p["pos"][:, 0], p["pos"][:, 1]
    return (x >= 11) & (x < 37) & (y >= 11) & (y < 29)


@pytest.mark.parametrize("grid_vel", [(0.0, 0.0), (30.0, 0.0), (0.0, -200.0), (0.0, 60.0)])
def test_current_relaxes_the_velocity_along_it(grid_vel):
    cur = N.Current.new(CENTRE, RADIUS, 0.0, 1.0)  # straight down at CURRENT_SPEED
    speed = N.CURRENT_SPEED
    assert cur.vel[0] == 0.0 and cur.vel[1] == pytest.approx(-speed) and cur.rate == N.CURRENT_RATE
    before, after, dt, params = current_pass([cur], grid_vel)
    sel = deep(before)
    r2 = ((before["pos"][:, 0].astype(np.float64) - CENTRE[0]) ** 2 +
          (before["pos"][:, 1].astype(np.float64) - CENTRE[1]) ** 2) / RADIUS ** 2
    inside = sel & (r2 < 1.0)
    outside = sel & (r2 >= 1.0)
    assert inside.sum() > 400 and outside.sum() > 400
    u0, v0 = grid_vel
    down = -v0  # the velocity along the current's direction (straight down)
    relax = 1.0 - np.exp(-N.CURRENT_RATE * (1.0 - r2[inside]) ** 2 * dt)
    want_down = down + max(speed - down, 0.0) * relax
    got = after["vel"][inside].astype(np.float64)
    assert got[:, 0] == pytest.approx(np.full(inside.sum(), u0), abs=1e-4)  # across it: unchanged
    assert -got[:, 1] == pytest.approx(want_down, rel=1e-5, abs=1e-4)
    if down >= speed:
        assert np.allclose(-got[:, 1], down, atol=1e-4)  # never braked
    out = after["vel"][outside].astype(np.float64)
    assert np.allclose(out[:, 0], u0, atol=1e-4) and np.allclose(out[:, 1], v0, atol=1e-4)


def test_no_current_no_change():
    before, after, dt, _ = current_pass([], (0.0, 0.0))
    sel = deep(before)
    assert np.allclose(after["vel"][sel], 0.0, atol=1e-5)


@pytest.mark.parametrize("strength", [1.0, 2.0])
def test_drain_removes_at_its_rate(strength):
    radius = 10.0
    dr = N.Current.drain(CENTRE, radius, strength)
    assert dr.rate == pytest.approx(-N.DRAIN_RATE * strength)
    before, after, dt, _ = current_pass([dr], (0.0, 0.0))
    r2 = ((before["pos"][:, 0].astype(np.float64) - CENTRE[0]) ** 2 +
          (before["pos"][:, 1].astype(np.float64) - CENTRE[1]) ** 2) / radius ** 2
    gone = after["kind"] == N.KIND_DEAD
    assert not np.any(gone[r2 >= 1.0])
    # in the whole zone, and ring by ring (the fall-off to a quarter at the rim)
    for lo, hi in ((0.0, 1.0), (0.0, 0.3), (0.6, 1.0)):
        ring = (r2 >= lo) & (r2 < hi)
        pr = N.DRAIN_RATE * strength * (1.0 - 0.75 * r2[ring]) * dt
        mean, sd = pr.sum(), np.sqrt((pr * (1.0 - pr)).sum())
        assert ring.sum() > 150
        assert abs(gone[ring].sum() - mean) <= 4.0 * sd, (lo, hi, gone[ring].sum(), mean, sd)
