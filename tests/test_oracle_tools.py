"""Pins of the oracle's mouse tools (particles.comp:594-631, air.glsl:24-37, 173-179), on
the CPU, by closed forms and by comparison with the same pool under the idle tool:

  * Hold (strength 600) and the stream nozzle (1000) saturate their relaxation where
    strength (1 - r^2) dt >= 1: the particles there are at rest, or at the launch velocity,
    exactly;
  * Grab relaxes each particle's velocity towards tool_vel + 4 d by strength (1 - r^2) dt,
    Push / pull towards the radial speed, Whirl towards the turn about the cursor, and
    nothing outside the radius changes;
  * Erase removes exactly the particles within the radius;
  * Heat adds heat_power (1 - r^2) dt kelvin for water (in proportion to the inverse heat
    capacity for another liquid), booked in ledger entry 4; Cool takes it away, not below
    -150 deg C, booked in entry 5.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim
from test_oracle_heat import insulated, ledger

GRID = (48, 32)
DT = 1.0 / 240.0


def pool(thermal=False, row=N.ROW_WATER, temp=None):
    if thermal:
        params, field, parts, rows = insulated(GRID, temp, row=row)
    else:
        params, field, parts, rows = cases.dam_break(GRID, frac=(1.0, 0.6))
        params.crowd_jitter = 0.0
    return params, field, parts, rows


def run(params, field, parts, rows, mode=0, pos=(24.0, 10.0), vel=(0.0, 0.0), radius=6.0, strength=40.0, thermal=False):
    """One substep under the idle tool, then the particle pass of the next with `mode`."""
    W, H = GRID
    g, mats = N.derive_grid_params(params, W, H, rows, thermal=thermal)
    g.n_slots = len(parts)
    o = OracleSim(GRID, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    o.begin_batch(DT, 1)
    o.substep(0)
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    g.tool_mode = mode
    g.tool_pos[0], g.tool_pos[1] = pos
    g.tool_vel[0], g.tool_vel[1] = vel
    g.tool_radius = radius
    g.tool_strength = strength
    o.set_gparams(g, mats)
    e0 = ledger(o)
    o.begin_batch(DT, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    return before, after, [a - b for a, b in zip(ledger(o), e0)], g, mats


def inside(before, pos, radius):
    d = np.float32(pos) - before["pos"]
    r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) / np.float32(radius * radius)
    return r2 < 1.0, r2, d


def test_hold_and_nozzle_saturate_where_strength_times_dt_reaches_one():
    built = pool()
    for mode, strength, vel in ((8, 600.0, (0.0, 0.0)), (4, 1000.0, (30.0, -12.0))):
        b, a, _, _, _ = run(*built, mode=mode, strength=strength, vel=vel)
        ins, r2, _ = inside(b, (24.0, 10.0), 6.0)
        full = ins & (strength * (1.0 - r2) * DT >= 1.0)
        assert full.sum() > 40, mode
        # the relaxation saturates: the gathered velocity is replaced, no midpoint is taken
        assert np.all(a["vel"][full] == np.float32(vel)), mode


def test_grab_push_whirl_relax_towards_their_targets_inside_the_radius_only():
    built = pool()
    b0, a0, _, _, _ = run(*built, mode=0)
    for mode, vel in ((1, (40.0, 10.0)), (3, (25.0, 0.0)), (3, (-25.0, 0.0)), (7, (30.0, 0.0))):
        b, a, _, _, _ = run(*built, mode=mode, vel=vel)
        assert np.array_equal(b, b0)
        ins, r2, d = inside(b, (24.0, 10.0), 6.0)
        assert np.array_equal(a[~ins], a0[~ins]), mode  # nothing outside the radius changes
        k = np.clip(40.0 * (1.0 - r2.astype(np.float64)) * DT, 0.0, 1.0)
        v0 = a0["vel"][ins].astype(np.float64)
        v1 = a["vel"][ins].astype(np.float64)
        dd = d[ins].astype(np.float64)
        l = np.hypot(dd[:, 0], dd[:, 1])[:, None]
        if mode == 1:
            target = np.array(vel) + 4.0 * dd
        elif mode == 3:
            target = -dd / np.maximum(l, 0.15 * 6.0) * vel[0]
        else:
            target = np.stack([dd[:, 1], -dd[:, 0]], 1) / np.maximum(l, 0.4 * 6.0) * vel[0] + 2.0 * dd
        # the change is k (target - v) for the particles free of a crowd, at most that
        change = v1 - v0
        want = k[ins][:, None] * (target - v0)
        free = np.linalg.norm(change - want, axis=1) <= 1e-3 * (1.0 + np.linalg.norm(want, axis=1))
        assert free.mean() > 0.9, (mode, free.mean())
        # where it differs the crowd held it back: the same direction, less of it
        cos = np.sum(change * want, 1) / (np.linalg.norm(change, axis=1) * np.linalg.norm(want, axis=1) + 1e-12)
        assert np.all(cos[~free] > 0.999), mode
        assert np.all(np.linalg.norm(change[~free], axis=1) < np.linalg.norm(want[~free], axis=1)), mode


def test_erase_removes_exactly_the_particles_within_the_radius():
    built = pool()
    b, a, _, _, _ = run(*built, mode=2)
    ins, _, _ = inside(b, (24.0, 10.0), 6.0)
    assert ins.sum() > 50
    assert np.array_equal(a["kind"] == 0, ins)


@pytest.mark.parametrize("row", [N.ROW_WATER, N.ROW_OIL])
def test_heat_and_cool_add_power_times_dt_and_book_it(row):
    built = pool(thermal=True, row=row, temp=lambda x, y: np.full(len(x), 40.0))
    params = built[0]
    pos, radius = (24.0, 8.0), 6.0
    b0, a0, led0, _, _ = run(*built, mode=0, thermal=True)
    b, a, led, g, mats = run(*built, mode=5, pos=pos, radius=radius, thermal=True)
    ins, r2, _ = inside(b, pos, radius)
    C = float(mats[row].th[0])
    e = params.heat_power * (1.0 - r2.astype(np.float64)) * DT
    dT = a["t"].astype(np.float64) - a0["t"]
    assert np.allclose(dT[ins], e[ins] / C, rtol=1e-4, atol=1e-5)
    assert np.all(dT[~ins] == 0.0)
    assert led[4] - led0[4] == pytest.approx(np.sum(e[ins]) / 4.0, rel=1e-5)
    # Cool takes the same away, booked in entry 5
    b, a, led, _, _ = run(*built, mode=6, pos=pos, radius=radius, thermal=True)
    dT = a["t"].astype(np.float64) - a0["t"]
    assert np.allclose(dT[ins], -e[ins] / C, rtol=1e-4, atol=1e-5)
    assert led[5] - led0[5] == pytest.approx(np.sum(e[ins]) / 4.0, rel=1e-5)


def test_cool_stops_at_minus_150():
    # frozen oil at -140 deg C: its enthalpy temperature is that less its latent heat
    p0 = N.default_params()
    lat = p0.mat[N.ROW_OIL].latent / p0.mat[N.ROW_OIL].heat_capacity
    built = pool(thermal=True, row=N.ROW_OIL, temp=lambda x, y: np.full(len(x), -140.0 - lat))
    params = built[0]
    params.heat_power = 5000.0  # up to 20.8 K in a substep: more than the 10 K left
    b, a, _, _, mats = run(*built, mode=6, pos=(24.0, 8.0), radius=6.0, thermal=True)
    ins, _, _ = inside(b, (24.0, 8.0), 6.0)
    th = mats[N.ROW_OIL].th
    # the enthalpy temperature of -150 deg C in frozen oil: -150 less its latent heat as degrees
    floor = -150.0 - mats[N.ROW_OIL].th2[0]
    assert np.all(a["t"][ins] >= floor - 1e-3) and np.any(np.abs(a["t"][ins] - floor) < 1e-3)
    assert th[2] > -150.0  # (frozen there: the latent heat is below it)
