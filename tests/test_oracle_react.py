"""Pins of the oracle's reaction restatement (particles.comp:740-770, the reaction mask of
grid_prep.comp:110-118, SimParams::closure's products, FluidSim::caps's CAP_REACT), on the
CPU:

  * the tank takes in the product of a reaction whose partner is present, and CAP_REACT is on
    only while both partners are in it;
  * a cell's mask has bit j where the tank's j-th row makes up more than REACT_MIN of it;
  * where both partners are present each particle reacts with probability rate x dt: the
    share that turns into the product in one substep is that probability within the
    binomial spread, and a partner with no rate does not react;
  * a liquid with no product is used up, and the ledger counts it.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim


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


def test_closure_and_caps():
    p = cases.reacting_mix((64, 48))[0]
    g, m = N.derive_grid_params(p, 64, 48, [cases.ROW_ACID, cases.ROW_LYE])
    rows = [g.present[k] for k in range(g.n_present)]
    assert rows == [cases.ROW_ACID, cases.ROW_LYE, cases.ROW_BRINE]  # the product joins
    assert g.caps & N.CAP_REACT
    assert m[cases.ROW_ACID].ic[0] == cases.ROW_LYE and m[cases.ROW_ACID].ic[1] == cases.ROW_BRINE
    assert m[cases.ROW_ACID].ic[2] == rows.index(cases.ROW_LYE)
    assert m[cases.ROW_LYE].ic[0] == -1  # no rate: no reaction
    g, m = N.derive_grid_params(p, 64, 48, [cases.ROW_ACID])
    assert not g.caps & N.CAP_REACT and m[cases.ROW_ACID].ic[0] == -1
    assert [g.present[k] for k in range(g.n_present)] == [cases.ROW_ACID]


def test_reaction_mask_and_rate():
    grid = (64, 48)
    W, H = grid
    rate = 12.0
    o, g, m = make(grid, cases.reacting_mix(grid, rate=rate))
    o.begin_batch(2 / 240, 2)
    o.substep(0)
    o.substep(1)
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    acc = o.field("ACC", np.int32, (W * H, g.acc_stride)).astype(np.float64)
    mask = o.field("RMASK", np.uint32, W * H)
    # the mask from the accumulator's weights: row j's share of the cell above REACT_MIN
    cw = acc[:, 4]
    rows = [g.present[k] for k in range(g.n_present)]
    for j, r in enumerate(rows):
        ch = m[r].chan_w
        w = acc[:, ch] if ch >= 0 else cw - sum(acc[:, m[q].chan_w] for q in rows if m[q].chan_w >= 0)
        share = np.where(cw > 0, np.minimum(w / np.maximum(cw, 1), 1.0), 0.0)
        bit = (mask >> j) & 1
        clear = np.abs(share - 0.02) > 1e-4
        assert np.array_equal(bit[clear] == 1, share[clear] > 0.02), r
    o.begin_batch(1 / 240, 1)
    o.substep(1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    acid = (before["kind"] & 255) == cases.ROW_ACID + 1
    lye = (before["kind"] & 255) == cases.ROW_LYE + 1
    qx = np.clip(before["pos"][:, 0].astype(int), 0, W - 1)
    qy = np.clip(before["pos"][:, 1].astype(int), 0, H - 1)
    j_lye = rows.index(cases.ROW_LYE)
    with_partner = acid & (((mask[qy * W + qx] >> j_lye) & 1) == 1)
    brine = (after["kind"] & 255) == cases.ROW_BRINE + 1
    turned = acid & brine
    assert not np.any(turned & ~with_partner)
    assert not np.any(lye & brine)  # lye has no rate of its own here
    n, k = int(with_partner.sum()), int((turned & with_partner).sum())
    q = rate / 240.0
    assert n > 500
    assert abs(k - n * q) <= 4.0 * np.sqrt(n * q * (1 - q)) + 1, (k, n * q)


def test_used_up_reactant_is_counted():
    grid = (64, 48)
    o, g, m = make(grid, cases.reacting_mix(grid, rate=20.0, into=-1))
    steps = 12
    o.begin_batch(steps / 240, steps)
    for k in range(steps):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    gone = int((p["kind"] == 0).sum())
    assert gone > 50
    counts = o.field("LEDGER", np.uint32, N.LEDGER_WORDS)[16:16 + N.MAX_MATERIALS]
    assert counts[cases.ROW_ACID] == gone and counts.sum() == gone


# ---- heat made inside (particles.comp:733-770), counted in ledger entry 7 -----------------

WATER_HEAT = 998.0 * 4182.0  # J / (m^3 K): the ledger's unit is one K of a water-filled cell


def one_pass(params, field, parts, rows):
    """A thermal tank's particles before and after one particle pass whose gather sees zero
    grid velocities and no conduction (no heat pass ran), with the ledger after it."""
    W, H = 48, 32
    g, mats = N.derive_grid_params(params, W, H, rows, thermal=True)
    g.n_slots = len(parts)
    o = OracleSim((W, H), W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    o.begin_batch(1 / 240, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    vel = o.field("VEL", np.float32, (W * H, 4)).copy()
    vel[:, :2] = 0.0
    o.put_field("VEL", vel)
    o.put_field("DHEAT", np.zeros((W * H, 4), np.float32))
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    dt = float(o.field("STEP", N.STEP_DTYPE, 1)[0]["dt"])
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    L = o.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16].astype(np.uint64)
    made = float(int(L[14]) | int(L[15]) << 32) / 65536.0
    return before, after, dt, made * WATER_HEAT * params.cell_size ** 2, g


def t_tol(t):
    """Half an ulp of the new temperature plus a relative 1e-6."""
    return 0.6 * np.spacing(np.abs(t).astype(np.float32)).astype(np.float64) + 1e-6 * np.abs(t)


@pytest.mark.parametrize("power", [2.0e4, 6.0e4])
def test_self_heating_warms_at_its_power_and_is_counted(power):
    """A pool whose liquid makes `power` W/kg of its own: in one substep every particle warms by
    power dt / c (K), and the ledger's heat made inside is sum m_p power dt (J per metre of
    depth), m_p = rho dx^2 / PPC, to half a ledger quantum per particle."""
    grid = (48, 32)
    params, field, parts, rows = cases.thermal_pool(grid, temp=lambda x, y: np.full(len(x), 30.0))
    m = params.mat[N.ROW_WATER]
    m.self_heat = power
    before, after, dt, made, g = one_pass(params, field, parts, rows)
    assert np.array_equal(before["id"], after["id"])
    dT = power * dt * params.heat_rate / m.heat_capacity
    t0 = before["t"].astype(np.float64)
    t1 = after["t"].astype(np.float64)
    assert np.all(np.abs(t1 - (t0 + dT)) <= t_tol(t1) + 1e-6 * dT)
    mass = m.density * params.cell_size ** 2 / N.PPC
    want = len(before) * mass * power * dt * params.heat_rate
    quantum = WATER_HEAT * params.cell_size ** 2 / 65536.0
    assert abs(made - want) <= 0.5 * quantum * len(before) + 1e-6 * want
    assert made == pytest.approx(want, rel=2e-3)


def test_heat_of_reaction_warms_what_reacts_and_is_counted():
    """Acid that turns into brine where lye is present gives off react_heat J/kg: each particle
    that reacts warms by react_heat / c of the acid (K), the others not at all, and the ledger
    counts sum m_p react_heat over the reacted ones."""
    grid = (48, 32)
    heat = 3.0e5
    params, field, parts, rows = cases.reacting_mix(grid, rate=60.0)
    a = params.mat[cases.ROW_ACID]
    a.react_heat = heat
    parts = parts.copy()
    parts["t"] = np.float32(20.5)  # off the ambient: a thermal tank
    before, after, dt, made, g = one_pass(params, field, parts, rows)
    assert g.caps & N.CAP_REACT
    was = (before["kind"] & 255) - 1
    now = (after["kind"] & 255) - 1
    reacted = (was == cases.ROW_ACID) & (now == cases.ROW_BRINE)
    assert reacted.sum() > 20
    t0 = before["t"].astype(np.float64)
    t1 = after["t"].astype(np.float64)
    dT = heat / a.heat_capacity
    assert np.all(np.abs(t1[reacted] - (t0[reacted] + dT)) <= t_tol(t1[reacted]) + 1e-6 * dT)
    assert np.array_equal(after["t"][~reacted], before["t"][~reacted])
    want = reacted.sum() * a.density * params.cell_size ** 2 / N.PPC * heat
    quantum = WATER_HEAT * params.cell_size ** 2 / 65536.0
    assert abs(made - want) <= 0.5 * quantum * reacted.sum() + 1e-6 * want
