"""Pins of the oracle's heat restatement (heat.comp, the thermal branches of grid_prep.comp
and particles.comp), on the CPU.

  * a pool at one temperature reads back that temperature in every liquid cell (the heat
    sums' fixed-point scales and heat capacities);
  * the viscosity and density of a cell follow the material's temperature law and thermal
    expansion at the cell's temperature, against their closed forms (Vogel for water);
  * conduction in an insulated pool conserves the particles' heat to rounding and lowers
    the spread of temperatures (a sign error runs it backwards), and the slowest cosine
    mode decays at the rate the heat equation gives for water's diffusivity;
  * the exchange inside a cell conserves heat and pulls particles together;
  * with air and a hot floor, what the particles gained is what the ledger booked from
    solids minus what it booked to air, and air cooling runs at the film coefficient.
"""
import math

import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim

WATER_HEAT = 998.0 * 4182.0


def sim(grid, params, field, parts, rows):
    W, H = grid
    g, mats = N.derive_grid_params(params, W, H, rows, thermal=True)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return o, g, mats


def parts_of(o):
    return o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())


def heat(o, mats):
    """The particles' heat content, K x water-filled cell volumes: sum C (t + 273.15) / PPC."""
    p = parts_of(o)
    rows = (p["kind"] & 255).astype(np.int64) - 1
    C = np.array([mats[r].th[0] for r in rows], np.float64)
    return float(np.sum(C * (p["t"].astype(np.float64) + 273.15)) / N.PPC)


def ledger(o):
    L = o.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16].astype(np.uint64)
    return [float((int(L[2 * e]) | int(L[2 * e + 1]) << 32) / 65536.0) for e in range(8)]


def run(o, n, first=True):
    o.begin_batch(n / 240.0, n)
    for k in range(n):
        o.substep(0 if (first and k == 0) else 1)


def gather_only(o):
    """The start of one more substep: the particles take the last heat pass's result."""
    o.begin_batch(1 / 240.0, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)


def liquid_cells(o, W, H):
    th = o.field("THERM", np.float32, (W * H, 4))
    cell = o.field("CELL", np.float32, (W * H, 4))
    return th, cell, (cell[:, 0] > 0.5) & (th[:, 1] > 0)


def test_uniform_pool_reads_its_temperature():
    grid = (40, 24)
    for t in (5.0, 20.0, 63.5, 95.0):
        params, field, parts, rows = cases.thermal_pool(grid, temp=lambda x, y: np.full(len(x), t))
        o, _, _ = sim(grid, params, field, parts, rows)
        o.begin_batch(1 / 240.0, 1)
        for s in ("STEP", "CLEAR_ACC"):
            o.stage(s)
        o.stage("PARTICLES", 0)
        o.stage("GRID_PREP")
        th, _, live = liquid_cells(o, *grid)
        assert live.sum() > 100
        # the sums are quantised at T_SCALE / MASS_SCALE = 1/256 K per unit of weight, rounded
        # per node: about 1e-2 K at most in a cell, unbiased on average
        assert np.max(np.abs(th[live, 0] - t)) < 1.2e-2, t
        assert abs(float(np.mean(th[live, 0])) - t) < 1e-3, t


def test_temperature_law_and_expansion_at_the_cell_temperature():
    grid = (40, 24)
    t = 60.0
    params, field, parts, rows = cases.thermal_pool(grid, temp=lambda x, y: np.full(len(x), t))
    o, g, mats = sim(grid, params, field, parts, rows)
    o.begin_batch(1 / 240.0, 1)
    for s in ("STEP", "CLEAR_ACC"):
        o.stage(s)
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    th, cell, live = liquid_cells(o, *grid)
    m = params.mat[N.ROW_WATER]
    # Vogel: ln nu(T) - ln nu(Tref) = B / (T - C) - B / (Tref - C), kelvin
    vft = m.temp_b / (t + 273.15 - m.temp_c) - m.temp_b / (m.temp_ref + 273.15 - m.temp_c)
    nu_cells = m.viscosity / params.cell_size ** 2 * math.exp(vft)
    nu = cell[live, 2] / cell[live, 1]
    assert np.allclose(nu, nu_cells, rtol=2e-5)
    # the cell's density (relative to water at rest) follows 1 - beta (T - Tref)
    rho = (1.0 - m.expansion * (t - m.temp_ref)) * m.density / 998.0
    assert np.allclose(cell[live, 1], rho, rtol=2e-5)


def insulated(grid, temp, **kw):
    kw.setdefault("air_cooling", 0.0)
    kw.setdefault("evaporation", 0)
    kw.setdefault("emissivity", 0.0)
    kw.setdefault("expansion", 0.0)
    kw.setdefault("eddy_heat", 0.0)
    return cases.thermal_pool(grid, temp=temp, **kw)


def test_conduction_conserves_heat_and_lowers_the_spread():
    grid = (40, 24)
    # mercury: thirty times water's diffusivity, so a dozen substeps show
    params, field, parts, rows = insulated(grid, lambda x, y: 20.0 + 40.0 * (x > 20.0), heat_rate=50.0,
                                           row=N.ROW_MERCURY)
    o, _, mats = sim(grid, params, field, parts, rows)
    h0 = heat(o, mats)
    s0 = float(np.std(parts_of(o)["t"]))
    run(o, 12)
    gather_only(o)
    h1 = heat(o, mats)
    s1 = float(np.std(parts_of(o)["t"]))
    assert abs(h1 - h0) <= 2e-6 * h0, (h0, h1)
    assert s1 < 0.99 * s0, (s0, s1)
    assert ledger(o)[:4] == [0.0, 0.0, 0.0, 0.0]


def test_slowest_mode_decays_at_the_heat_equation_rate():
    """T = 40 + 10 cos(pi (x - x0) / L) across the open width of a mercury pool: the heat
    equation decays it as exp(-a pi^2 t / L^2), a = k / (rho c) (in cells^2 / s) times the
    heat rate."""
    grid = (24, 20)
    wall = 3.0
    x0, L = wall, grid[0] - 2 * wall
    rate = 50.0
    params, field, parts, rows = insulated(grid, lambda x, y: 40.0 + 10.0 * np.cos(np.pi * (x - x0) / L), heat_rate=rate,
                                           row=N.ROW_MERCURY)
    o, _, _ = sim(grid, params, field, parts, rows)

    def amplitude():
        p = parts_of(o)
        c = np.cos(np.pi * (p["pos"][:, 0].astype(np.float64) - x0) / L)
        t = p["t"].astype(np.float64)
        return float(np.sum((t - t.mean()) * c) / np.sum(c * c))

    a0 = amplitude()
    steps = 240
    run(o, steps)
    gather_only(o)
    a1 = amplitude()
    m = params.mat[N.ROW_MERCURY]
    kappa = m.conductivity / (m.density * m.heat_capacity) / params.cell_size ** 2 * rate  # cells^2 / s
    want = math.exp(-kappa * math.pi ** 2 * (steps / 240.0) / L ** 2)
    got = a1 / a0
    assert abs(math.log(got) / math.log(want) - 1.0) < 0.3, (got, want)


def test_exchange_inside_cells_conserves_heat_and_pulls_together():
    """Particles of alternating temperature within each cell, cell means equal everywhere:
    conduction between cells has nothing to do; the exchange inside the cells does."""
    grid = (32, 20)
    params, field, parts, rows = insulated(grid, None, heat_rate=20.0)
    order = np.arange(len(parts))
    parts["t"] = np.where(order % 2 == 0, 30.0, 50.0).astype(np.float32)
    o, _, mats = sim(grid, params, field, parts, rows)
    h0 = heat(o, mats)
    s0 = float(np.std(parts_of(o)["t"]))
    run(o, 6)
    gather_only(o)
    assert abs(heat(o, mats) - h0) <= 2e-6 * h0
    assert float(np.std(parts_of(o)["t"])) < 0.9 * s0


def test_ledger_balances_the_particles_with_air_and_a_hot_floor():
    grid = (40, 24)
    params, field, parts, rows = cases.thermal_pool(grid, temp=lambda x, y: np.full(len(x), 40.0),
                                                    hot=(10.0, 30.0, 76.85), heat_rate=20.0)
    assert field.temp is not None and (field.temp > -1e8).sum() > 10
    o, _, mats = sim(grid, params, field, parts, rows)
    h0 = heat(o, mats)
    run(o, 8)
    gather_only(o)
    h1 = heat(o, mats)
    L = ledger(o)
    booked = L[0] - L[1] + L[2] - L[3]
    assert L[1] > 0.0 and L[2] > 0.0     # lost to air, gained from the floor
    assert abs((h1 - h0) - booked) <= 1e-3 * max(abs(booked), 1.0), (h1 - h0, booked)


def test_air_cooling_runs_at_the_film_coefficient():
    """A pool above the air's temperature, film only (no radiation, no evaporation): the
    first pass's loss to air is linear in the film coefficient and in the temperature
    difference, and at least what a flat surface of the tank's width loses; the exposed
    parts of cells that hold more liquid than their neighbours add to it (heat.comp:
    150-162 counts them), so at most twice that."""
    grid = (40, 24)
    rate = 20.0

    def loss(film, t):
        params, field, parts, rows = cases.thermal_pool(grid, temp=lambda x, y: np.full(len(x), t), heat_rate=rate,
                                                        air_cooling=film, evaporation=0, emissivity=0.0, eddy_heat=0.0)
        o, _, _ = sim(grid, params, field, parts, rows)
        run(o, 1)
        return ledger(o)[1], params

    l10, params = loss(10.0, 50.0)
    l20, _ = loss(20.0, 50.0)
    l10_hot, _ = loss(10.0, 80.0)
    assert abs(l20 / l10 - 2.0) < 1e-2          # cells whose loss hits the clamp at their own excess fall short
    assert abs(l10_hot / l10 - 2.0) < 2e-2     # 60 K over 30 K
    width = grid[0] - 6                         # open cells across the tank
    film = params.air_cooling / (WATER_HEAT * params.cell_size)
    flat = film * 30.0 * (1 / 240.0) * rate * width
    assert flat < l10 < 2.0 * flat, (l10, flat)
