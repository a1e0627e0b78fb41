"""Pins of the oracle's boiling, residue, steam and puff branches (particles.comp:98-134,
303-305, 505-517, 633-731, 771-807; grid_prep.comp:163-167, 247; heat.comp:95-104;
air.glsl:81-175), on the CPU.

  * the tables: steam and the water it condenses into join a thermal tank whose aqueous
    liquid can boil while steam is on, the residue joins the rows, the steam row carries
    water's heat capacity, the steam parameters in grid units;
  * a liquid that makes no steam leaves exactly when it has taken up its latent heat of
    vaporisation (enthalpy temperature past boil + L / c), and the ledger books the heat of
    what left;
  * a boiling concentrate leaves its residue: the share of the particles that stay behind is
    the residue share, each a solid at the boiling point, and the ledger books the
    difference of the heat contents;
  * aqueous liquid turns into steam at the vapour quality steam_boil, remembering its row;
    steam at the free surface escapes as a puff with probability x, else it stays as its
    liquid at the boiling point, giving its latent heat to the air; steam below
    steam_condense condenses into the liquid it boiled from with its heat content kept;
  * steam drifts up through water at the terminal velocity of a bubble of the unresolved
    size (the drag balance of test_oracle_drift), and the share of steam reaches grid_prep
    and speeds up the exchange inside the cell to steam_mixing;
  * a puff takes up the air's rise exponentially with its time constant, dies at the end of
    its life, and the eddies of the air are divergence-free (the value noise's gradient is
    the derivative of the noise).
"""
import math

import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim
from test_oracle_drift import terminal_root
from test_oracle_heat import insulated, ledger

PPC = 4
WATER_LC = 2257.0e3 / 4182.0  # water's latent heat of vaporisation over its heat capacity, K


def sim(grid, params, field, parts, rows, thermal=True):
    W, H = grid
    g, mats = N.derive_grid_params(params, W, H, rows, thermal=thermal)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return o, g, mats


def parts_of(o):
    return o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())


def counts(o):
    L = o.field("LEDGER", np.uint32, N.LEDGER_WORDS)
    m = N.MAX_MATERIALS
    return L[16:16 + m], L[16 + m:16 + 2 * m], L[16 + 2 * m:16 + 3 * m]


def first_then_gather(o, grid):
    """One substep without a gather (nothing moves), then the particle pass of the next one:
    (particles and CELL field between the two)."""
    o.begin_batch(1 / 240.0, 1)
    o.substep(0)
    before = parts_of(o).copy()
    cell = o.field("CELL", np.float32, (grid[0] * grid[1], 4))
    o.begin_batch(1 / 240.0, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    return before, cell


def steam_kind(src_row):
    return (N.ROW_STEAM + 1) | ((src_row + 1) << 8)


def test_tables_steam_layout_caps_and_units():
    p = N.default_params()
    W, H = 64, 48
    g, m = N.derive_grid_params(p, W, H, [N.ROW_WATER])
    assert not g.caps & (N.CAP_STEAM | N.CAP_BOILS | N.CAP_PUFFS)
    g, m = N.derive_grid_params(p, W, H, [N.ROW_WATER], thermal=True)
    rows = list(g.present[:g.n_present])
    assert rows == [N.ROW_WATER, N.ROW_STEAM]
    assert g.caps & N.CAP_STEAM and g.caps & N.CAP_BOILS and g.caps & N.CAP_PUFFS
    # steam carries the mass and heat of the water it boiled from; its density is the vapour's
    assert m[N.ROW_STEAM].th[0] == m[N.ROW_WATER].th[0]
    assert m[N.ROW_STEAM].rho == pytest.approx(0.6 / 998.0, rel=1e-6)
    assert m[N.ROW_STEAM].group == 0 and m[N.ROW_STEAM].chan_w >= 0
    dx = p.cell_size
    assert list(g.steam) == pytest.approx([p.steam_condense, p.steam_rise / dx, p.steam_life, p.steam_boil], rel=1e-6)
    assert list(g.steam2) == pytest.approx([p.steam_mixing, 1.0, p.steam_swirl / dx, p.steam_eddy / dx], rel=1e-6)
    # steam off: boiled liquid leaves at once, no steam row
    p.steam = 0
    g, _ = N.derive_grid_params(p, W, H, [N.ROW_WATER], thermal=True)
    assert list(g.present[:g.n_present]) == [N.ROW_WATER]
    assert g.caps & N.CAP_BOILS and not g.caps & (N.CAP_STEAM | N.CAP_PUFFS)
    # an oily liquid makes no steam; brine brings its diluent and its residue
    p.steam = 1
    g, _ = N.derive_grid_params(p, W, H, [N.ROW_OIL], thermal=True)
    assert list(g.present[:g.n_present]) == [N.ROW_OIL]
    g, m = N.derive_grid_params(p, W, H, [N.ROW_BRINE], thermal=True)
    assert list(g.present[:g.n_present]) == [N.ROW_WATER, N.ROW_BRINE, N.ROW_SALT, N.ROW_STEAM]
    assert m[N.ROW_BRINE].residue == N.ROW_SALT and m[N.ROW_BRINE].residue_share == pytest.approx(0.144)


def test_liquid_that_makes_no_steam_leaves_at_its_latent_heat():
    grid = (40, 24)
    row = N.ROW_OIL
    params, field, parts, rows = insulated(grid, None, row=row)
    m = params.mat[row]
    m.boil, m.latent_vap = 150.0, 300.0e3
    lc = m.latent_vap / m.heat_capacity
    hi = np.arange(len(parts)) % 2 == 1
    parts["t"] = np.where(hi, m.boil + 1.05 * lc, m.boil + 0.95 * lc).astype(np.float32)
    o, g, mats = sim(grid, params, field, parts, rows)
    assert g.caps & N.CAP_BOILS and not g.caps & N.CAP_STEAM
    e0 = ledger(o)[6]
    first_then_gather(o, grid)
    after = parts_of(o)
    gone = after["kind"] == N.KIND_DEAD if hasattr(N, "KIND_DEAD") else after["kind"] == 0
    assert np.array_equal(gone, hi)  # exactly those past boil + L / c
    assert np.allclose(after["t"][~hi], parts["t"][~hi], atol=2e-2)
    boiled, residue, steam = counts(o)
    assert boiled[row] == hi.sum() and residue.sum() == 0 and steam.sum() == 0
    C = float(mats[row].th[0])
    heat = float(np.sum(C * (parts["t"][hi].astype(np.float64) + 273.15))) / PPC
    assert ledger(o)[6] - e0 == pytest.approx(heat, rel=1e-5)


def test_boiling_brine_leaves_its_residue():
    grid = (48, 32)
    row = N.ROW_BRINE
    params, field, parts, rows = insulated(grid, None, row=row)
    params.steam = 0  # the brine leaves at once: no steam
    mb = params.mat[row]
    lc = mb.latent_vap / mb.heat_capacity
    parts["s"] = 1.0
    parts["t"] = np.float32(mb.boil + 1.05 * lc)
    o, g, mats = sim(grid, params, field, parts, rows)
    e0 = ledger(o)[6]
    first_then_gather(o, grid)
    after = parts_of(o)
    salt = (after["kind"] & 255) == N.ROW_SALT + 1
    dead = after["kind"] == 0
    assert np.all(salt | dead)  # every brine particle left; some left salt behind
    n = len(parts)
    share = salt.sum() / n
    assert abs(share - mb.residue_share) < 4.0 * math.sqrt(mb.residue_share * (1 - mb.residue_share) / n), share
    # the residue is a solid at the boiling point
    th = mats[N.ROW_SALT].th
    lat = mats[N.ROW_SALT].th2[0]
    t_salt = after["t"][salt].astype(np.float64)
    lo = th[2] - lat  # phase_of: below the solidus less the latent heat, te = T - L / c
    assert np.all(t_salt <= lo) and np.allclose(t_salt + lat, mb.boil, atol=1e-3)
    assert np.all(after["s"][salt] == 1.0)
    boiled, residue, steam = counts(o)
    assert boiled[row] == n and residue[N.ROW_SALT] == salt.sum() and steam.sum() == 0
    # what left: the brine's heat content less the residue's (from absolute zero)
    Cb, Cs = float(mats[row].th[0]), float(mats[N.ROW_SALT].th[0])
    t0 = parts["t"].astype(np.float64)
    heat = (np.sum(Cb * (t0 + 273.15)) - np.sum(Cs * (t_salt + 273.15))) / PPC
    assert ledger(o)[6] - e0 == pytest.approx(heat, rel=1e-5)


def test_water_turns_into_steam_at_steam_boil():
    grid = (40, 24)
    params, field, parts, rows = insulated(grid, None, row=N.ROW_WATER)
    hi = np.arange(len(parts)) % 3 == 0
    x = np.where(hi, params.steam_boil + 0.002, params.steam_boil - 0.002)
    parts["t"] = (100.0 + x * WATER_LC).astype(np.float32)
    o, g, mats = sim(grid, params, field, parts, rows)
    assert g.caps & N.CAP_STEAM
    first_then_gather(o, grid)
    after = parts_of(o)
    k = after["kind"]
    assert np.all((k[hi] & 255) == N.ROW_STEAM + 1) and np.all(((k[hi] >> 8) & 255) == N.ROW_WATER + 1)
    assert np.all(k[~hi] == N.ROW_WATER + 1)
    # the same heat content: steam carries water's heat capacity
    assert np.allclose(after["t"], parts["t"], atol=2e-2)
    boiled, residue, steam = counts(o)
    assert steam[N.ROW_WATER] == hi.sum() and boiled.sum() == 0


def steam_pool(grid, x, src=N.ROW_WATER, depth=0.5, s=1.0, **kw):
    params, field, parts, _ = insulated(grid, None, frac=(1.0, depth), row=N.ROW_WATER, **kw)
    parts["kind"] = steam_kind(src)
    parts["s"] = s
    parts["t"] = np.float32(100.0 + x * WATER_LC)
    rows = sorted({N.ROW_WATER, N.ROW_STEAM, src})
    return params, field, parts, rows


def surface_of(before, cell, grid):
    W, H = grid
    px = np.clip(before["pos"][:, 0].astype(np.int64), 0, W - 1)
    py = np.clip(before["pos"][:, 1].astype(np.int64), 0, H - 1)
    up = np.clip(py + 1, 0, H - 1)
    return (cell[py * W + px, 3] <= 0.0) | (cell[up * W + px, 3] <= 0.0)


def test_steam_escapes_at_the_surface_with_probability_x():
    grid = (96, 24)
    x = 0.3
    params, field, parts, rows = steam_pool(grid, x)
    o, g, mats = sim(grid, params, field, parts, rows)
    e0 = ledger(o)[6]
    before, cell = first_then_gather(o, grid)
    after = parts_of(o)
    surf = surface_of(before, cell, grid)
    puff = after["kind"] == N.KIND_PUFF
    water = after["kind"] == N.ROW_WATER + 1
    assert surf.sum() > 300
    assert np.all(~puff[~surf] & ~water[~surf])  # below the surface it stays steam
    assert np.all((after["kind"][~surf] & 255) == N.ROW_STEAM + 1)
    assert np.all(puff[surf] | water[surf])
    share = puff[surf].sum() / surf.sum()
    assert abs(share - x) < 4.0 * math.sqrt(x * (1 - x) / surf.sum()), share
    # puffs: age 0, the vapour of steam; those that stay are water at the boiling point
    assert np.all(after["s"][puff] == 0.0) and np.all(after["c"][puff][:, 0] == 1.0)
    assert np.all(after["t"][water] == np.float32(100.0))
    boiled, _, _ = counts(o)
    assert boiled[N.ROW_STEAM] == puff.sum()
    C = float(mats[N.ROW_STEAM].th[0])
    t0 = parts["t"].astype(np.float64)
    heat = (np.sum(C * (t0[puff] + 273.15)) + np.sum(C * (t0[water] - 100.0))) / PPC
    assert ledger(o)[6] - e0 == pytest.approx(heat, rel=1e-5)
    # the share of steam reaches the grid: cells inside the steam pool are all steam
    gb = o.field("GBASE", np.float32, (grid[0] * grid[1], 4))
    inside = cell[:, 0] > 0.99
    assert inside.sum() > 100 and np.allclose(gb[inside, 3], 1.0, atol=1e-6)


def test_steam_below_steam_condense_condenses_into_its_liquid():
    grid = (48, 24)
    x = 0.003
    params, field, parts, rows = steam_pool(grid, x, src=N.ROW_BRINE)
    o, g, mats = sim(grid, params, field, parts, rows)
    before, cell = first_then_gather(o, grid)
    after = parts_of(o)
    surf = surface_of(before, cell, grid)
    inner = ~surf
    assert inner.sum() > 100
    assert np.all(after["kind"][inner] == N.ROW_BRINE + 1)  # its own row again, KIND_SRC cleared
    # the same heat content at the brine's heat capacity
    Cs = float(mats[N.ROW_STEAM].th[0])
    Cb = float(mats[N.ROW_BRINE].th[0])  # s = 1: the concentrate's own
    h0 = Cs * (parts["t"][inner].astype(np.float64) + 273.15)
    h1 = Cb * (after["t"][inner].astype(np.float64) + 273.15)
    assert np.allclose(h1, h0, rtol=1e-6)


def test_steam_rises_through_water_at_the_bubble_terminal_velocity():
    grid = (40, 32)
    params, field, parts, rows = insulated(grid, None, frac=(1.0, 0.8), row=N.ROW_WATER)
    params.crowd_jitter = 0.0
    params.advection_order = 1
    parts["t"] = np.float32(100.0)  # at the boiling point: no exchange, no steam made
    # a steam bubble (x = 0.01, between steam_condense and steam_boil) on top of some interior
    # water particles
    pick = np.flatnonzero((parts["pos"][:, 0] > 8) & (parts["pos"][:, 0] < 32) & (parts["pos"][:, 1] > 8) &
                          (parts["pos"][:, 1] < 18))[::17]
    bub = parts[pick].copy()
    bub["kind"] = steam_kind(N.ROW_WATER)
    bub["t"] = np.float32(100.0 + 0.01 * WATER_LC)
    bub["id"] = np.arange(len(parts), len(parts) + len(bub), dtype=np.uint32)
    allp = np.concatenate([parts, bub])
    o, g, mats = sim(grid, params, field, allp, rows)
    before, _ = first_then_gather(o, grid)
    after = parts_of(o)
    n = len(parts)
    dt = 1.0 / 240.0
    dy_w = after["pos"][pick, 1].astype(np.float64) - before["pos"][pick, 1]
    dy_s = after["pos"][n:, 1].astype(np.float64) - before["pos"][n:, 1]
    dx_w = after["pos"][pick, 0].astype(np.float64) - before["pos"][pick, 0]
    dx_s = after["pos"][n:, 0].astype(np.float64) - before["pos"][n:, 0]
    assert np.all((after["kind"][n:] & 255) == N.ROW_STEAM + 1)
    rw, rs = float(mats[N.ROW_WATER].rho), float(mats[N.ROW_STEAM].rho)
    vt = terminal_root(rs - rw, rw, math.exp(float(mats[N.ROW_WATER].ln_nu)), float(g.sep_tau), float(g.gravity))
    vt = min(vt, g.cfl / dt)
    # the bubble moves up by vt dt relative to the water it sits on, whatever the share of
    # steam around it; sideways they move together
    assert np.allclose(dy_s - dy_w, vt * dt, rtol=1e-3, atol=1e-5)
    assert np.allclose(dx_s, dx_w, atol=1e-6)
    # the share of steam reaches grid_prep where bubbles are, and nowhere else
    gb = o.field("GBASE", np.float32, (grid[0] * grid[1], 4))
    assert np.all(gb[:, 3] >= 0.0) and np.all(gb[:, 3] < 1.0) and (gb[:, 3] > 0.0).sum() >= len(bub)
    W = grid[0]
    cx = np.floor(before["pos"][n:, 0]).astype(int)
    cy = np.floor(before["pos"][n:, 1]).astype(int)
    near = np.zeros(grid[0] * grid[1], bool)
    for a in (-2, -1, 0, 1, 2):
        for b in (-2, -1, 0, 1, 2):
            near[(cy + b) * W + cx + a] = True
    assert np.all(gb[~near, 3] == 0.0)


def test_exchange_inside_a_cell_holding_steam_runs_at_steam_mixing():
    grid = (40, 32)
    params, field, parts, rows = insulated(grid, None, frac=(1.0, 0.8), row=N.ROW_WATER, heat_rate=1.0)
    params.steam_mixing = 500.0
    parts["t"] = np.float32(100.0)
    steam = (np.arange(len(parts)) % 7 == 0) & (parts["pos"][:, 0] < 0.4 * grid[0])  # bubbles in the left part
    parts["kind"][steam] = steam_kind(N.ROW_WATER)
    o, g, mats = sim(grid, params, field, parts, rows)
    o.begin_batch(1 / 240.0, 1)
    o.substep(0)
    gb = o.field("GBASE", np.float32, (grid[0] * grid[1], 4))
    dh = o.field("DHEAT", np.float32, (grid[0] * grid[1], 4))
    hc = o.field("HCELL", np.int32, (grid[0] * grid[1], 2))
    k = (1.0 / 240.0) * params.heat_rate
    fast = np.float32(1.0 - math.exp(-k * params.steam_mixing))
    with_steam = (gb[:, 3] > 0.0) & (hc[:, 0] > 0)
    without = (gb[:, 3] == 0.0) & (hc[:, 0] > 0)
    assert with_steam.sum() > 100 and without.sum() > 20
    assert np.allclose(dh[with_steam, 1], fast, rtol=1e-5)
    assert np.all(dh[without, 1] < 0.5 * fast)


def puff_tank(grid, n=8, swirl=0.0, life=2.0):
    params = cases.dam_params(cell_size=0.005)
    params.steam_swirl = swirl
    params.steam_life = life
    import sprout_amd.scene as sc
    scene = sc.Scene(grid)
    field = scene.rasterize()
    p = np.zeros(n, N.PARTICLE_DTYPE)
    # over the middle of the floor, which is the nearest solid: the air's rise is not turned
    p["pos"][:, 0] = np.linspace(0.45, 0.55, n) * grid[0]
    p["pos"][:, 1] = 0.25 * grid[1]
    p["kind"] = N.KIND_PUFF
    p["c"][:, 0] = 1.0
    p["id"] = np.arange(n, dtype=np.uint32)
    return params, field, p, [N.ROW_WATER, N.ROW_STEAM]


def test_puff_takes_up_the_air_rise_and_dies_at_the_end_of_its_life():
    grid = (192, 96)
    params, field, p, rows = puff_tank(grid, life=0.1)
    o, g, mats = sim(grid, params, field, p, rows)
    dt = 1.0 / 240.0
    rise = params.steam_rise / params.cell_size
    k = 1.0 - math.exp(-dt / 0.08)
    y0 = p["pos"][:, 1].astype(np.float64)
    ys = y0.copy()
    v_model = 0.0
    for step in range(23):
        o.begin_batch(dt, 1)
        o.substep(1 if step > 0 else 0)
        q = parts_of(o)
        assert np.all(q["kind"] == N.KIND_PUFF), step
        v_model = v_model * (1.0 - k) + rise * k
        ys += v_model * dt
        # the vertical velocity: rise (1 - (1 - k)^n)
        assert np.allclose(q["vel"][:, 1], rise * (1.0 - (1.0 - k) ** (step + 1)), rtol=1e-4), step
        assert np.allclose(q["pos"][:, 1], ys, rtol=1e-5), step
        lean = q["vel"][:, 0] / q["vel"][:, 1]
        assert np.all(np.abs(lean) <= 0.15 + 1e-6)
    assert np.allclose(q["s"], 23 * dt, rtol=1e-5)
    for step in range(2):
        o.begin_batch(dt, 1)
        o.substep(1)
    assert np.all(parts_of(o)["kind"] == 0)  # gone after steam_life = 0.1 s


def test_air_eddies_are_divergence_free():
    grid = (160, 160)
    params, field, p, rows = puff_tank(grid, n=1, swirl=0.3)
    o, g, _ = sim(grid, params, field, p, rows)
    o.begin_batch(1 / 240.0, 1)
    o.substep(0)  # some time in the step state for the noise to evolve
    h = 0.05
    worst = 0.0
    scale = 0.0
    rng = np.random.default_rng(3)
    for _ in range(40):
        x, y = rng.uniform(70, 90, 2)
        vxp = o.air_velocity(x + h, y, 0.0)[0]
        vxm = o.air_velocity(x - h, y, 0.0)[0]
        vyp = o.air_velocity(x, y + h, 0.0)[1]
        vym = o.air_velocity(x, y - h, 0.0)[1]
        div = (vxp - vxm + vyp - vym) / (2 * h)
        dvx = abs(vxp - vxm) / (2 * h)
        worst = max(worst, abs(div))
        scale = max(scale, dvx)
    assert scale > 0.05  # the eddies are there
    assert worst < 1e-2 * scale, (worst, scale)
