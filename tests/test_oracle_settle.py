"""Pins of the oracle's settling restatement (settle.comp, the family occupancy of
grid_prep.comp:69-70 and 147, the particle side of particles.comp:332-351, FluidSim::settles),
on the CPU:

  * tables: the grain in cells, CAP_SETTLE only for a concentrate with grains under gravity;
  * the family occupancy is the concentrate's share of the cell times its fill;
  * in a uniform suspension the flux through every inner face is the Kynch flux
    v_t s (1 - s), v_t the terminal speed of a grain from the drag law (Stokes to Newton,
    closed form): an inner cell's gain factor is dt v_t s and its loss factor dt v_t (1 - s),
    the top cell gains nothing and the bottom cell loses nothing; a lighter concentrate
    creams upwards instead;
  * the limiter: no cell gives more than it holds or takes more than it has room for;
  * a particle takes s + (1 - s) gain - s loss from uniform factors, and a water particle of
    the primary family's diluent becomes the concentrate at the gain;
  * whole substeps: the concentrate is conserved and stays in [0, 1], and it moves down.
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


def set_dt(o, dt):
    st = o.field("STEP", N.STEP_DTYPE, 1)
    st["dt"] = dt
    o.put_field("STEP", st)


def terminal(drho, rho_c, nu_c, d, g):
    """The drag law's terminal speed (common.glsl:370-374): the positive root of
    0.44 rho_c v^2 + 24 nu_c rho_c / d v = 4/3 |drho| g d."""
    qa, qb, qc = 0.44 * rho_c, 24.0 * nu_c * rho_c / d, 4.0 / 3.0 * abs(drho) * g * d
    return (-qb + np.sqrt(qb * qb + 4.0 * qa * qc)) / (2.0 * qa)


def test_tables_and_caps():
    grid = (64, 48)
    p = cases.mud_params(cases.dam_params(), grain=100e-6)
    g, m = N.derive_grid_params(p, 64, 48, [N.ROW_WATER, N.ROW_HONEY])
    assert g.caps & N.CAP_SETTLE and g.n_fam == 1
    assert m[N.ROW_HONEY].grain == pytest.approx(100e-6 / p.cell_size)
    p.gravity = 0.0
    g, _ = N.derive_grid_params(p, 64, 48, [N.ROW_WATER, N.ROW_HONEY])
    assert not g.caps & N.CAP_SETTLE
    p = cases.mud_params(cases.dam_params(), grain=0.0)
    g, _ = N.derive_grid_params(p, 64, 48, [N.ROW_WATER, N.ROW_HONEY])
    assert not g.caps & N.CAP_SETTLE
    g, _ = N.derive_grid_params(cases.mud_params(cases.dam_params()), 64, 48, [N.ROW_WATER])
    assert not g.caps & N.CAP_SETTLE


def test_family_occupancy_is_share_times_fill():
    grid = (64, 48)
    W, H = grid
    o, g, m = make(grid, cases.mud_column(grid, water_above=0.2))
    set_dt(o, 1.0 / 240.0)
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    acc = o.field("ACC", np.int32, (W * H, g.acc_stride)).astype(np.float64)
    cell = o.field("CELL", np.float32, (W * H, 4))
    occ = o.field("FAMOCC", np.float32, W * H)
    cw = acc[:, 4]
    water = acc[:, m[N.ROW_WATER].chan_w] if m[N.ROW_WATER].chan_w >= 0 else np.zeros_like(cw)
    share = np.where(cw > 0, np.clip((cw - water) / np.maximum(cw, 1), 0, 1), 0.0)
    want = share * cell[:, 0]
    assert np.count_nonzero(occ) > 100
    assert np.allclose(occ, want, rtol=1e-5, atol=1e-6)
    # pure mud cells: their whole fill
    pure = (water == 0) & (cw > 0)
    assert np.allclose(occ[pure], cell[pure, 0], rtol=1e-6)


def uniform_column(grid, s, occ=1.0, x=(20, 44), y=(4, 30)):
    """comp.z = s and famocc = occ in a block of cells, nothing elsewhere."""
    W, H = grid
    comp = np.zeros((H, W, 4), np.float32)
    comp[..., 2] = -1.0
    fo = np.zeros((H, W), np.float32)
    comp[y[0]:y[1], x[0]:x[1], 2] = s
    comp[y[0]:y[1], x[0]:x[1], 0] = 1.0
    fo[y[0]:y[1], x[0]:x[1]] = occ
    return comp, fo


@pytest.mark.parametrize("density,s", [(2000.0, 0.5), (2000.0, 0.2), (500.0, 0.4)])
def test_uniform_suspension_carries_the_kynch_flux(density, s):
    grid = (64, 48)
    W, H = grid
    o, g, m = make(grid, cases.mud_column(grid, density=density, grain=100e-6))
    comp, fo = uniform_column(grid, s)
    o.put_field("COMP", comp.reshape(-1, 4))
    o.put_field("FAMOCC", fo.reshape(-1))
    dt = 1.0 / 240.0
    set_dt(o, dt)
    o.stage("SETTLE")
    ds = o.field("DSETTLE", np.float32, (H, W, 2)).astype(np.float64)
    c, dl = m[N.ROW_HONEY], m[N.ROW_WATER]
    vt = terminal(c.rho - dl.rho, dl.rho, np.exp(dl.ln_nu), c.grain, g.gravity)
    inner = ds[5:29, 21:43]
    assert np.allclose(inner[..., 0], dt * vt * s, rtol=1e-5)
    assert np.allclose(inner[..., 1], dt * vt * (1.0 - s), rtol=1e-5)
    top, bottom = ds[29, 21:43], ds[4, 21:43]
    if density > 1000.0:  # grains sink: the top cell gains nothing, the bottom loses nothing
        assert np.all(top[:, 0] == 0.0) and np.allclose(top[:, 1], dt * vt * (1 - s), rtol=1e-5)
        assert np.all(bottom[:, 1] == 0.0) and np.allclose(bottom[:, 0], dt * vt * s, rtol=1e-5)
    else:  # lighter than water: it creams upwards
        assert np.all(top[:, 1] == 0.0) and np.allclose(top[:, 0], dt * vt * s, rtol=1e-5)
        assert np.all(bottom[:, 0] == 0.0) and np.allclose(bottom[:, 1], dt * vt * (1 - s), rtol=1e-5)
    outside = np.ones((H, W), bool)
    outside[4:30, 20:44] = False
    assert np.all(ds[outside] == 0.0)


def test_flux_is_limited_by_holding_and_room():
    grid = (64, 48)
    W, H = grid
    # grains of a cell (about 77 cells/s) over a long substep (0.05 s): v_t s (1 - s) occ dt
    # exceeds what a cell holds, and the limiter decides
    o, g, m = make(grid, cases.mud_column(grid, grain=5e-3))
    s = 0.5
    comp, fo = uniform_column(grid, s, occ=0.6)
    o.put_field("COMP", comp.reshape(-1, 4))
    o.put_field("FAMOCC", fo.reshape(-1))
    dt = 0.05
    set_dt(o, dt)
    o.stage("SETTLE")
    ds = o.field("DSETTLE", np.float32, (H, W, 2)).astype(np.float64)
    inner = ds[5:29, 21:43]
    # flux = min(s occ, (1 - s) occ) / dt: the gain factor F dt / (occ (1 - s)) and the loss
    # factor F dt / (occ s), neither above 1 (no more than the room, no more than the holding)
    F = min(s * 0.6, (1 - s) * 0.6) / dt
    assert np.allclose(inner[..., 0], dt * F / (0.6 * (1 - s)), rtol=1e-5)
    assert np.allclose(inner[..., 1], dt * F / (0.6 * s), rtol=1e-5)
    assert inner[..., 0].max() <= 1.0 + 1e-6 and inner[..., 1].max() <= 1.0 + 1e-6


def test_particles_take_the_factors():
    grid = (64, 48)
    W, H = grid
    o, g, m = make(grid, cases.mud_column(grid, water_above=0.2))
    set_dt(o, 1.0 / 240.0)
    # a still grid, full cells everywhere (so nothing else moves the particles' s)
    vel = np.zeros((W * H, 4), np.float32)
    o.put_field("VEL", vel)
    comp = np.zeros((W * H, 4), np.float32)
    comp[:, 2] = -1.0  # no family-0 concentration to mix with (mix_rate is 0 anyway)
    o.put_field("COMP", comp)
    a, b = 0.03, 0.05
    ds = np.tile(np.array([a, b], np.float32), (W * H, 1))
    o.put_field("DSETTLE", ds)
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    mud = (before["kind"] & 255) == N.ROW_HONEY + 1
    s0 = before["s"][mud].astype(np.float64)
    assert np.allclose(after["s"][mud], s0 + (1 - s0) * a - s0 * b, rtol=1e-5)
    # water (the diluent of family 0) becomes mud at the gain
    water = (before["kind"] & 255) == N.ROW_WATER + 1
    assert water.sum() > 100
    assert np.all((after["kind"][water] & 255) == N.ROW_HONEY + 1)
    assert np.allclose(after["s"][water], a, rtol=1e-5)


def concentrate(p):
    return float(np.sum(p["s"][(p["kind"] & 255) == N.ROW_HONEY + 1], dtype=np.float64))


def test_whole_substeps_conserve_and_settle():
    grid = (64, 48)
    built = cases.mud_column(grid, grain=200e-6, water_above=0.2)
    o, g, m = make(grid, built)
    p0 = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    total0 = concentrate(p0)
    mud0 = (p0["kind"] & 255) == N.ROW_HONEY + 1
    y0 = float(np.sum(p0["s"][mud0] * p0["pos"][mud0, 1], dtype=np.float64)) / total0
    steps = 48
    o.begin_batch(steps / 240, steps)
    for k in range(steps):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    mud = (p["kind"] & 255) == N.ROW_HONEY + 1
    assert np.all((p["s"] >= 0.0) & (p["s"] <= 1.0))
    assert concentrate(p) == pytest.approx(total0, rel=2e-4)
    y1 = float(np.sum(p["s"][mud] * p["pos"][mud, 1], dtype=np.float64)) / concentrate(p)
    assert y1 < y0 - 0.05  # the grains went down
    # and the clear water above took none of them from the top down: the topmost mud
    # particles lost concentration
    top = mud & (p["pos"][:, 1] > np.percentile(p["pos"][mud, 1], 90))
    assert np.mean(p["s"][top]) < 0.5
