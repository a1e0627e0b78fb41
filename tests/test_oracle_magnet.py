"""Pins of the oracle's magnetism restatement (magnet.comp, magnet_field of common.glsl,
the susceptibility and resting rule of grid_prep.comp, the tables of mod.rs:2258-2264,
2313 and 2749-2751), on the CPU, against what the physics fixes:

  * tables: CAP_MAGNET only with a magnet and a susceptible liquid; the largest
    susceptibility, its saturation and 1 / (rho_w dx^2) in `mag`;
  * the field of a point dipole, B = B0 (R / r)^3 (3 (m.e) e - m) / 2: on its axis B0 (R / r)^3,
    across it half that and opposite, constant inside the rim, a superposition for two magnets;
  * the energy density of a saturating linear liquid, psi = -mu0 Ms^2 / chi_e (x - ln(1 + x)),
    x = chi_e H / Ms, chi_e = chi / (1 + chi / 3), in grid units, at every cell (both the
    series and the logarithm branch);
  * the face force is the discrete gradient -dt s (psi_c - psi_n) / rho, s the face's share of
    the largest susceptibility (the mean of two liquid cells', the liquid cell's own at a
    surface); dead, closed and coefficient-free faces get none;
  * whole substeps: a ferrofluid pool rises towards a magnet above it.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim

MU0 = 1.25663706e-6
ROW = cases.ROW_FERRO


def make(grid, built):
    params, field, parts, rows, magnets = built
    W, H = grid
    g, mats = N.derive_grid_params(params, W, H, rows, magnets=magnets)
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


def dipole(x, y, magnets):
    """B (T) of the magnets at points (x, y), float64."""
    Bx = np.zeros_like(x)
    By = np.zeros_like(x)
    for m in magnets:
        dx, dy = x - m.pos[0], y - m.pos[1]
        d = np.hypot(dx, dy)
        R = max(m.radius, 0.5)
        r = np.maximum(d, R)
        ex = np.where(d > 1e-4, dx / np.maximum(d, 1e-30), m.dir[0])
        ey = np.where(d > 1e-4, dy / np.maximum(d, 1e-30), m.dir[1])
        me = m.dir[0] * ex + m.dir[1] * ey
        k = m.strength * (R / r) ** 3 * 0.5
        Bx += k * (3 * me * ex - m.dir[0])
        By += k * (3 * me * ey - m.dir[1])
    return Bx, By


def test_tables_and_caps():
    p = cases.ferro_params(cases.dam_params())
    mg = [N.Magnet.new((10.0, 20.0), 0.2, 90.0, 0.3)]
    g, m = N.derive_grid_params(p, 64, 48, [ROW], magnets=mg)
    assert g.caps & N.CAP_MAGNET
    assert tuple(g.mag) == pytest.approx((1.0, 2.0, 0.05, 1.0 / (998.0 * p.cell_size ** 2)), rel=1e-6)
    assert tuple(g.magnet[0]) == pytest.approx((10.0, 20.0, 0.5, 0.3))  # the radius at least half a cell
    assert tuple(g.magnet_dir[0]) == pytest.approx((1.0, 0.0), abs=1e-7)  # 90 degrees clockwise from up
    assert m[ROW].gr[3] == 2.0
    g, _ = N.derive_grid_params(p, 64, 48, [ROW])
    assert not g.caps & N.CAP_MAGNET  # no magnet
    g, _ = N.derive_grid_params(p, 64, 48, [N.ROW_WATER], magnets=mg)
    assert not g.caps & N.CAP_MAGNET  # nothing magnetisable


def psi_closed_form(Bx, By, chi, sat, dx):
    H = np.hypot(Bx, By) / MU0
    ce = chi / (1.0 + chi / 3.0)
    Ms = sat / MU0
    x = ce * H / Ms
    return -MU0 * Ms * Ms / ce * (x - np.log1p(x)) / (998.0 * dx * dx)


@pytest.mark.parametrize("two", [False, True])
def test_energy_density_of_the_dipole_field(two):
    grid = (96, 64)
    W, H = grid
    built = list(cases.ferro_pool(grid))
    if two:
        built[4] = built[4] + [N.Magnet.new((20.0, 40.0), 3.0, 135.0, 0.2)]
    o, g, _ = make(grid, tuple(built))
    o.stage("MAGNET", 0)
    psi = o.field("MAGC", np.float32, (H, W, 2))[..., 1].astype(np.float64)
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    Bx, By = dipole(x, y, built[4])
    want = psi_closed_form(Bx, By, 2.0, 0.05, built[0].cell_size)
    ce = 2.0 / (1.0 + 2.0 / 3.0)
    xs = ce * np.hypot(Bx, By) / MU0 / (0.05 / MU0)
    # float32 x - ln(1 + x) just above the series branch (x >= 0.02) loses the rounding of
    # 1 + x (half an ulp of 1) against x^2 / 2: a relative error up to 6e-8 / (x^2 / 2)
    tol = 2e-5 + np.where(xs >= 0.02, 1.5e-7 / np.maximum(xs, 0.02) ** 2, 0.0)
    assert np.all(np.abs(psi - want) <= tol * np.abs(want) + 1e-12)
    # both branches: near the magnet x is large, far away it is below 0.02
    assert (xs < 0.02).sum() > 100 and (xs > 0.5).sum() > 10
    if not two:
        m = built[4][0]
        # on the axis above the magnet B0 (R / r)^3, across it half that
        r = 10.0
        Ba = dipole(np.array([m.pos[0]]), np.array([m.pos[1] + r]), built[4])
        Bc = dipole(np.array([m.pos[0] + r]), np.array([m.pos[1]]), built[4])
        assert Ba[1][0] == pytest.approx(m.strength * (m.radius / r) ** 3)
        assert Bc[1][0] == pytest.approx(-0.5 * m.strength * (m.radius / r) ** 3)


def test_face_force_is_the_discrete_gradient():
    grid = (64, 48)
    W, H = grid
    o, g, _ = make(grid, cases.ferro_pool(grid))
    rng = np.random.default_rng(2)
    psi = rng.normal(0.0, 50.0, (H, W)).astype(np.float32)
    chi = np.full((H, W), 2.0, np.float32)
    chi[:, 40:] = 1.0  # a weaker liquid on the right
    magc = np.stack([chi, psi], -1)
    o.put_field("MAGC", magc.reshape(-1, 2))
    cell = np.zeros((H, W, 4), np.float32)
    cell[..., 0] = 1.0
    cell[..., 1] = 1.2
    cell[..., 3] = 0.5
    cell[30:, :, 3] = -0.5  # air above row 30
    o.put_field("CELL", cell.reshape(-1, 4))
    coef = np.zeros((H, W, 4), np.float32)
    coef[..., 0] = 1.0
    coef[..., 1] = 1.0
    coef[..., 2] = 1.3
    coef[..., 3] = 0.9
    coef[10, 10, 0] = 0.0  # no coefficient: no force
    o.put_field("COEF", coef.reshape(-1, 4))
    vel = np.zeros((H, W, 4), np.float32)
    vel[20, 20, 2] = -1e30  # a closed face
    o.put_field("VEL", vel.reshape(-1, 4))
    dt = 1.0 / 240.0
    set_dt(o, dt)
    o.stage("MAGNET", 1)
    got = o.field("VEL", np.float32, (H, W, 4)).astype(np.float64)
    c64 = chi.astype(np.float64)
    p64 = psi.astype(np.float64)
    liquid = cell[..., 3] > 0
    # u faces between (i - 1, j) and (i, j)
    su = np.where(liquid[:, 1:] == liquid[:, :-1], 0.5 * (c64[:, 1:] + c64[:, :-1]),
                  np.where(liquid[:, 1:], c64[:, 1:], c64[:, :-1])) / 2.0
    du = -dt * su * (p64[:, 1:] - p64[:, :-1]) / 1.3
    sv = np.where(liquid[1:] == liquid[:-1], 0.5 * (c64[1:] + c64[:-1]), np.where(liquid[1:], c64[1:], c64[:-1])) / 2.0
    dv = -dt * sv * (p64[1:] - p64[:-1]) / 0.9
    mu = np.ones((H, W - 1), bool)
    mu[0, :] = False  # row 0 is skipped (cells with y < 1)
    mu[10, 9] = mu[20, 19] = False
    assert np.allclose(got[:, 1:, 0][mu], du[mu], rtol=1e-5, atol=1e-9)
    assert got[10, 10, 0] == 0.0 and got[20, 20, 0] == 0.0
    mv = np.ones((H - 1, W), bool)
    mv[:, 0] = False
    assert np.allclose(got[1:, :, 1][mv], dv[mv], rtol=1e-5, atol=1e-9)


def test_pool_rises_towards_the_magnet():
    grid = (64, 48)
    W, H = grid
    built = cases.ferro_pool(grid, strength=0.5, magnet_gap=5.0)
    o, g, _ = make(grid, built)
    p0 = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    mid0 = np.abs(p0["pos"][:, 0] - W / 2) < 4
    top0 = p0["pos"][mid0, 1].max()
    steps = 48
    o.begin_batch(steps / 240, steps)
    for k in range(steps):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    assert np.all(np.isfinite(p["pos"]))
    mid = np.abs(p["pos"][:, 0] - W / 2) < 4
    far = np.abs(p["pos"][:, 0] - W / 2) > 20
    # under the magnet the surface rises; far from it, it does not
    assert p["pos"][mid, 1].max() > top0 + 1.0
    assert np.percentile(p["pos"][far, 1], 99) < top0 + 0.6


def resting_level(pull_x, with_magnet=True):
    """grid_prep's level of a cell a little under half full of ferrofluid beside a packed
    column of it (cell weights put straight into the accumulator: the column at x = 20, the
    cell at (21, 15), nothing below it), with the last substep's magnetic energy density
    psi = A x, A chosen so that the horizontal pull -km dpsi/dx is `pull_x` times gravity
    (positive: to the right)."""
    grid = (48, 32)
    W, H = grid
    params, field, parts, rows, magnets = cases.ferro_pool(grid)
    o, g, mats = make(grid, (params, field, parts, rows, magnets if with_magnet else []))
    assert bool(g.caps & N.CAP_MAGNET) == with_magnet
    o.begin_batch(1 / 240, 1)
    o.stage("STEP")
    acc = np.zeros((H, W, g.acc_stride), np.int32)
    acc[10:21, 20, 4] = int(1.0 * N.PPC * 1048576)
    phi = 0.4
    acc[15, 21, 4] = int(phi * N.PPC * 1048576)
    o.put_field("ACC", acc.reshape(-1))
    row = cases.ROW_FERRO
    chi = mats[row].gr[3]
    rho = mats[row].rho
    km = chi / g.mag[1] / max(rho, 0.05) if with_magnet else 1.0
    A = -pull_x * g.gravity / km
    x = np.arange(W, dtype=np.float64)
    magc = np.zeros((H, W, 2), np.float32)
    magc[..., 1] = (A * x)[None, :]
    o.put_field("MAGC", magc.reshape(-1, 2))
    o.stage("GRID_PREP")
    cell = o.field("CELL", np.float32, (H, W, 4))
    return float(cell[15, 21, 3]), float(cell[15, 21, 0]), float(cell[15, 20, 3])


def test_magnetic_resting_rule():
    """grid_prep.comp:189-222: a cell not full enough to count as liquid rests on a packed
    neighbour when the pull on it points there; with a magnet the pull is gravity plus the
    magnetic force -chi/rho grad psi of the last substep (211-221), along its larger component.
    Here the cell's only packed neighbour is on its left: pulled left harder than gravity it
    rests (level 0.05), pulled right, or left but weaker than gravity (so down, onto nothing),
    or with no magnet in the tank, it keeps its own level phi - 1/2."""
    lvl, phi, col = resting_level(-4.0)
    assert col > 0.0 and 0.25 < phi <= 0.5
    assert lvl == np.float32(0.05)
    for pull, mag in ((+4.0, True), (-0.25, True), (-4.0, False)):
        lvl, phi, _ = resting_level(pull, mag)
        assert lvl == pytest.approx(phi - 0.5, abs=1e-6), (pull, mag)
