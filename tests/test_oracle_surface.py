"""Pins of the oracle's surface-tension curvature (surf_smooth.comp + surf_curv.comp, the
curvature of the smoothed liquid fraction from its node normals), on the CPU, against
geometry:

  * a disc of liquid of radius R: over the interface band (cells within 1.5 of the rim) the
    mean of curvature x distance to the centre is 1 (the curvature of a circle, 1 / r) within
    5 %, for two radii, positive for a convex blob;
  * a straight surface (a half-plane of liquid, at 0 and 30 degrees): no curvature in the band.
A dropped half, a sign error or a stencil slip fails one of them. Of the pressure jump
(surf_apply.comp) the curvature drives, by Laplace's law at the reference's own bar
(validate.rs:1270-1360): a resting water drop holds sigma / R inside at two radii, and an
oil disc in water holds the oil-water pair tension over R. And of the explicit viscosity
step (visc_setup + visc_apply mode 2): a shear wave in the bulk decays by the discrete
Laplacian's Fourier factor, on the u and on the v faces.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim


def curvature_of(phi):
    H, W = phi.shape
    grid = (W, H)
    params, field, parts, rows = cases.dam_break(grid, frac=(0.1, 0.1))
    g, mats = N.derive_grid_params(params, W, H, rows)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    cell = np.zeros((H, W, 4), np.float32)
    cell[..., 0] = phi
    cell[..., 1] = 1.0
    cell[..., 3] = phi - 0.5
    o.put_field("CELL", cell.reshape(-1, 4))
    comp = np.zeros((H, W, 4), np.float32)
    comp[..., 0] = np.where(phi > 0, 1.0, 0.0)
    comp[..., 2] = -1.0
    comp[..., 3] = phi
    o.put_field("COMP", comp.reshape(-1, 4))
    o.stage("SURF")
    return o.field("SURF", np.float32, (H, W, 4))[..., 2].astype(np.float64)


@pytest.mark.parametrize("R", [12.0, 20.0])
def test_disc_has_the_curvature_of_a_circle(R):
    W = H = 64
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    r = np.hypot(x - 32.0, y - 32.0)
    k = curvature_of(np.clip(R + 0.5 - r, 0.0, 1.0).astype(np.float32))
    band = (k != 0.0) & (np.abs(r - R) < 1.5)
    assert band.sum() > 100
    assert np.mean(k[band] * r[band]) == pytest.approx(1.0, abs=0.05)


@pytest.mark.parametrize("deg", [0.0, 30.0])
def test_straight_surface_has_none(deg):
    W = H = 64
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    a = np.radians(deg)
    d = (y - 30.0) * np.cos(a) - (x - 32.0) * np.sin(a)  # signed distance above the surface
    k = curvature_of(np.clip(0.5 - d, 0.0, 1.0).astype(np.float32))
    inner = (np.abs(d) < 1.5) & (x > 8) & (x < W - 8) & (y > 8) & (y < H - 8)
    assert inner.sum() > 40
    assert np.abs(k[inner]).max() < 0.02


def shear_wave_step(component, nu, rho, kw, dt=1.0 / 240.0):
    """One explicit viscosity step (visc_setup + visc_apply mode 2) of a shear wave in a full
    tank: u = sin(k y) on the u faces, or v = sin(k x) on the v faces. Returns (before,
    after) of that component."""
    W = H = 64
    params, field, parts, rows = cases.dam_break((W, H), frac=(0.1, 0.1), viscosity=1e-6)
    g, mats = N.derive_grid_params(params, W, H, rows)
    assert g.visc_mode == 1  # explicit
    o = OracleSim((W, H), W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    st = o.field("STEP", N.STEP_DTYPE, 1)
    st["dt"] = dt
    o.put_field("STEP", st)
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    cell = np.zeros((H, W, 4), np.float32)
    cell[..., 0] = 1.0
    cell[..., 1] = rho
    cell[..., 2] = nu * rho  # dynamic viscosity over water's density
    cell[..., 3] = 0.5
    o.put_field("CELL", cell.reshape(-1, 4))
    vel = np.zeros((H, W, 4), np.float32)
    if component == 0:
        vel[..., 0] = np.sin(kw * (y + 0.5))
    else:
        vel[..., 1] = np.sin(kw * (x + 0.5))
    o.put_field("VEL", vel.reshape(-1, 4))
    o.stage("COEF")
    o.stage("VISC")
    got = o.field("VEL", np.float32, (H, W, 4)).astype(np.float64)
    return vel[..., component].astype(np.float64), got[..., component]


@pytest.mark.parametrize("component,nu,rho,period", [(0, 0.5, 1.0, 16.0), (1, 0.3, 1.3, 10.0)])
def test_explicit_viscosity_decays_a_shear_wave(component, nu, rho, period):
    """Pin of the explicit viscosity step (visc_setup + visc_apply mode 2): in the bulk a
    shear wave of wavenumber k loses dt nu (2 - 2 cos k) of itself per step, the decay of the
    discrete Laplacian's Fourier mode (nu = mu / rho)."""
    kw = 2.0 * np.pi / period
    dt = 1.0 / 240.0
    before, after = shear_wave_step(component, nu, rho, kw, dt)
    inner = (slice(8, 56), slice(8, 56))
    change = after[inner] - before[inner]
    want = -dt * nu * (2.0 - 2.0 * np.cos(kw)) * before[inner]
    assert np.abs(want).max() > 1e-4
    assert np.allclose(change, want, rtol=0.0, atol=2e-3 * np.abs(want).max())


def test_collide_keeps_particles_out_of_the_floor():
    """Pin of `collide` (particles.comp:51-65 and the advect loop): particles driven into the
    floor end at least the wall margin above its surface (the solid border is WALL cells
    thick), with no velocity into it; those that never reach it move freely."""
    W = H = 48
    params, field, parts, rows = cases.dam_break((W, H), frac=(0.5, 0.5))
    g, mats = N.derive_grid_params(params, W, H, rows)
    o = OracleSim((W, H), W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    st = o.field("STEP", N.STEP_DTYPE, 1)
    st["dt"] = 1.0 / 240.0
    o.put_field("STEP", st)
    # a grid of empty faces: the gather finds no live face and the particles keep their
    # own velocity
    vel = np.zeros((W * H, 4), np.float32)
    vel[:, 2:] = 1e30
    o.put_field("VEL", vel)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    p["vel"][:, 0] = 0.0
    p["vel"][:, 1] = -400.0  # 1.67 cells per substep downwards, below the speed cap
    p["c"][:] = 0.0
    o.put_field("PARTICLES", p)
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    q = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    floor = float(cases.sc.WALL)
    step = 400.0 / 240.0
    assert 400.0 < g.ext[0]
    hit = p["pos"][:, 1] - step < floor + g.wall_margin
    assert hit.sum() > 50 and (~hit).sum() > 50
    assert np.all(q["pos"][:, 1] >= floor + g.wall_margin - 1e-4)
    free = p["pos"][:, 1] - step > floor + 0.5  # ends well clear of the floor's margin
    assert np.allclose(q["pos"][free, 1], p["pos"][free, 1] - step, atol=1e-4)
    assert np.all(q["vel"][hit, 1] >= 0.0)  # no velocity into the floor


# ---- the surface-tension pressure jump (surf_apply.comp), against Laplace's law -------------

def capillary_case(kind, grid, rad):
    """validate.rs:1270-1360 on a smaller grid: 1 mm cells, no gravity; a water drop of
    radius `rad` at the centre of an empty tank, or an oil disc of that radius inside water
    filling 260/288 of the tank."""
    from sprout_amd import scene as sc
    W, H = grid
    params = N.default_params()
    params.cell_size = 0.001
    params.gravity = 0.0
    c = (W / 2, H / 2)
    scene = sc.Scene(grid)
    if kind == "drop":
        ps = sc.fill_rect(scene, (c[0] - rad, c[1] - rad, c[0] + rad, c[1] + rad), N.ROW_WATER)
        ps = ps[np.hypot(ps["pos"][:, 0] - c[0], ps["pos"][:, 1] - c[1]) < rad]
        rows = [N.ROW_WATER]
    else:
        ps = sc.fill_rect(scene, (0.0, 0.0, float(W), H * 260.0 / 288.0), N.ROW_WATER)
        ps["kind"][np.hypot(ps["pos"][:, 0] - c[0], ps["pos"][:, 1] - c[1]) < rad] = N.ROW_OIL + 1
        rows = [N.ROW_WATER, N.ROW_OIL]
    ps = cases.settle(ps, params)
    g, mats = N.derive_grid_params(params, W, H, rows)
    g.n_slots = len(ps)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(scene.rasterize())
    o.reset(ps)
    return o, g, params, c


def run_for(o, g, seconds):
    first = True
    while True:
        o.begin_batch(16 * g.dt_max, 16)
        for _ in range(16):
            o.substep(0 if first else 1)
            first = False
        S = o.field("STEP", N.STEP_DTYPE, 1)[0]
        if S["advanced"] >= seconds:
            return S


def pressure_rings(o, S, c, W, H):
    pres = o.field("PRES", np.float32, (W * H, 2))[:, 0].reshape(H, W).astype(np.float64) / float(S["dt"])
    j, i = np.mgrid[0:H, 0:W]
    return pres, np.hypot(i + 0.5 - c[0], j + 0.5 - c[1])


@pytest.mark.parametrize("rad", [16.0, 24.0])
def test_resting_drop_holds_the_laplace_pressure(rad):
    """Inside a resting drop (the mean over r < R/2 after 0.1 s) the pressure is sigma / R
    within 10 % (the reference's own bar, validate.rs:1301-1306), at two radii, so a jump
    that does not scale as 1/R fails; the spurious currents stay under 3 % of
    sqrt(sigma / (rho R))."""
    W, H = 128, 96
    o, g, params, c = capillary_case("drop", (W, H), rad)
    S = run_for(o, g, 0.1)
    pres, d = pressure_rings(o, S, c, W, H)
    sigma = params.mat[N.ROW_WATER].tension / (1000.0 * params.cell_size ** 3)
    ratio = pres[d < 0.5 * rad].mean() / (sigma / rad)
    assert abs(ratio - 1.0) < 0.10, ratio
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    rms = np.sqrt(np.mean((p["vel"].astype(np.float64) ** 2).sum(1)))
    assert rms / np.sqrt(sigma / rad) < 0.03


def test_oil_disc_in_water_holds_the_pair_tension_jump():
    """An oil disc inside water: the pressure inside (r < R/2) less that of a ring outside
    (R + 8 <= r < R + 16) is the oil-water tension over R within 10 % (validate.rs:
    1320-1355); the pair tension, not either liquid's own, sets it."""
    W, H = 128, 96
    rad = 16.0
    o, g, params, c = capillary_case("oil", (W, H), rad)
    S = run_for(o, g, 0.1)
    pres, d = pressure_rings(o, S, c, W, H)
    aq, oily = params.mat[N.ROW_WATER].group, params.mat[N.ROW_OIL].group
    sigma = params.pair_tension[aq][oily] / (1000.0 * params.cell_size ** 3)
    jump = pres[d < 0.5 * rad].mean() - pres[(d >= rad + 8) & (d < rad + 16)].mean()
    assert abs(jump / (sigma / rad) - 1.0) < 0.10, jump / (sigma / rad)
    assert abs(params.pair_tension[aq][oily] - params.mat[N.ROW_WATER].tension) > 0.01
