"""Pins of the oracle's rheology restatement (rheo.comp, shear_rate / nn_eval in common.glsl,
the non-Newtonian share of grid_prep.comp, the plug and jammed cells of rigid.comp), on the
CPU.

  * the shear rate sqrt(2 D:D) of a uniform simple shear u = gamma y is gamma, and of a pure
    extension u = a x, v = -a y it is 2a (a stencil slip or a wrong neighbour fails both);
  * at those rates each model's effective viscosity is its constitutive law: Carreau
    mu (r + (1 - r)(1 + (lambda g)^2)^((n-1)/2)), Herschel-Bulkley mu g^(n-1) + tau_y / g,
    shear thickening mu min((1 + (lambda g)^2)^((n-1)/2), ratio), weighted by the cell's share
    of the material and capped at rigid_viscosity;
  * a yield-stress cell whose stress mu g is below its yield stress is a plug (rigid), above
    it flows; a still skin cell at the free surface holds;
  * whole substeps: a Bingham block slumps less than a Newtonian block of the same viscosity.
"""
import math

import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim

GRID = (40, 30)


def tank(rheology=2, **kw):
    grid = GRID
    W, H = grid
    params, field, parts, rows = cases.rheo_block(grid, rheology=rheology, block=(0.1, 0.2, 0.1, 0.2), **kw)
    g, mats = N.derive_grid_params(params, W, H, rows)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    o.begin_batch(1 / 240, 1)
    o.stage("STEP")
    return o, g, mats


def put_cells(o, mu0, rho=1.0, share=1.0, level=0.5, row=N.ROW_HONEY):
    W, H = GRID
    n = W * H
    cell = np.zeros((n, 4), np.float32)
    cell[:, 0] = 1.0
    cell[:, 1] = rho
    cell[:, 2] = mu0
    cell[:, 3] = level
    o.put_field("CELL", cell)
    rh = np.zeros((n, 4), np.float32)
    rh[:, 0] = share
    rh[:, 1] = row
    rh[:, 2] = 1.0
    o.put_field("RH", rh)


def put_vel(o, u, v):
    W, H = GRID
    vel = np.zeros((H, W, 4), np.float32)
    vel[..., 0] = u
    vel[..., 1] = v
    o.put_field("VEL", vel)


def shear_field(gamma):
    W, H = GRID
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    return gamma * (y + 0.5), np.zeros_like(x)  # u on u faces at (i, j + 1/2)


def extension_field(a, g_dt):
    W, H = GRID
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    # u at (i, j + 1/2), v at (i + 1/2, j); the pre-solve rate adds gravity's step to v
    return a * x, -a * y - g_dt


def mu_after(o):
    W, H = GRID
    o.stage("RHEO")
    return o.field("CELL", np.float32, (H, W, 4))[..., 2]


def law(model, mu0, g, n, lam, ratio, tau_y, share, cap):
    if model == 1:
        F = ratio + (1 - ratio) * (1 + (lam * g) ** 2) ** (0.5 * (n - 1))
        tau = 0.0
    elif model == 2:
        F = max(g, 1e-6) ** (n - 1)
        tau = tau_y
    else:
        F = min((1 + (lam * g) ** 2) ** (0.5 * (n - 1)), ratio)
        tau = 0.0
    F = 1 + share * (F - 1)
    return min(mu0 * F + share * tau / max(g, 1e-6), cap)


CASES = [
    (1, dict(flow_index=0.4, shear_time=0.5, visc_ratio=0.05, yield_stress=0.0)),   # Carreau, thinning
    (2, dict(flow_index=0.6, yield_stress=40.0)),                                   # Herschel-Bulkley
    (2, dict(flow_index=1.0, yield_stress=40.0)),                                   # Bingham
    (3, dict(flow_index=2.5, shear_time=0.3, visc_ratio=30.0, yield_stress=0.0)),   # thickening
]


@pytest.mark.parametrize("model,kw", CASES)
@pytest.mark.parametrize("flow", ["shear", "extension"])
@pytest.mark.parametrize("rate", [0.5, 3.0, 20.0])
@pytest.mark.parametrize("share", [1.0, 0.6])
def test_effective_viscosity_is_the_constitutive_law(model, kw, flow, rate, share):
    o, g, mats = tank(rheology=model, **kw)
    m = mats[N.ROW_HONEY]
    mu0 = 3.0
    rho = 1.1
    put_cells(o, mu0, rho=rho, share=share)
    S = o.field("STEP", N.STEP_DTYPE, 1)[0]
    g_dt = float(g.gravity) * float(S["dt"])
    if flow == "shear":
        u, v = shear_field(rate)
        gamma = rate
    else:
        u, v = extension_field(rate / 2.0, g_dt)
        gamma = rate
    put_vel(o, u, v)
    mu = mu_after(o)
    cap = rho * math.exp(float(g.ln_rigid))
    want = law(model, mu0, gamma, m.rheo_p[1], m.rheo_p[2], m.rheo_p[3], m.rheo_p[0], share, cap)
    inner = mu[5:-5, 5:-5]
    assert np.allclose(inner, want, rtol=2e-5), (inner.min(), inner.max(), want)


def test_yield_stress_cells_are_plugs_below_their_yield_stress():
    tau_y = 40.0  # Pa: 1603 cells^2/s^2 at 5 mm cells
    for mu0, gamma, plug in ((0.3, 0.2, True), (50.0, 60.0, False)):
        o, g, mats = tank(rheology=2, flow_index=1.0, yield_stress=tau_y)
        put_cells(o, mu0)
        u, v = shear_field(gamma)
        put_vel(o, u, v)
        o.stage("RIGID", 0)
        W, H = GRID
        rl = o.field("RL", np.int32, (H, W, 2))[..., 0]
        ty = mats[N.ROW_HONEY].rheo_p[0]
        assert (mu0 * gamma < ty) == plug
        inner = rl[5:-5, 5:-5]
        assert np.all(inner >= 0) if plug else np.all(inner < 0)


def test_still_skin_at_the_free_surface_holds():
    o, g, mats = tank(rheology=0, skin=0.5, yield_stress=0.0)
    W, H = GRID
    put_cells(o, 0.3)
    cell = o.field("CELL", np.float32, (H, W, 4))
    cell[H // 2:, :, 3] = -0.5  # liquid below, air above: a free surface at y = H / 2
    o.put_field("CELL", cell.reshape(-1, 4))
    put_vel(o, np.zeros((H, W)), np.zeros((H, W)) - 0.0)
    o.stage("RHEO")
    rh = o.field("RH", np.float32, (H, W, 4))
    surface = rh[H // 2 - 1, 5:-5, 3]
    deep = rh[H // 4, 5:-5, 3]
    assert np.all(surface == 1.0) and np.all(deep == 0.0)


def spread(o, parts0, steps):
    o.begin_batch(steps / 240.0, steps)
    for k in range(steps):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    x = p["pos"][:, 0]
    return float(np.percentile(x, 98) - np.percentile(x, 2)), float(np.max(p["pos"][:, 1]))


def test_bingham_block_slumps_less_than_newtonian():
    grid = (64, 48)
    W, H = grid
    out = {}
    for name, rheology, ty in (("newtonian", 0, 0.0), ("bingham", 2, 600.0)):
        params, field, parts, rows = cases.rheo_block(grid, rheology=rheology, yield_stress=ty, viscosity=0.02,
                                                      block=(0.4, 0.6, 0.07, 0.6))
        g, mats = N.derive_grid_params(params, W, H, rows)
        g.n_slots = len(parts)
        o = OracleSim(grid, W * H * 4)
        o.set_gparams(g, mats)
        o.set_solids(field)
        o.reset(parts)
        out[name] = spread(o, parts, 72)
        if rheology == 2:
            assert g.caps & 8 and g.caps & 4
    w0 = 0.2 * W
    assert out["newtonian"][0] > w0 + 2.0          # it slumped
    assert out["bingham"][0] < out["newtonian"][0] - 2.0
    assert out["bingham"][1] > out["newtonian"][1]  # and stands taller
