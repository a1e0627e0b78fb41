"""Pins of the oracle's drift-flux and family-exchange branches of the particle pass
(particles.comp:374-570, terminal velocity common.glsl:366-374), on the CPU.

Every cell of a block holds the same four particles on an exact 2 x 2 lattice, so the
composition the particles gather in the block's interior is the same everywhere. The grid
velocities are zeroed after grid_prep, so a particle's displacement over one particle pass is
its drift alone: dt times the slip.

* Between phase groups: water and oil three to one. A particle drifts at the terminal velocity
  of a droplet of the unresolved size (the positive root of (4/3) |drho| g d = 24 nu_c rho_c v
  / d + 0.44 rho_c v^2, solved here independently), towards where its density belongs, scaled
  by the share of the other group; so the volume flux of the drift cancels. A wrong drag
  coefficient, share, sign or continuous phase fails it.
* Within one group (oil and a lighter oily liquid, no interfacial tension): each drifts at the
  terminal velocity against the group's local mean density; in the Stokes regime the volume
  flux cancels.
* Family exchange (honey in water, concentrations 0.2 and 0.6 alternating): each particle's
  concentration relaxes towards the local mean as s + (m - s)(1 - exp(-k dt)) in one pass (the
  deviation from the mean shrinks by exp(-k dt)), and the concentration summed over the
  particles is conserved.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd import scene as sc
from oracle.binding import OracleSim

W, H = 48, 40
BLOCK = (8, 8, 40, 32)  # x0, y0, x1, y1 (cells) of the block of liquid
INNER = (14, 14, 34, 26)  # cells whose 3 x 3 neighbourhood lies inside the block, with margin


def lattice_particles(rows_of_slot, conc_of_slot=None):
    """Four particles per cell of BLOCK at the 2 x 2 lattice points; slot k = (k & 1, k >> 1)
    gets row rows_of_slot[k] (and concentration conc_of_slot[k])."""
    out = []
    for k in range(4):
        ox, oy = 0.25 + 0.5 * (k & 1), 0.25 + 0.5 * (k >> 1)
        y, x = np.mgrid[BLOCK[1]:BLOCK[3], BLOCK[0]:BLOCK[2]].astype(np.float32)
        pos = np.stack([(x + ox).ravel(), (y + oy).ravel()], axis=-1).astype(np.float32)
        p = sc.new_particles(rows_of_slot[k], pos)
        if conc_of_slot is not None:
            p["s"] = np.float32(conc_of_slot[k])
        out.append(p)
    return np.concatenate(out)


def drift_pass(params, parts, rows):
    """One particle pass with the grid velocities zeroed after grid_prep: (particles before,
    after, gparams, mats, dt)."""
    params.crowd_jitter = 0.0
    parts = cases.settle(parts, params)
    scene = sc.Scene((W, H))
    g, mats = N.derive_grid_params(params, W, H, rows)
    o = OracleSim((W, H), W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(scene.rasterize())
    o.reset(parts)
    o.begin_batch(1 / 240, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    vel = o.field("VEL", np.float32, (H * W, 4)).copy()
    vel[:, :2] = 0.0
    o.put_field("VEL", vel)
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    aux = o.field("AUX", np.float32, (H, W, 4)).copy()
    cell = o.field("CELL", np.float32, (H, W, 4)).copy()
    dt = float(o.field("STEP", N.STEP_DTYPE, 1)[0]["dt"])
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 1)
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
    return before, after, g, mats, dt, aux, cell


def assert_drift(before, after, k, dt, want):
    """Each particle's vertical displacement over the pass is dt * want, to the float32
    rounding of its new position (half an ulp) and 1e-5 of the drift; no horizontal one."""
    y0 = before["pos"][k, 1].astype(np.float64)
    y1 = after["pos"][k, 1].astype(np.float64)
    tol = 0.6 * np.spacing(np.abs(after["pos"][k, 1])).astype(np.float64) + 1e-5 * abs(want) * dt
    assert np.all(np.abs((y1 - y0) - dt * want) <= tol), np.abs((y1 - y0) / dt - want).max()
    assert np.array_equal(after["pos"][k, 0], before["pos"][k, 0])
    return (y1 - y0) / dt


def inner(p):
    x, y = p["pos"][:, 0], p["pos"][:, 1]
    return (x >= INNER[0]) & (x < INNER[2]) & (y >= INNER[1]) & (y < INNER[3])


def terminal_root(drho, rho_c, nu_c, d, g):
    """Positive root of the drag balance a v^2 + b v = c, a = 0.44 rho_c, b = 24 nu_c rho_c / d,
    c = 4/3 |drho| g d (float64; in the form without cancellation when b^2 >> a c), checked
    against the balance itself."""
    a, b, c = 0.44 * rho_c, 24.0 * nu_c * rho_c / d, 4.0 / 3.0 * abs(drho) * g * d
    v = 2.0 * c / (b + np.sqrt(b * b + 4.0 * a * c))
    assert abs(a * v * v + b * v - c) <= 1e-12 * c
    return v


@pytest.mark.parametrize("nu_w", [1e-6, 1e-3])
def test_drift_between_groups_is_the_terminal_velocity(nu_w):
    params = N.default_params()
    params.fixed_hz = 240.0
    params.cell_size = 0.01
    params.mat[N.ROW_WATER].viscosity = nu_w
    rows = [N.ROW_WATER, N.ROW_OIL]
    parts = lattice_particles([N.ROW_OIL, N.ROW_WATER, N.ROW_WATER, N.ROW_WATER])
    before, after, g, mats, dt, aux, cell = drift_pass(params, parts, rows)
    assert np.array_equal(before["id"], after["id"])
    sel = inner(before)
    rw, ro = mats[N.ROW_WATER].rho, mats[N.ROW_OIL].rho
    # the bond hold: this tension and cell size leave the unresolved drift on
    cap = g.cfl / dt if g.corr[2] > 0.0 else np.inf
    nu_c = float(aux[INNER[1]:INNER[3], INNER[0]:INNER[2], 0].mean())
    assert np.allclose(aux[INNER[1]:INNER[3], INNER[0]:INNER[2], 0], nu_c, rtol=1e-6)
    flux = 0.0
    for row, share, drho in ((N.ROW_WATER, 0.75, rw - ro), (N.ROW_OIL, 0.25, ro - rw)):
        k = sel & ((before["kind"] & 255) == row + 1)
        assert k.sum() > 100
        vt = min(terminal_root(drho, rw, nu_c, g.sep_tau, g.gravity), cap)
        want = -vt * np.sign(drho) * (1.0 - share)
        dy = assert_drift(before, after, k, dt, want)
        flux += share * float(dy.mean())
        # water sinks through the lighter oil, oil rises
        assert (dy.mean() < 0) == (row == N.ROW_WATER)
    assert abs(flux) < 1e-2 * 0.75 * abs(want)


def test_drift_within_a_group_cancels_in_stokes_flow():
    params = N.default_params()
    params.fixed_hz = 240.0
    params.cell_size = 0.01
    wax = 3  # a lighter oily liquid in a free row: same phase group as the oil, no tension between them
    m = params.mat[wax]
    o = params.mat[N.ROW_OIL]
    for f, _ in N.Material._fields_:
        setattr(m, f, getattr(o, f))
    m.density = 700.0
    params.mat[N.ROW_OIL].viscosity = 5e-3
    m.viscosity = 5e-3
    rows = [N.ROW_OIL, wax]
    parts = lattice_particles([wax, N.ROW_OIL, wax, N.ROW_OIL])
    before, after, g, mats, dt, aux, cell = drift_pass(params, parts, rows)
    sel = inner(before)
    rmean = 0.5 * (mats[N.ROW_OIL].rho + mats[wax].rho)
    cq = cell[INNER[1]:INNER[3], INNER[0]:INNER[2]]
    nu_c = float((cq[..., 2] / cq[..., 1]).mean())  # the cell's viscosity over its density
    assert np.allclose(cq[..., 2] / cq[..., 1], nu_c, rtol=1e-6)
    cap = g.cfl / dt if g.corr[2] > 0.0 else np.inf
    flux = 0.0
    for row in (N.ROW_OIL, wax):
        k = sel & ((before["kind"] & 255) == row + 1)
        assert k.sum() > 100
        drho = mats[row].rho - rmean
        vt = min(terminal_root(drho, rmean, nu_c, g.sep_tau, g.gravity), cap)
        dy = assert_drift(before, after, k, dt, -vt * np.sign(drho))
        # Stokes regime: the speed is close to |drho| g d^2 / (18 nu rho)
        stokes = abs(drho) * g.gravity * g.sep_tau ** 2 / (18.0 * nu_c * rmean)
        assert vt == pytest.approx(stokes, rel=0.05)
        flux += 0.5 * float(dy.mean())
    assert abs(flux) < 0.05 * vt


def test_family_exchange_relaxes_to_the_local_mean_and_conserves():
    params = N.default_params()
    params.fixed_hz = 240.0
    params.cell_size = 0.005
    honey = N.ROW_HONEY
    rows = [N.ROW_WATER, honey]
    parts = lattice_particles([honey] * 4, conc_of_slot=[0.2, 0.6, 0.6, 0.2])
    before, after, g, mats, dt, aux, cell = drift_pass(params, parts, rows)
    k = mats[honey].mix_rate
    assert k > 0.0
    decay = np.exp(-k * dt)
    s0 = before["s"].astype(np.float64)
    s1 = after["s"].astype(np.float64)
    sel = inner(before)
    # the local mean of every interior neighbourhood is 0.4
    want = s0[sel] + (0.4 - s0[sel]) * (1.0 - decay)
    assert s1[sel] == pytest.approx(want, abs=2e-5)
    assert np.abs(s1[sel] - 0.4) == pytest.approx(np.abs(s0[sel] - 0.4) * decay, abs=2e-5)
    # concentration is conserved over the whole block (boundary particles included)
    assert s1.sum() == pytest.approx(s0.sum(), rel=2e-5)


def test_crowd_jitter_is_a_bounded_uniform_shake():
    """Crowd jitter (particles.comp:896-900, air.glsl:4-11): a particle in a crowd (gathered fill
    above 1.5) is shaken by (u - 1/2) jitter excess per axis, u from an integer hash of its id,
    position and the substep count. Eight particles per cell (fill 2, excess 1/2) with zero
    grid velocities: every displacement lies within +-jitter excess / 2, each axis is uniform
    over that range (mean 0, variance 1/12 of its square, within sampling error), the two axes
    are uncorrelated, the same state and substep give the same shake and the next substep a
    different one."""
    params = N.default_params()
    params.fixed_hz = 240.0
    params.cell_size = 0.01
    rows = [N.ROW_WATER]
    dense = np.concatenate([lattice_particles([N.ROW_WATER] * 4), lattice_particles([N.ROW_WATER] * 4)])
    dense["pos"][len(dense) // 2:] += np.float32(0.125)  # the second lattice shifted inside each cell
    jitter = 0.3

    def shake(substeps_before):
        params.crowd_jitter = jitter
        parts = cases.settle(dense, params)
        scene = sc.Scene((W, H))
        g, mats = N.derive_grid_params(params, W, H, rows)
        o = OracleSim((W, H), W * H * 4)
        o.set_gparams(g, mats)
        o.set_solids(scene.rasterize())
        o.reset(parts)
        st = o.field("STEP", N.STEP_DTYPE, 1).copy()
        st["substeps"] = substeps_before
        o.put_field("STEP", st)
        o.begin_batch(1 / 240, 1)
        o.stage("STEP")
        o.stage("CLEAR_ACC")
        o.stage("PARTICLES", 0)
        o.stage("GRID_PREP")
        vel = o.field("VEL", np.float32, (H * W, 4)).copy()
        vel[:, :2] = 0.0
        o.put_field("VEL", vel)
        comp = o.field("COMP", np.float32, (H, W, 4))
        before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
        o.stage("CLEAR_ACC")
        o.stage("PARTICLES", 1)
        after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()).copy()
        return before, after, comp

    before, after, comp = shake(7)
    fill = comp[INNER[1]:INNER[3], INNER[0]:INNER[2], 3]
    assert np.allclose(fill, 2.0, atol=1e-5)
    amp = jitter * (2.0 - 1.5)
    sel = inner(before)
    d = after["pos"][sel].astype(np.float64) - before["pos"][sel]
    n = len(d)
    assert n > 1500
    ulp = np.spacing(np.float32(64.0))
    assert np.abs(d).max() <= 0.5 * amp + ulp
    u = d / amp
    for axis in (0, 1):
        assert abs(u[:, axis].mean()) < 4.0 / np.sqrt(12.0 * n)
        assert u[:, axis].var() == pytest.approx(1.0 / 12.0, rel=0.08)
        hist, _ = np.histogram(u[:, axis], bins=10, range=(-0.5, 0.5))
        assert hist.min() > 0.7 * n / 10 and hist.max() < 1.3 * n / 10
    assert abs(np.corrcoef(u[:, 0], u[:, 1])[0, 1]) < 4.0 / np.sqrt(n)
    again_before, again, _ = shake(7)
    assert np.array_equal(again["pos"], after["pos"])
    _, other, _ = shake(8)
    moved = other["pos"][sel].astype(np.float64) - before["pos"][sel]
    assert np.mean(np.abs(moved - d) > 1e-4) > 0.9


def test_diluent_joins_its_primary_family_at_the_exchange_rate():
    """Water next to honey (honey's diluent, the tank's primary family): water takes part in the
    exchange as concentration 0 (particles.comp:524-540) and becomes a honey particle once it
    gains any, at s = m (1 - exp(-k dt)) with k honey's mixing rate; honey relaxes towards the
    same mean m (here 0.4: two honey particles at 0.8 and two water per cell); the concentration
    summed over the particles (water as 0) is conserved."""
    params = N.default_params()
    params.fixed_hz = 240.0
    params.cell_size = 0.005
    honey, water = N.ROW_HONEY, N.ROW_WATER
    rows = [water, honey]
    parts = lattice_particles([honey, water, water, honey], conc_of_slot=[0.8, 0.0, 0.0, 0.8])
    before, after, g, mats, dt, aux, cell = drift_pass(params, parts, rows)
    assert g.primary >= 0
    decay = np.exp(-mats[honey].mix_rate * dt)
    sel = inner(before)
    was_water = (before["kind"] & 255) == water + 1
    w = sel & was_water
    h = sel & ~was_water
    assert w.sum() > 100 and h.sum() > 100
    assert np.all((after["kind"][w] & 255) == honey + 1)
    assert after["s"][w].astype(np.float64) == pytest.approx(np.full(w.sum(), 0.4 * (1 - decay)), abs=2e-5)
    assert after["s"][h].astype(np.float64) == pytest.approx(np.full(h.sum(), 0.8 + (0.4 - 0.8) * (1 - decay)), abs=2e-5)

    def total(p):
        return float(np.where((p["kind"] & 255) == honey + 1, p["s"].astype(np.float64), 0.0).sum())

    assert total(after) == pytest.approx(total(before), rel=2e-5)
