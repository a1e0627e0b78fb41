"""Pins of the oracle's rigid-body restatement (rigid.comp, body_motion / body_contact in
common.glsl, the CAP_RIGID branches of grid_prep.comp and particles.comp), on the CPU.

  * the labels are the 8-connected components of the rigid cells (scipy.ndimage.label), each
    rooted at its smallest cell index, and every body has a slot;
  * the body sums are the physical totals: mass, centre of mass and momentum of the body's
    cells from their densities, fill and cell-centred velocities;
  * a body already moving rigidly (translation and rotation) keeps its motion through the
    passes, up to the cell's own inertia (1/6 of its mass), which the closed form includes;
  * a body pressed onto the floor while sliding stops moving into it and slides with Coulomb
    friction (the closed form of body_contact for a flat support), or sticks below the
    friction limit;
  * whole substeps: an ice block floating in warm water, and a block of a liquid stiffer than
    rigid_viscosity falling in air, keep the distances between their particles.
"""
import numpy as np
import pytest
from scipy import ndimage

import cases
from sprout_amd import _native as N
from sprout_amd import scene as sc
from oracle.binding import OracleSim

FIX = N.RIGID_FIX


def make(grid, built, thermal=False):
    params, field, parts, rows = built
    W, H = grid
    g, mats = N.derive_grid_params(params, W, H, rows, thermal=thermal)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return o, g


def empty_tank(grid):
    """An oracle tank with walls and nothing in it, whose rigid passes are driven by the
    fields the test puts in (cell fill and density, stiff flags, face velocities)."""
    built = cases.dam_break(grid, frac=(0.1, 0.1))
    return make(grid, built)[0]


def put_body(o, grid, mask, rho=1.0, phi=1.0, vel=None):
    W, H = grid
    n = W * H
    cell = np.zeros((n, 4), np.float32)
    m = mask.reshape(-1)
    cell[:, 0] = np.where(m, phi, 0.0)
    cell[:, 1] = np.where(m, rho, 0.0)
    cell[:, 3] = np.where(m, 0.5, -0.5)
    o.put_field("CELL", cell)
    rh = np.zeros((n, 4), np.float32)
    rh[:, 1] = -1.0
    rh[:, 2] = 1.0
    rh[:, 3] = m.astype(np.float32)
    o.put_field("RH", rh)
    v = np.zeros((n, 4), np.float32) if vel is None else vel.reshape(n, 4).astype(np.float32)
    o.put_field("VEL", v)


def rigid(o, modes):
    for m in modes:
        o.stage("RIGID", m)


def labels(o, grid):
    W, H = grid
    return o.field("RL", np.int32, (W * H, 2))


def body_sums(o, slot):
    words = o.field("RBODY", np.uint32, N.RIGID_WORDS).astype(np.uint64)
    base = 16 + N.BODY_WORDS * slot
    out = []
    for k in range(16):
        v = int(words[base + 2 * k]) | int(words[base + 2 * k + 1]) << 32
        if v >= 1 << 63:
            v -= 1 << 64
        out.append(v / FIX)
    return out


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_labels_are_the_8_connected_components(seed):
    grid = (40, 30)
    W, H = grid
    rng = np.random.default_rng(seed)
    mask = np.zeros((H, W), bool)
    mask[4:H - 4, 4:W - 4] = rng.random((H - 8, W - 8)) < 0.45
    o = empty_tank(grid)
    put_body(o, grid, mask)
    rigid(o, (0, 1, 2))
    rl = labels(o, grid)
    comp, n = ndimage.label(mask, structure=np.ones((3, 3), int))
    flat = comp.reshape(-1)
    assert np.all(rl[flat == 0, 0] == -1)
    roots = []
    for k in range(1, n + 1):
        idx = np.flatnonzero(flat == k)
        assert np.all(rl[idx, 0] == idx.min()), f"component {k}"
        roots.append(idx.min())
    roots = np.sort(roots)
    # every body has a slot, claimed at its root; other cells carry none
    assert sorted(rl[roots, 1]) == list(range(n))
    others = np.setdiff1d(np.arange(W * H), roots)
    assert np.all(rl[others, 1] == -1)
    assert o.field("RBODY", np.uint32, N.RIGID_WORDS)[0] == n


def rigid_field(grid, U, V, w, xc, yc):
    """Face velocities of a rigid motion (U, V) + w x (x - xc): u at (i, j + 0.5), v at
    (i + 0.5, j); correction velocities 0 (live faces)."""
    W, H = grid
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    vel = np.zeros((H, W, 4), np.float64)
    vel[..., 0] = U - w * (y + 0.5 - yc)
    vel[..., 1] = V + w * (x + 0.5 - xc)
    return vel


def test_body_sums_are_mass_centre_and_momentum():
    grid = (48, 40)
    W, H = grid
    rng = np.random.default_rng(5)
    mask = np.zeros((H, W), bool)
    mask[12:27, 10:31] = True
    mask[27:30, 18:23] = True  # a knob: not a rectangle
    vel = np.zeros((H, W, 4), np.float32)
    vel[..., :2] = rng.normal(0.0, 3.0, (H, W, 2))
    rho = np.float32(1.3)
    o = empty_tank(grid)
    put_body(o, grid, mask, rho=rho, phi=1.0, vel=vel)
    rigid(o, (0, 1, 2, 3))
    rl = labels(o, grid)
    root = int(np.flatnonzero(mask.reshape(-1)).min())
    slot = int(rl[root, 1])
    s = body_sums(o, slot)
    ys, xs = np.nonzero(mask)
    m = float(rho)
    M = m * len(xs)
    rx, ry = root % W, root // W
    uc = 0.5 * (vel[ys, xs, 0].astype(np.float64) + vel[ys, xs + 1, 0])
    vc = 0.5 * (vel[ys, xs, 1].astype(np.float64) + vel[ys + 1, xs, 1])
    # each cell's mass is rounded to 1 / FIX
    assert s[0] == pytest.approx(M, abs=len(xs) * 0.5 / FIX)
    assert s[1] / s[0] == pytest.approx(np.mean(xs - rx), abs=1e-4)
    assert s[2] / s[0] == pytest.approx(np.mean(ys - ry), abs=1e-4)
    assert s[3] / s[0] == pytest.approx(np.mean(uc), abs=1e-4)
    assert s[4] / s[0] == pytest.approx(np.mean(vc), abs=1e-4)
    # no solid within a cell of the body: no contacts
    assert s[7] == 0.0


@pytest.mark.parametrize("U,V,w", [(1.5, -0.7, 0.0), (0.0, 0.0, 0.3), (0.8, 0.4, -0.25)])
def test_rigid_motion_is_kept(U, V, w):
    grid = (64, 48)
    W, H = grid
    mask = np.zeros((H, W), bool)
    mask[14:34, 18:46] = True
    ys, xs = np.nonzero(mask)
    xc, yc = xs.mean() + 0.5, ys.mean() + 0.5
    vel = rigid_field(grid, U, V, w, xc, yc)
    o = empty_tank(grid)
    put_body(o, grid, mask, rho=1.0, phi=1.0, vel=vel)
    rigid(o, range(5))
    got = o.field("VEL", np.float32, (H, W, 4)).astype(np.float64)
    # the cells' own inertia (M / 6) takes part in the angular momentum but not in the sampled
    # motion: the body turns at w * S2 / (S2 + M / 6)
    S2 = float(np.sum((xs + 0.5 - xc) ** 2 + (ys + 0.5 - yc) ** 2))
    w1 = w * S2 / (S2 + len(xs) / 6.0)
    want = rigid_field(grid, U, V, w1, xc, yc)
    inner_u = mask & np.roll(mask, 1, axis=1)  # faces between two body cells
    inner_v = mask & np.roll(mask, 1, axis=0)
    assert np.allclose(got[..., 0][inner_u], want[..., 0][inner_u], atol=2e-4)
    assert np.allclose(got[..., 1][inner_v], want[..., 1][inner_v], atol=2e-4)
    # outside the body nothing changes
    assert np.array_equal(got[..., 0][~inner_u], vel[..., 0][~inner_u].astype(np.float32))


@pytest.mark.parametrize("ux,friction_stops", [(2.0, False), (0.3, True)])
def test_body_on_the_floor_stops_and_slides_with_coulomb_friction(ux, friction_stops):
    grid = (64, 40)
    W, H = grid
    o = empty_tank(grid)
    floor = int(sc.WALL)  # the solid border's thickness (cells): the floor's surface at y = floor
    mask = np.zeros((H, W), bool)
    mask[floor:floor + 6, 20:40] = True
    vy = 1.0
    vel = np.zeros((H, W, 4), np.float32)
    vel[..., 0] = ux
    vel[..., 1] = -vy
    put_body(o, grid, mask, rho=1.0, phi=1.0, vel=vel)
    rigid(o, range(5))
    got = o.field("VEL", np.float32, (H, W, 4))
    mu = 0.6
    slide = np.sign(ux) * max(abs(ux) - mu * vy, 0.0)
    inner_u = mask & np.roll(mask, 1, axis=1)
    inner_v = mask & np.roll(mask, 1, axis=0)
    assert np.allclose(got[..., 0][inner_u], slide, atol=1e-4)
    assert np.allclose(got[..., 1][inner_v], 0.0, atol=1e-4)
    assert (slide == 0.0) == friction_stops


def pairwise(p):
    d = p[:, None, :] - p[None, :, :]
    return np.sqrt((d ** 2).sum(-1))


def test_ice_block_floats_rigidly():
    grid = (64, 48)
    built = cases.ice_block(grid)
    o, g = make(grid, built, thermal=True)
    assert g.caps & 4  # rigid bodies on: water can freeze
    parts0 = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    ice = parts0["t"] < -50.0
    o.begin_batch(24 / 240.0, 24)
    for k in range(24):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    by_id = np.argsort(p["id"])
    a = parts0["pos"][ice].astype(np.float64)
    b = p["pos"][by_id][ice].astype(np.float64)
    assert np.all(np.isfinite(b))
    moved = np.abs(b - a).max()
    assert moved > 1e-3  # it moved (it is lighter than the water)
    # distances between the block's particles are kept
    sel = np.random.default_rng(0).choice(len(a), min(300, len(a)), replace=False)
    da, db = pairwise(a[sel]), pairwise(b[sel])
    assert np.abs(da - db).max() < 2e-3 * max(1.0, moved)
    rl = o.field("RL", np.int32, (grid[0] * grid[1], 2))
    assert (rl[:, 1] >= 0).sum() >= 1  # at least one body


def test_stiff_block_falls_rigidly():
    grid = (64, 48)
    built = cases.stiff_block(grid)
    o, g = make(grid, built)
    assert g.caps & 4
    parts0 = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    o.begin_batch(24 / 240.0, 24)
    for k in range(24):
        o.substep(0 if k == 0 else 1)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    by_id = np.argsort(p["id"])
    a = parts0["pos"].astype(np.float64)
    b = p["pos"][by_id].astype(np.float64)
    drop = (a[:, 1] - b[:, 1]).mean()
    assert drop > 0.5  # it fell
    rh = o.field("RH", np.float32, (grid[0] * grid[1], 4))
    assert rh[:, 3].sum() > 10
    sel = np.random.default_rng(1).choice(len(a), min(300, len(a)), replace=False)
    da, db = pairwise(a[sel]), pairwise(b[sel])
    assert np.abs(da - db).max() < 2e-3 * max(1.0, drop)
