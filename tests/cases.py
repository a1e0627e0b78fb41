"""Seeded inputs of the parity tests and benchmarks.

Dam break (the reference's headless dam-break workload): one block of water, x < 0.4 W,
y < 0.6 H, on a 2 x 2 jittered lattice per cell (fill_rect, maps.rs:468-490), in the solid
tank frame (scene.rs:8, WALL = 3 cells), zero initial velocity, fixed dt 1/240 s.

Four liquids: oil, water, honey (miscible with water) and mercury in four seeded rectangles
tiling a block of 80 % x 80 % of the tank, densities 800 / 1000 / 1300 / 13500 kg/m^3,
kinematic viscosities log-uniform over 1e-6 .. 1e-1 m^2/s and surface tensions uniform over
0.02 .. 0.07 N/m, both drawn from the seed; fixed dt 1/240 s, pressure tolerance 1e-6.
"""
import numpy as np

from sprout_amd import _native as N
from sprout_amd import scene as sc


def dam_params(cell_size=0.005, viscosity=0.0, pressure_tol=1e-6, fixed_hz=240.0):
    p = N.default_params()
    p.cell_size = cell_size
    p.fixed_hz = fixed_hz
    p.pressure_tol = pressure_tol
    p.mat[N.ROW_WATER].viscosity = viscosity
    return p


def settle(parts, params):
    """Particle::settle_spawn (params.rs:70-88) for ids 0..n, isothermal rows."""
    q = parts.copy()
    q["id"] = np.arange(len(q), dtype=np.uint32)
    rows = (q["kind"] & 255).astype(np.int64) - 1
    for row in np.unique(rows):
        m = params.mat[int(row)]
        sel = rows == row
        t = q["t"][sel]
        q["t"][sel] = np.where(t == N.SPAWN_TEMP, np.float32(m.spawn_temp), t)
        sv = q["s"][sel]
        load = np.float32(min(max(m.load, 0.0), 1.0)) if m.diluent >= 0 else np.float32(1.0)
        q["s"][sel] = np.where(sv == N.SPAWN_S, load, sv)
    return q


def dam_break(grid, frac=(0.4, 0.6), **kw):
    """(params, solid field, settled particles, rows) of a dam break on `grid`."""
    params = dam_params(**kw)
    scene, parts = sc.dam_break(grid, frac=frac)
    field = scene.rasterize()
    parts = settle(parts, params)
    return params, field, parts, [N.ROW_WATER]


MIXED_SIZES = ((128, 72), (256, 144), (384, 216), (512, 288))


def mixed_tanks(n=64, seed=0, mg_precision=1, pressure_tol=1e-6):
    """Independent dam-break tanks of uneven sizes: each tank's grid is drawn from
    MIXED_SIZES, its block fills 30-60 % of the tank (area drawn, then the width fraction,
    the height following), its liquid is water or oil (the built-in presets), cell 5 mm,
    dt 1/240 s; pressure multigrid storage `mg_precision` (1: BF16 iterates, FP16 coarse
    matrices). List of (grid, params, solid field, settled particles, rows)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        grid = MIXED_SIZES[int(rng.integers(len(MIXED_SIZES)))]
        area = float(rng.uniform(0.3, 0.6))
        fw = float(rng.uniform(max(area / 0.9, 0.35), 0.9))
        row = (N.ROW_WATER, N.ROW_OIL)[int(rng.integers(2))]
        params = dam_params(pressure_tol=pressure_tol)
        params.mg_precision = mg_precision
        scene, parts = sc.dam_break(grid, frac=(fw, area / fw), row=row)
        field = scene.rasterize()
        out.append((grid, params, field, settle(parts, params), [row]))
    return out


FOUR_ROWS = (N.ROW_OIL, N.ROW_WATER, N.ROW_HONEY, N.ROW_MERCURY)
FOUR_DENSITY = {N.ROW_OIL: 800.0, N.ROW_WATER: 1000.0, N.ROW_HONEY: 1300.0, N.ROW_MERCURY: 13500.0}


def four_liquids(grid, seed=0, fill=(0.8, 0.8), cell_size=0.005, pressure_tol=1e-6, viscosity_tol=1e-3):
    """(params, solid field, settled particles, rows) of the four-liquid tank."""
    rng = np.random.default_rng(seed)
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    params.viscosity_tol = viscosity_tol
    for r in FOUR_ROWS:
        m = params.mat[r]
        m.density = FOUR_DENSITY[r]
        m.viscosity = float(10.0 ** rng.uniform(-6.0, -1.0))
        m.tension = float(rng.uniform(0.02, 0.07))
    W, H = grid
    scene = sc.Scene(grid)
    X, Y = fill[0] * W, fill[1] * H
    xs = float(rng.uniform(0.35, 0.65)) * X
    y0, y1 = (float(rng.uniform(0.35, 0.65)) * Y for _ in range(2))
    rects = [(0.0, 0.0, xs, y0), (0.0, y0, xs, Y), (xs, 0.0, X, y1), (xs, y1, X, Y)]
    order = [FOUR_ROWS[k] for k in rng.permutation(4)]
    parts = []
    n = 0
    for rect, row in zip(rects, order):
        q = sc.fill_rect(scene, rect, row, n_before=n)
        n += len(q)
        parts.append(q)
    parts = settle(np.concatenate(parts), params)
    return params, scene.rasterize(), parts, sorted(FOUR_ROWS)


def floor_strip(scene, x0, x1):
    """A box filling the tank floor between x0 and x1 (cells), flush with its top."""
    w = float(scene.wall)
    return sc.Box(((x0 + x1) / 2.0, w / 2.0), ((x1 - x0) / 2.0, w / 2.0))


def thermal_pool(grid, frac=(1.0, 0.6), temp=None, row=N.ROW_WATER, hot=None, cell_size=0.005, heat_rate=1.0,
                 air_cooling=10.0, evaporation=1, emissivity=None, expansion=None, eddy_heat=0.17, pressure_tol=1e-6):
    """(params, solid field, settled particles, rows) of a pool of one liquid, x < frac[0] W,
    y < frac[1] H, in the tank frame, whose particles hold temperatures: `temp` is a
    function of the particle positions (x, y arrays, cells) giving deg C, or None for the
    ambient 20 deg C. `hot` = (x0, x1, deg C): a strip of the floor held at a temperature."""
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    params.heat_rate = heat_rate
    params.air_cooling = air_cooling
    params.evaporation = evaporation
    params.eddy_heat = eddy_heat
    if emissivity is not None:
        params.mat[row].emissivity = emissivity
    if expansion is not None:
        params.mat[row].expansion = expansion
    W, H = grid
    scene = sc.Scene(grid)
    if hot is not None:
        x0, x1, t = hot
        scene.add(floor_strip(scene, x0, x1), temp=t)
    parts = sc.fill_rect(scene, (0.0, 0.0, frac[0] * W, frac[1] * H), row)
    field = scene.rasterize()
    parts = settle(parts, params)
    if temp is not None:
        parts["t"] = np.asarray(temp(parts["pos"][:, 0], parts["pos"][:, 1]), np.float32)
    return params, field, parts, [row]


def ice_block(grid, block=(0.35, 0.65, 0.35, 0.55), ice_temp=-10.0, water_temp=20.0, pool=0.5, cell_size=0.005,
              heat_rate=1.0, pressure_tol=1e-6):
    """A block of ice floating in water (a thermal tank; water can freeze, so it has rigid
    bodies): water at `water_temp` up to `pool` x H, and a block of frozen water at `ice_temp`
    (fractions of W and H: x0, x1, y0, y1) partly above it. An ice particle carries the
    enthalpy temperature T - L / c (phase_of). (params, solid field, particles, rows)."""
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    params.heat_rate = heat_rate
    W, H = grid
    m = params.mat[N.ROW_WATER]
    scene = sc.Scene(grid)
    x0, x1, y0, y1 = block[0] * W, block[1] * W, block[2] * H, block[3] * H
    water = sc.fill_rect(scene, (0.0, 0.0, float(W), pool * H), N.ROW_WATER)
    inside = (water["pos"][:, 0] >= x0) & (water["pos"][:, 0] < x1) & (water["pos"][:, 1] >= y0)
    water = water[~inside]
    ice = sc.fill_rect(scene, (x0, y0, x1, y1), N.ROW_WATER, len(water))
    water["t"] = np.float32(water_temp)
    ice["t"] = np.float32(ice_temp - m.latent / m.heat_capacity)
    parts = settle(np.concatenate([water, ice]), params)
    return params, scene.rasterize(), parts, [N.ROW_WATER]


def stiff_block(grid, block=(0.3, 0.6, 0.4, 0.7), viscosity=5.0, cell_size=0.005, pressure_tol=1e-6):
    """A block of a liquid more viscous than rigid_viscosity (honey at `viscosity` m^2/s, an
    isothermal tank) falling in an empty tank: every cell it fills is stiff and the block
    moves as one rigid body. (params, solid field, particles, rows)."""
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    params.mat[N.ROW_HONEY].viscosity = viscosity
    W, H = grid
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (block[0] * W, block[2] * H, block[1] * W, block[3] * H), N.ROW_HONEY)
    return params, scene.rasterize(), settle(parts, params), [N.ROW_HONEY]


def rheo_block(grid, rheology=2, yield_stress=300.0, flow_index=1.0, shear_time=1.0, visc_ratio=1.0,
               viscosity=0.05, skin=0.0, friction=0.0, block=(0.35, 0.65, 0.05, 0.55), cell_size=0.005,
               pressure_tol=1e-6):
    """A block of a non-Newtonian liquid (the honey row with a rheology: 1 Carreau, 2
    Herschel-Bulkley, 3 shear thickening, 0 Newtonian) slumping on the floor of an empty tank,
    isothermal. `block` in fractions of W and H. (params, solid field, particles, rows)."""
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    m = params.mat[N.ROW_HONEY]
    m.viscosity = viscosity
    m.rheology = rheology
    m.yield_stress = yield_stress
    m.flow_index = flow_index
    m.shear_time = shear_time
    m.visc_ratio = visc_ratio
    m.skin = skin
    m.friction = friction
    W, H = grid
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (block[0] * W, block[2] * H, block[1] * W, block[3] * H), N.ROW_HONEY)
    return params, scene.rasterize(), settle(parts, params), [N.ROW_HONEY]


def two_fluid_heat(grid=(512, 288), seed=0, cell_size=0.01, strip_cells=10, strip_temp=76.85):
    """The two-fluid tank with heat: water below (density 1000, 1e-6 m^2/s), oil above
    (900, 1e-4), a honey-like blob (1400, 0.09) near the top centre, everything at 20 deg C
    (293 K), on a floor strip of `strip_cells` cells at the centre held at 350 K.
    (params, solid field, particles, rows)."""
    W, H = grid
    strip = (0.5 - 0.5 * strip_cells / W, 0.5 + 0.5 * strip_cells / W)
    p = dam_params(cell_size=cell_size)
    p.mat[N.ROW_WATER].density = 1000.0
    p.mat[N.ROW_WATER].viscosity = 1e-6
    p.mat[N.ROW_OIL].density = 900.0
    p.mat[N.ROW_OIL].viscosity = 1e-4
    p.mat[N.ROW_HONEY].density = 1400.0
    p.mat[N.ROW_HONEY].viscosity = 0.09
    scene = sc.Scene(grid)
    scene.add(floor_strip(scene, W * strip[0], W * strip[1]), temp=strip_temp)
    water = sc.fill_rect(scene, (0.0, 0.0, float(W), 0.35 * H), N.ROW_WATER)
    oil = sc.fill_rect(scene, (0.0, 0.35 * H, float(W), 0.55 * H), N.ROW_OIL, len(water))
    rng = np.random.default_rng(seed)
    cx = W * (0.5 + 0.05 * (rng.random() - 0.5))
    honey = sc.fill_rect(scene, (cx - 0.06 * W, 0.72 * H, cx + 0.06 * W, 0.86 * H), N.ROW_HONEY, len(water) + len(oil))
    parts = np.concatenate([water, oil, honey])
    field = scene.rasterize()
    parts = settle(parts, p)
    return p, field, parts, [N.ROW_WATER, N.ROW_HONEY, N.ROW_OIL]


LAYER_ROWS = (N.ROW_WATER, N.ROW_OIL, N.ROW_HONEY, 3, 4, 5)


def layered_liquids(grid, k, seed=0, cell_size=0.005, fill=(0.6, 0.8), pressure_tol=1e-6, viscosity_tol=1e-4):
    """k (2 .. 6) liquids in horizontal layers of seeded thicknesses, stacked in a seeded
    order (heavy over light where the draw puts it so), filling a seeded 60-80 % of the tank
    height; densities uniform over 900 .. 1400 kg/m^3 and kinematic viscosities log-uniform
    over 1e-6 .. 0.09 m^2/s (the two-fluid tank's ranges), isothermal.
    (params, solid field, particles, rows)."""
    assert 2 <= k <= len(LAYER_ROWS)
    rng = np.random.default_rng(seed)
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    params.viscosity_tol = viscosity_tol
    rows = list(LAYER_ROWS[:k])
    for r in rows:
        m = params.mat[r]
        m.density = float(rng.uniform(900.0, 1400.0))
        m.viscosity = float(10.0 ** rng.uniform(-6.0, np.log10(0.09)))
    W, H = grid
    scene = sc.Scene(grid)
    top = float(rng.uniform(*fill)) * H
    cuts = np.sort(rng.uniform(0.0, 1.0, k - 1))
    edges = np.concatenate([[0.0], cuts, [1.0]]) * top
    parts = []
    n = 0
    for row, y0, y1 in zip([rows[j] for j in rng.permutation(k)], edges[:-1], edges[1:]):
        q = sc.fill_rect(scene, (0.0, float(y0), float(W), float(y1)), row, n_before=n)
        n += len(q)
        parts.append(q)
    parts = settle(np.concatenate(parts), params)
    return params, scene.rasterize(), parts, sorted(rows)


ROW_SLIME = 3  # a built-in row with the default material, made viscoelastic by slime_params


def slime_params(params, row=ROW_SLIME, shear_modulus=100.0, relax_time=1.0, viscosity=1.0e-2, density=1030.0,
                 tension=0.06):
    """Makes `row` the reference's slime (presets.rs:428-445): a Maxwell liquid, G 100 Pa,
    relaxation 1 s, over a 0.01 m^2/s solvent, 1030 kg/m^3 (isothermal fields only)."""
    m = params.mat[row]
    m.shear_modulus = shear_modulus
    m.relax_time = relax_time
    m.viscosity = viscosity
    m.density = density
    m.tension = tension
    return params


def slime_drop(grid, block=(0.35, 0.65, 0.35, 0.75), cell_size=0.005, pressure_tol=1e-6, **mat):
    """A block of slime falling onto the floor of an empty tank (the reference's slime drop),
    isothermal. `block` in fractions of W and H. (params, solid field, particles, rows)."""
    params = slime_params(dam_params(cell_size=cell_size, pressure_tol=pressure_tol), **mat)
    W, H = grid
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (block[0] * W, block[2] * H, block[1] * W, block[3] * H), ROW_SLIME)
    return params, scene.rasterize(), settle(parts, params), [ROW_SLIME]


def shear_layer(grid=(96, 48), y0=16.0, h=16.0, x=(8.0, 88.0), amp=2.0, cell_size=0.01, **mat):
    """The fundamental shear mode of a free slime layer in zero gravity (the reference's
    viscoelastic check, checks.rs:456-503): a layer y0 <= y < y0 + h of the tank, x[0] <= x <
    x[1], whose particles move as u = amp cos(k (y - y0)), k = pi / h, v = 0 (cells / s), with
    the matching APIC gradient du/dy. No surface tension; adaptive substeps (no fixed rate).
    (params, solid field, particles, rows, k)."""
    params = slime_params(dam_params(cell_size=cell_size, fixed_hz=0.0), tension=0.0, **mat)
    params.gravity = 0.0
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (x[0], y0, x[1], y0 + h), ROW_SLIME)
    k = np.pi / h
    yy = parts["pos"][:, 1].astype(np.float64) - y0
    parts["vel"][:, 0] = (amp * np.cos(k * yy)).astype(np.float32)
    parts["vel"][:, 1] = 0.0
    parts["c"][:] = 0.0
    parts["c"][:, 1] = (-amp * k * np.sin(k * yy)).astype(np.float32)
    return params, scene.rasterize(), settle(parts, params), [ROW_SLIME], k


def mud_params(params, grain=100e-6, density=2000.0, load=0.5, viscosity=0.05):
    """Makes the honey row (a concentrate of water) the reference's mud (presets.rs:119-140):
    silt grains of `grain` m at 2000 kg/m^3 suspended in water, emitted at concentration
    `load`; Newtonian here (the preset's Carreau thinning is left out)."""
    m = params.mat[N.ROW_HONEY]
    m.density = density
    m.viscosity = viscosity
    m.grain = grain
    m.load = load
    m.mix_rate = 0.0
    return params


def mud_column(grid, block=(0.3, 0.7, 0.0, 0.6), cell_size=0.005, pressure_tol=1e-6, water_above=0.0, **mat):
    """A column of mud (a suspension of grains in water) standing on the floor of a tank,
    optionally under a layer of clear water `water_above` (fraction of H) deep, isothermal.
    (params, solid field, particles, rows)."""
    params = mud_params(dam_params(cell_size=cell_size, pressure_tol=pressure_tol), **mat)
    W, H = grid
    scene = sc.Scene(grid)
    x0, x1, y0, y1 = block[0] * W, block[1] * W, block[2] * H, block[3] * H
    mud = sc.fill_rect(scene, (x0, y0, x1, y1), N.ROW_HONEY)
    parts = [mud]
    if water_above > 0.0:
        parts.append(sc.fill_rect(scene, (x0, y1, x1, y1 + water_above * H), N.ROW_WATER, len(mud)))
    parts = np.concatenate(parts)
    return params, scene.rasterize(), settle(parts, params), [N.ROW_WATER, N.ROW_HONEY]


ROW_FERRO = 4  # a built-in row with the default material, made magnetisable by ferro_params


def ferro_params(params, row=ROW_FERRO, susceptibility=2.0, saturation=0.05, density=1200.0, viscosity=5e-3):
    """Makes `row` a ferrofluid (magnetisable liquid): volume susceptibility `susceptibility`,
    saturation mu0 Ms `saturation` T."""
    m = params.mat[row]
    m.susceptibility = susceptibility
    m.saturation = saturation
    m.density = density
    m.viscosity = viscosity
    return params


def ferro_pool(grid, depth=0.3, magnet_gap=6.0, strength=0.3, radius=4.0, cell_size=0.005, pressure_tol=1e-6, **mat):
    """A pool of ferrofluid `depth` (fraction of H) deep on the floor of a tank, and a magnet
    `magnet_gap` cells above its surface at the middle, north pole up. (params, solid field,
    particles, rows, magnets)."""
    params = ferro_params(dam_params(cell_size=cell_size, pressure_tol=pressure_tol), **mat)
    W, H = grid
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (0.0, 0.0, float(W), depth * H), ROW_FERRO)
    top = max(depth * H, float(sc.WALL))
    magnets = [N.Magnet.new((0.5 * W, top + magnet_gap), radius, 0.0, strength)]
    return params, scene.rasterize(), settle(parts, params), [ROW_FERRO], magnets


def volatile_pool(grid, volatility=2.0, depth=0.3, cell_size=0.005, pressure_tol=1e-6):
    """A pool of a volatile liquid (the default material's row 5 with `volatility` 1/s) on the
    floor of a tank, isothermal at 20 deg C. (params, solid field, particles, rows)."""
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    params.mat[ROW_VOLATILE].volatility = volatility
    W, H = grid
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (0.0, 0.0, float(W), depth * H), ROW_VOLATILE)
    return params, scene.rasterize(), settle(parts, params), [ROW_VOLATILE]


ROW_VOLATILE = 5  # a built-in row with the default material, made volatile by volatile_pool


ROW_ACID, ROW_LYE, ROW_BRINE = 5, 6, 7  # built-in rows with the default material, made reactive by reacting_mix


def reacting_mix(grid, rate=6.0, into=ROW_BRINE, partner_rate=0.0, block=(0.2, 0.8, 0.0, 0.4), seed=0, cell_size=0.005,
                 pressure_tol=1e-6):
    """Two liquids that react, mixed particle by particle in one block on the floor: each
    particle of the block is acid or lye at random (seed). Acid turns into `into` (-1: it is
    used up) at `rate` per second where lye is present; lye turns into brine at
    `partner_rate`. Isothermal. (params, solid field, particles, rows)."""
    params = dam_params(cell_size=cell_size, pressure_tol=pressure_tol)
    a, b = params.mat[ROW_ACID], params.mat[ROW_LYE]
    a.react_with, a.react_into, a.react_rate = ROW_LYE, into, rate
    b.react_with, b.react_into, b.react_rate = ROW_ACID, ROW_BRINE, partner_rate
    W, H = grid
    scene = sc.Scene(grid)
    parts = sc.fill_rect(scene, (block[0] * W, block[2] * H, block[1] * W, block[3] * H), ROW_ACID)
    lye = np.random.default_rng(seed).random(len(parts)) < 0.5
    parts["kind"][lye] = ROW_LYE + 1
    rows = [ROW_ACID, ROW_LYE]
    return params, scene.rasterize(), settle(parts, params), rows


def boiling_pot(grid, x_range=(0.0, 0.06), row=N.ROW_WATER, steam=1, hot=350.0, fume=None, boil=None, latent_vap=None,
                depth=0.5, seed=0, cell_size=0.005, pressure_tol=1e-6):
    """A pot of `row` on a floor held at `hot` deg C, its particles near their boiling point:
    each at the vapour quality x (the share of its latent heat of vaporisation taken up)
    drawn uniformly from x_range by seed, so that within a few substeps some turn into steam,
    rise, condense, escape as puffs or boil away. `steam` 0: boiled liquid leaves at once.
    `fume`: the liquid's fume (puffs where it boils away); `boil`, `latent_vap`: its boiling
    point (deg C) and latent heat (J/kg) if given. Air cooling and radiation off.
    (params, solid field, particles, rows)."""
    W = grid[0]
    params, field, parts, rows = thermal_pool(grid, frac=(1.0, depth), row=row, hot=(0.3 * W, 0.7 * W, hot),
                                              cell_size=cell_size, air_cooling=0.0, evaporation=0, emissivity=0.0,
                                              pressure_tol=pressure_tol)
    params.steam = steam
    m = params.mat[row]
    if fume is not None:
        m.fume = fume
    if boil is not None:
        m.boil = boil
    if latent_vap is not None:
        m.latent_vap = latent_vap
    x = np.random.default_rng(seed).uniform(x_range[0], x_range[1], len(parts))
    parts["t"] = (m.boil + x * m.latent_vap / m.heat_capacity).astype(np.float32)
    return params, field, parts, rows

