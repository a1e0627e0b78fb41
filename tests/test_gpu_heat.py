"""Parity of the heat path on the device with the oracle: the thermal scatter (the heat sums
and each cell's particle sums), the thermal grid_prep (temperatures, heat capacities,
conductivities, the temperature law and expansion), the heat pass (heat.comp) with its
ledger, the particles' heat gather and exchange, and whole substeps of thermal tanks.

Integer work (accumulators, cell sums, ledger words) is bit-exact, and so is every float
that no transcendental function touches. What passes through exp differs by a few ulp
between the host libm and the device: the cell viscosity, the exchange share inside a cell
(1 - exp(-k rate)), the evaporation (saturation pressure) and the contact exchange
(1 - exp(-4 x)); those are compared at rtol 1e-5, and the ledger, which adds them up in
quanta of 1/65536, within one quantum per contributing cell.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu


def pool(grid, **kw):
    kw.setdefault("temp", lambda x, y: 20.0 + 40.0 * (x > grid[0] / 2) + 0.1 * y)
    kw.setdefault("hot", (grid[0] * 0.3, grid[0] * 0.7, 76.85))
    kw.setdefault("heat_rate", 20.0)
    return Pair(grid, case=cases.thermal_pool, **kw)


def prefix(p, gather=0):
    p.ora.begin_batch(1 / 240, 1)
    p.sync_to_gpu("STEP")
    p.both("STEP")
    p.both("CLEAR_ACC")
    p.both("PARTICLES", gather)


def ledger_words(p):
    a = p.gpu.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16].astype(np.uint64)
    b = p.ora.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16].astype(np.uint64)
    la = [int(a[2 * e]) | int(a[2 * e + 1]) << 32 for e in range(8)]
    lb = [int(b[2 * e]) | int(b[2 * e + 1]) << 32 for e in range(8)]
    return la, lb


@pytest.mark.parametrize("grid", [(48, 32), (96, 54)])
def test_thermal_scatter_and_grid_prep_bitexact(grid):
    p = pool(grid)
    try:
        assert p.g.caps & 2 and p.g.chan_t >= 5
        prefix(p)
        a, b = p.f("ACC", np.int32, p.N * p.stride)
        assert np.array_equal(a, b), f"{np.count_nonzero(a != b)} accumulator words differ"
        assert a.reshape(-1, p.stride)[:, p.g.chan_t].sum() > 0
        a, b = p.f("HCELL", np.int32, 2 * p.N)
        assert np.array_equal(a, b) and a.sum() > 0
        p.both("GRID_PREP")
        a, b = p.f("THERM", np.float32, (p.N, 4))
        assert np.array_equal(a, b)
        a, b = p.f("DHEAT", np.float32, (p.N, 4))
        assert np.array_equal(a, b)
        a, b = p.f("CELL", np.float32, (p.N, 4))
        assert np.array_equal(a[:, [0, 1, 3]], b[:, [0, 1, 3]])
        np.testing.assert_allclose(a[:, 2], b[:, 2], rtol=4e-7, atol=0)
    finally:
        p.close()


@pytest.mark.parametrize("evaporation", [0, 1])
def test_heat_pass_matches(evaporation):
    grid = (64, 40)
    p = pool(grid, evaporation=evaporation)
    try:
        prefix(p)
        p.both("GRID_PREP")
        p.both("SURF")
        p.sync_to_gpu("CELL", "THERM", "DHEAT", "SPROP", "VEL")
        p.both("HEAT")
        a, b = p.f("DHEAT", np.float32, (p.N, 4))
        assert np.array_equal(a[:, 2], b[:, 2])
        np.testing.assert_allclose(a[:, 1], b[:, 1], rtol=1e-5, atol=1e-12)
        scale = float(np.max(np.abs(b[:, 0])))
        assert scale > 0
        np.testing.assert_allclose(a[:, 0], b[:, 0], rtol=1e-5, atol=1e-6 * scale)
        if not evaporation:
            assert np.array_equal(a[:, 0], b[:, 0])
        la, lb = ledger_words(p)
        cells = int(np.count_nonzero(b[:, 0]))
        for e in range(4):
            assert abs(la[e] - lb[e]) <= (cells if evaporation else 0), (e, la[e], lb[e])
        assert lb[1] > 0 and lb[2] > 0
    finally:
        p.close()


def test_heat_gather_matches():
    """The particles take the heat pass's result and exchange inside their cells."""
    grid = (64, 40)
    p = pool(grid)
    try:
        prefix(p)
        for s in ("GRID_PREP", "SURF", "HEAT"):
            p.both(s)
        p.sync_to_gpu("DHEAT", "PARTICLES", "VEL", "COMP", "AUX", "GRHO", "CELL", "SURF", "SURF2", "STEP")
        p.ora.begin_batch(1 / 240, 1)
        p.sync_to_gpu("STEP")
        p.both("STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 1)
        pa = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        assert rel_l2(pa["t"] + 273.15, pb["t"] + 273.15) <= 1e-7
        assert np.max(np.abs(pa["t"] - pb["t"])) < 1e-3
    finally:
        p.close()


@pytest.mark.parametrize("grid,case", [((64, 40), "pool"), ((128, 72), "two_fluid")])
def test_whole_thermal_substeps_match(grid, case):
    if case == "pool":
        p = pool(grid)
    else:
        p = Pair(grid, case=cases.two_fluid_heat, cell_size=0.01)
    try:
        steps = 12
        p.gpu.record(steps, steps / 240)
        p.gpu.wait()
        p.ora.begin_batch(steps / 240, steps)
        for k in range(steps):
            p.ora.substep(1 if k > 0 else 0)
        pa = p.gpu.download_particles()
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        assert np.array_equal(pa["id"], pb["id"])
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["t"] + 273.15, pb["t"] + 273.15) <= 1e-6
        assert np.max(np.abs(pa["t"] - pb["t"])) < 0.05
        ea = np.zeros(8)
        assert p.gpu._lib.sp_download_ledger(p.gpu._h, ea.ctypes.data) == 0
        la, lb = ledger_words(p)
        assert ea[2] > 0.0
        for e in range(4):
            assert abs(la[e] - lb[e]) <= 1e-4 * max(lb[e], 65536), (e, la[e], lb[e])
    finally:
        p.close()


def test_floor_above_the_boiling_point_is_carried_with_steam():
    """A floor held above the boiling point (once refused): the tank takes steam into its
    rows and records like any other."""
    from sprout_amd.fluid_sim import FluidSim
    grid = (48, 32)
    params, field, parts, _ = cases.thermal_pool(grid, hot=(10.0, 30.0, 120.0))
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    try:
        s.set_solids(field)
        s.reset(parts)
        s.record(4, 4 / 240)
        s.wait()
        g, _ = s.grid_params()
        assert g.caps & N.CAP_STEAM and N.ROW_STEAM in list(g.present[:g.n_present])
        assert s.solver.substeps == 4
    finally:
        s.close()
