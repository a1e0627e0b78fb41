"""Tank eviction (sp_evict): an evicted tank comes back on its next call and continues bit
for bit as if it had stayed on the device."""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidSim

pytestmark = pytest.mark.gpu

SPECS = [((64, 36), "dam", 0, (0.4, 0.6)), ((64, 36), "four", 1, None), ((96, 54), "dam", 0, (0.5, 0.5)),
         ((64, 36), "dam", 0, (0.6, 0.3)), ((96, 54), "four", 2, None), ((64, 36), "four", 3, None)]


def build(spec):
    grid, kind, seed, frac = spec
    if kind == "four":
        params, field, parts, _ = cases.four_liquids(grid, seed=seed, viscosity_tol=1e-5)
    else:
        params, field, parts, _ = cases.dam_break(grid, frac=frac)
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


def state(s):
    p = s.download_particles()
    vel, cell = s.download_grid()
    pres, _ = s.download_pressure()
    return p, vel, cell, pres


def assert_same(a, b, k):
    sa, sb = state(a), state(b)
    assert sa[0].tobytes() == sb[0].tobytes(), f"scene {k}: particles"
    for name, x, y in zip(("vel", "cell", "pres"), sa[1:], sb[1:]):
        assert np.array_equal(x, y), f"scene {k}: {name}"
    assert a.sim_time() == b.sim_time()
    for f in ("substeps", "pressure_iters_max", "viscosity_iters_max", "pressure_unconverged",
              "viscosity_unconverged", "pressure_residual_max", "viscosity_residual_max"):
        assert getattr(a.solver, f) == getattr(b.solver, f), f"scene {k}: {f}"


def test_evicted_tank_continues_bit_for_bit():
    a, b = build(SPECS[1]), build(SPECS[1])
    try:
        for s in (a, b):
            s.record(4, 4 / 240)
            s.wait()
        a.evict()
        assert not a.resident
        a.evict()  # evicting twice is harmless
        for s in (a, b):
            s.record(4, 4 / 240)
            s.wait()
        assert a.resident
        assert_same(a, b, 0)
    finally:
        a.close()
        b.close()
