"""Batches of tanks (sp_batch_*): each tank's results in a batch are bit-identical to the
same tank stepped alone. Tanks of the same grid and solver schedule share every launch of
a substep; a converged tank's blocks return at entry while the rest of its group iterates,
so the extra recorded iterations of a group change nothing, and the FP64 partial sums of a
tank are formed in the same order whoever shares its launches.
"""
import numpy as np
import pytest

import cases
from sprout_amd.fluid_sim import FluidBatch, FluidSim

pytestmark = pytest.mark.gpu


def tank(grid, kind, seed=0, frac=(0.4, 0.6)):
    if kind == "four":
        params, field, parts, _ = cases.four_liquids(grid, seed=seed, viscosity_tol=1e-5)
    else:
        params, field, parts, _ = cases.dam_break(grid, frac=frac)
    W, H = grid
    s = FluidSim(grid, W * H * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


def state(s):
    p = s.download_particles()
    vel, cell = s.download_grid()
    pres, _ = s.download_pressure()
    return p, vel, cell, pres


SPECS = [((64, 36), "dam", 0, (0.4, 0.6)), ((64, 36), "four", 1, None), ((96, 54), "dam", 0, (0.5, 0.5)),
         ((64, 36), "dam", 0, (0.6, 0.3)), ((96, 54), "four", 2, None), ((64, 36), "four", 3, None)]


def build(spec):
    grid, kind, seed, frac = spec
    return tank(grid, kind, seed, frac if frac else (0.4, 0.6))


def test_batch_matches_each_tank_alone():
    alone = [build(sp) for sp in SPECS]
    inbatch = [build(sp) for sp in SPECS]
    try:
        steps = 20  # two batches of 10: the iteration adaptation between them included
        for s in alone:
            for _ in range(2):
                s.record(steps // 2, steps / 2 / 240)
                s.wait()
        b = FluidBatch(inbatch)
        for _ in range(2):
            b.record(steps // 2, steps / 2 / 240)
            b.wait()
        # 64x36 water (2 tanks), 64x36 four liquids (2), 96x54 water, 96x54 four liquids
        assert sorted(b.groups()) == [1, 1, 2, 2]
        for k, (x, y) in enumerate(zip(alone, inbatch)):
            sa, sb = state(x), state(y)
            assert sa[0].tobytes() == sb[0].tobytes(), f"tank {k}: particles"
            for name, a, c in zip(("vel", "cell", "pres"), sa[1:], sb[1:]):
                assert np.array_equal(a, c), f"tank {k}: {name}"
            assert x.solver.substeps == y.solver.substeps == steps
            assert x.sim_time() == y.sim_time()
        b.close()
    finally:
        for s in alone + inbatch:
            s.close()


def test_batch_splits_groups_beyond_one_launch():
    """34 tanks of one size: two launches per kernel (32 + 2 tanks); every tank as alone."""
    grid = (32, 24)
    fracs = [(0.3 + 0.01 * k, 0.5) for k in range(34)]
    inbatch = [tank(grid, "dam", frac=f) for f in fracs]
    alone = [tank(grid, "dam", frac=fracs[k]) for k in (0, 33)]
    try:
        b = FluidBatch(inbatch)
        b.record(6, 6 / 240)
        b.wait()
        assert b.groups() == [32, 2]
        for s in alone:
            s.record(6, 6 / 240)
            s.wait()
        for s, k in zip(alone, (0, 33)):
            assert s.download_particles().tobytes() == inbatch[k].download_particles().tobytes()
        b.close()
    finally:
        for s in alone + inbatch:
            s.close()


def test_batch_tank_stops_at_its_own_recorded_iterations():
    """A tank whose solves run out of recorded iterations (pressure capped at 3, viscosity at
    2) shares its launches with tanks that record many more; it still stops where it stops
    alone, unconverged, and its state and statistics are those of the tank alone."""
    grid = (64, 36)

    def make(seed, capped):
        params, field, parts, _ = cases.four_liquids(grid, seed=seed, viscosity_tol=1e-5)
        if capped:
            params.pressure_max_iters = 3
            params.viscosity_max_iters = 2
        s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
        s.set_solids(field)
        s.reset(parts)
        return s

    alone = [make(1, True), make(1, False)]
    inbatch = [make(1, True), make(1, False), make(2, False)]
    try:
        for s in alone:
            for _ in range(2):
                s.record(5, 5 / 240)
                s.wait()
        b = FluidBatch(inbatch)
        for _ in range(2):
            b.record(5, 5 / 240)
            b.wait()
        assert b.groups() == [3]
        assert alone[0].solver.pressure_unconverged > 0
        assert alone[0].solver.viscosity_unconverged > 0
        for k in range(2):
            sa, sb = state(alone[k]), state(inbatch[k])
            assert sa[0].tobytes() == sb[0].tobytes(), f"tank {k}: particles"
            for name, a, c in zip(("vel", "cell", "pres"), sa[1:], sb[1:]):
                assert np.array_equal(a, c), f"tank {k}: {name}"
            x, y = alone[k].solver, inbatch[k].solver
            for f in ("pressure_unconverged", "viscosity_unconverged", "pressure_iters_max",
                      "viscosity_iters_max", "pressure_residual_max", "viscosity_residual_max"):
                assert getattr(x, f) == getattr(y, f), f"tank {k}: {f}"
        b.close()
    finally:
        for s in alone + inbatch:
            s.close()
