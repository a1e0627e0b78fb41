"""The particle pass's launch-shape autotuner (sp_autotune*): every candidate shape (work-items
per group, LDS scatter-box capacity, registers per work-item) gives bit-identical results,
on a one-liquid tank and on a four-liquid tank (a wider accumulator stride, so more blocks
fall back to global atomics), and tuning on the launches themselves settles on a shape for
each key."""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidSim

pytestmark = pytest.mark.gpu


def run(grid, kind, substeps):
    if kind == "four":
        params, field, parts, _ = cases.four_liquids(grid, seed=2, viscosity_tol=1e-5)
    else:
        params, field, parts, _ = cases.dam_break(grid, frac=(0.5, 0.6))
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    try:
        s.set_solids(field)
        s.reset(parts)
        s.record(substeps, substeps / 240)
        s.wait()
        vel, cell = s.download_grid()
        return s.download_particles(), vel, cell, s.download_pressure()[0]
    finally:
        s.close()


@pytest.mark.parametrize("grid,kind", [((96, 54), "dam"), ((96, 54), "four")])
def test_every_particle_shape_is_bit_identical(grid, kind):
    shapes = N.autotune_shapes()
    assert len(shapes) >= 2
    try:
        ref = None
        for k in range(len(shapes)):
            N.autotune(k)
            print(f"shape {k}: {shapes[k]} on {kind} {grid}", flush=True)
            got = run(grid, kind, 6)
            if ref is None:
                ref = got
                continue
            assert got[0].tobytes() == ref[0].tobytes(), f"shape {k}: particles"
            for name, a, b in zip(("vel", "cell", "pres"), got[1:], ref[1:]):
                assert np.array_equal(a, b), f"shape {k}: {name}"
    finally:
        N.autotune(-1)


def test_tuning_settles_on_a_shape():
    N.autotune(-2, 2)
    try:
        grid = (128, 72)
        params, field, parts, _ = cases.dam_break(grid, frac=(0.5, 0.6))
        s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
        try:
            s.set_solids(field)
            s.reset(parts)
            n = len(N.autotune_shapes())
            for _ in range(n):  # 8 timed substeps per wait, 2 timings per shape
                s.record(8, 8 / 240)
                s.wait()
        finally:
            s.close()
        rep = [e for e in N.autotune_report() if (e["grid_w"], e["grid_h"], e["tanks"]) == (128, 72, 1)]
        print(rep, flush=True)
        assert len(rep) == 1
        e = rep[0]
        assert 0 <= e["best"] < n
        assert e["samples"] >= 2 * n
        assert 0 < e["best_ms"] <= e["default_ms"]
    finally:
        N.autotune(-2, 3)
