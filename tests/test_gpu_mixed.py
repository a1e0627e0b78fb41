"""Mixed-precision multigrid (mg_precision 1: BF16 iterates and coarse right-hand sides,
FP16 coarse matrices, FP32 smoothing arithmetic, FP32 CG vectors, FP64 reductions) against
FP64: each pressure solve's solution against the exact FP64 solution of the same system,
its true residual, and the face velocities it projects against those of the exact pressure.

The bar is the FP32 solver's own. Both solvers stop on the recurrence residual in FP64
(r.r <= tol^2 b.b); the true residual ||b - A x|| / ||b|| of an FP32 CG then sits at the
FP32 floor (about 2e-5 on these tanks, above the 1e-6 tolerance), and its error against the
exact solution near what an FP64 CG stopped at the same tolerance reaches. The mixed solve
must converge (both potentials done), keep its true residual within twice the FP32 solve's,
its error within a few times the FP32 solve's, and its projected face velocities within
1e-5 rel-L2 of those of the exact FP64 pressure.
"""
import numpy as np
import pytest

import cases
import fp64_ref
from sprout_amd.fluid_sim import FluidSim

pytestmark = pytest.mark.gpu

TOL = 1e-6


def tank(grid, kind, precision, seed=0):
    if kind == "four":
        params, field, parts, _ = cases.four_liquids(grid, seed=seed, pressure_tol=TOL)
    else:
        params, field, parts, _ = cases.dam_break(grid, frac=(0.5, 0.6))
        params.pressure_tol = TOL
    params.mg_precision = precision
    W, H = grid
    s = FluidSim(grid, W * H * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


@pytest.mark.parametrize("grid,kind", [((96, 54), "dam"), ((192, 108), "four"), ((256, 144), "four")])
def test_mixed_solve_against_fp64(grid, kind):
    out = {}
    for prec in (0, 1):
        s = tank(grid, kind, prec)
        try:
            # a few substeps in, so the warm start and the flow are not trivial
            s.record(6, 6 / 240)
            s.wait()
            out[prec] = fp64_ref.pressure_accuracy(s, TOL, gather=1)
        finally:
            s.close()
    (e32, v32, k32), (e16, v16, k16) = out[0], out[1]
    print(f"\n{kind} {grid}: fp32 {e32} vel {v32:.3g} iters {k32['iters']} | mixed {e16} vel {v16:.3g} iters {k16['iters']}")
    assert k16["done"].tolist() == [1, 1]
    for a, b in zip(e32, e16):
        if a is None:
            assert b is None
            continue
        assert b["true_residual"] <= max(2 * a["true_residual"], 4 * TOL)
        assert b["err_vs_fp64"] <= max(4 * a["err_vs_fp64"], 4 * b["fp64_cg_same_tol_err"], 1e-6)
    assert v32 <= 1e-5
    assert v16 <= 1e-5
    # the narrower storage costs at most a couple of iterations
    assert int(k16["iters"].max()) <= int(k32["iters"].max()) + 2


def test_mixed_substeps_converge():
    """Whole four-liquid substeps with the mixed-precision multigrid: every pressure and
    second-pressure solve converges, and the flow stays close to the FP32 solver's."""
    grid = (128, 72)
    s32, s16 = tank(grid, "four", 0, seed=2), tank(grid, "four", 1, seed=2)
    try:
        for s in (s32, s16):
            for _ in range(3):
                s.record(8, 8 / 240)
                s.wait()
        assert s16.solver.pressure_unconverged == 0
        assert s16.solver.pressure_residual_max <= TOL * 1.0001
        v32, _ = s32.download_grid()
        v16, _ = s16.download_grid()
        d = fp64_ref.live_faces_rel_l2(v16, v32)
        print(f"\nvelocity after 24 substeps, mixed vs fp32: rel-L2 {d:.3g}")
        assert 0 < d <= 1e-3
    finally:
        s32.close()
        s16.close()
