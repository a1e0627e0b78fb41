"""The pressure solve's launch structure at grids with coarse levels between level 0 and the
tail (64 x 36 has none): the V-cycle against the oracle bit for bit, coarse levels two per
launch bit-identical to one per launch (FP32 and mixed storage), and whole solves against the
oracle within rel-L2 1e-5 (their FP64 reductions sum in a different order).
"""
import os

import numpy as np
import pytest

from sprout_amd import _native as N
from test_gpu_parity import Pair, first_substep_prefix, rel_l2

pytestmark = pytest.mark.gpu


def prepared(grid, **opts):
    """A dam break after its first substep's grid passes, the device holding the oracle's
    matrix and tile flags."""
    p = Pair(grid, opts=opts)
    first_substep_prefix(p)
    for s in ["GRID_PREP", "COEF", "SURF", "RHS", "MG_BUILD"]:
        p.ora.stage(s)
    p.sync_to_gpu("STEP", "VEL", "CELL", "COMP", "COEF", "SURF", "PJUMP", "PRHS", "MGA", "MGD", "PRES", "PTILE", "SPROP")
    return p


@pytest.mark.parametrize("grid", [(256, 144), (384, 216), (512, 288), (1024, 576)])
def test_vcycle_bitexact_with_coarse_levels(grid):
    p = prepared(grid)
    try:
        assert p.L.tail >= 2
        rng = np.random.default_rng(11)
        r = np.zeros((p.L.total, 2), np.float32)
        r[: p.N] = rng.standard_normal((p.N, 2)).astype(np.float32)
        mgd = p.ora.field("MGD", np.float32, (p.L.total,))
        r[: p.N][mgd[: p.N] <= 0] = 0
        p.gpu.put_field("MGR", r)
        p.ora.put_field("MGR", r)
        p.both("VCYCLE")
        a, b = p.f("MGZ", np.float32, (p.L.total, 2))
        unk = mgd > 0
        assert np.array_equal(a[unk], b[unk])  # every level's finished iterate
        a, b = p.f("MGX", np.float32, (p.L.total, 2))
        assert np.array_equal(a[unk], b[unk])  # and the pre-smoothed one
    finally:
        p.close()


def solve_on(grid, fuse, precision):
    old = os.environ.get("SPROUT_MG_FUSE")
    os.environ["SPROUT_MG_FUSE"] = "1" if fuse else "0"
    try:
        p = prepared(grid, mg_precision=precision)
    finally:
        if old is None:
            del os.environ["SPROUT_MG_FUSE"]
        else:
            os.environ["SPROUT_MG_FUSE"] = old
    try:
        p.gpu.stage("MG_BUILD")
        p.gpu.stage("PRESSURE")
        out = {n: p.gpu.field(n, np.float32, (p.L.total if n.startswith("MG") else p.N, 2))
               for n in ("PRES", "MGZ", "MGX")}
        out["SOLVE"] = p.gpu.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
        return out
    finally:
        p.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("grid", [(512, 288), (1024, 576)])
def test_coarse_levels_two_per_launch_are_bit_identical(grid, precision):
    a = solve_on(grid, True, precision)
    b = solve_on(grid, False, precision)
    for n in ("PRES", "MGZ", "MGX"):
        assert a[n].tobytes() == b[n].tobytes(), n
    assert a["SOLVE"]["iters"].tolist() == b["SOLVE"]["iters"].tolist()
    assert a["SOLVE"]["rr"].tolist() == b["SOLVE"]["rr"].tolist()


@pytest.mark.parametrize("grid", [(384, 216), (1024, 576)])
def test_pressure_solve_matches_oracle_with_coarse_levels(grid):
    p = prepared(grid)
    try:
        p.both("PRESSURE")
        a, b = p.f("PRES", np.float32, (p.N, 2))
        assert rel_l2(a[:, 0], b[:, 0]) <= 1e-5
        assert rel_l2(a[:, 1], b[:, 1]) <= 1e-5
        ka, kb = p.f("SOLVE", N.SOLVE_DTYPE, 1)
        assert ka[0]["done"].tolist() == [1, 1]
        assert abs(int(ka[0]["iters"].max()) - int(kb[0]["iters"].max())) <= 1
        rel = np.sqrt(ka[0]["rr"][0] / ka[0]["bb"][0])
        assert rel <= 1e-6 * (1 + 1e-9)  # the recurrence residual at the tolerance
    finally:
        p.close()
