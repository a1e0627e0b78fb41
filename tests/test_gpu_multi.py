"""Parity of the HIP fluid step with the CPU oracle on the four-liquid tank: oil, water,
honey (a miscible family with water) and mercury, three phase groups with pair tensions,
drift between the groups, the family exchange and the implicit viscosity of the viscous
rows (the 1024 x 576 workload's physics at a size the oracle runs in seconds).

The scatter accumulators are integers (bit-exact); the per-cell grid stages have the same
operation order on both sides (bit-exact, except fields through exp(), a few ulp); whole
substeps pass through the FP64-reduced solves (rel-L2 1e-5 on positions).
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, first_substep_prefix, rel_l2

pytestmark = pytest.mark.gpu


def four(grid=(128, 72), seed=0, **kw):
    return Pair(grid, case=cases.four_liquids, seed=seed, viscosity_tol=1e-5, **kw)


def test_four_liquids_tables():
    p = four()
    try:
        g = p.g
        assert g.visc_mode == 2 and g.n_present == 4 and g.n_fam == 1
        assert g.group_mask == 0b111
    finally:
        p.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_four_liquids_scatter_and_grid_bitexact(seed):
    p = four(seed=seed)
    try:
        first_substep_prefix(p)
        a, b = p.f("ACC", np.int32, p.N * p.stride)
        assert np.array_equal(a, b), np.count_nonzero(a != b)
        p.both("GRID_PREP")
        for name in ["VEL", "COMP", "SPROP", "GRHO", "GBASE"]:
            a, b = p.f(name, np.float32, (p.N, 4))
            assert np.array_equal(a, b), name
        # mu (cell.z) and the major group's viscosity (aux.x) pass through exp()
        for name, k in (("CELL", 2), ("AUX", 0)):
            a, b = p.f(name, np.float32, (p.N, 4))
            rest = [j for j in range(4) if j != k]
            assert np.array_equal(a[:, rest], b[:, rest]), name
            np.testing.assert_allclose(a[:, k], b[:, k], rtol=4e-7, atol=0)
        p.sync_to_gpu("CELL", "AUX")
        p.both("COEF")
        p.both("SURF")
        for name in ["COEF", "SURF", "SURF2"]:
            a, b = p.f(name, np.float32, (p.N, 4))
            assert np.array_equal(a, b), name
        a, b = p.f("PJUMP", np.float32, (p.N, 2))
        assert np.array_equal(a, b)
        assert np.abs(a).max() > 0
    finally:
        p.close()


@pytest.mark.parametrize("seed", [0, 1])
def test_four_liquids_whole_substeps_match(seed):
    """Eight substeps: particles by id (positions rel-L2 1e-5, velocities 1e-4, the
    concentration of the honey-water family 1e-4), every solve converged."""
    p = four(seed=seed)
    try:
        steps = 8
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
        assert np.array_equal(pa["kind"], pb["kind"])
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
        assert rel_l2(pa["s"], pb["s"]) <= 1e-4
        st = p.gpu.solver
        assert st.pressure_unconverged == 0 and st.viscosity_unconverged == 0
        assert st.pressure_residual_max <= 1e-6
    finally:
        p.close()
