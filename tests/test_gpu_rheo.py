"""Parity of the rheology on the device with the oracle (rheo.comp, the non-Newtonian share
of grid_prep.comp, the plug and jammed cells of rigid.comp, the yield-stress particles held
on a plug): a block of each non-Newtonian model slumping on the floor.

From identical inputs, grid_prep's non-Newtonian share (rh) is bit-exact. The rheology pass
goes through pow and exp, which differ by a few ulp between the host libm and the device, so
the effective viscosity is compared at rtol 1e-5 and its stiff flags exactly. Whole substeps
are compared at the tolerance of the other parity tests.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu

MODELS = {
    "carreau": dict(rheology=1, flow_index=0.4, shear_time=0.5, visc_ratio=0.05, yield_stress=0.0, viscosity=0.05),
    "bingham": dict(rheology=2, flow_index=1.0, yield_stress=300.0, viscosity=0.02),
    "herschel_bulkley": dict(rheology=2, flow_index=0.6, yield_stress=150.0, viscosity=0.05),
    "thickening": dict(rheology=3, flow_index=2.5, shear_time=0.3, visc_ratio=30.0, yield_stress=0.0, viscosity=0.01),
    "skin": dict(rheology=0, skin=0.3, yield_stress=0.0, viscosity=0.01),
}


def block(grid, kind):
    return Pair(grid, case=cases.rheo_block, block=(0.35, 0.65, 0.07, 0.6), **MODELS[kind])


@pytest.mark.parametrize("kind", list(MODELS))
def test_rheo_pass_matches(kind):
    grid = (64, 48)
    p = block(grid, kind)
    try:
        assert p.g.caps & 8
        # a few substeps of the oracle first, so that the block is sheared as it slumps
        p.ora.begin_batch(6 / 240, 6)
        for k in range(6):
            p.ora.substep(0 if k == 0 else 1)
        p.sync_to_gpu("PARTICLES", "VEL", "COMP", "AUX", "GRHO", "CELL", "SURF", "SURF2", "RL", "RBODY", "RH",
                      "PRES", "STEP")
        p.ora.begin_batch(1 / 240, 1)
        p.sync_to_gpu("STEP")
        p.both("STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 1)
        pa = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-6
        p.sync_to_gpu("PARTICLES", "ACC")
        p.both("GRID_PREP")
        a, b = p.f("RH", np.float32, (p.N, 4))
        assert np.array_equal(a, b), f"{np.count_nonzero(a != b)} rh words differ"
        assert (b[:, 0] > 0).sum() > 50
        p.sync_to_gpu("CELL", "VEL")
        p.both("RHEO")
        a, b = p.f("CELL", np.float32, (p.N, 4))
        assert np.allclose(a[:, 2], b[:, 2], rtol=1e-5, atol=0.0)
        ra, rb = p.f("RH", np.float32, (p.N, 4))
        assert np.array_equal(ra, rb)
    finally:
        p.close()


@pytest.mark.parametrize("kind", list(MODELS))
def test_whole_rheo_substeps_match(kind):
    grid = (64, 48)
    p = block(grid, kind)
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
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
    finally:
        p.close()
