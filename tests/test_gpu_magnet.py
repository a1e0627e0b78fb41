"""Parity of magnetism on the device with the oracle (magnet.comp's energy density and face
force, the susceptibility and magnetic resting rule of grid_prep): a ferrofluid pool under
a magnet.

The energy density goes through pow and log, which differ by ulps between the host libm and
the device, so it is compared at rtol 1e-5 (near the series switch, where float32 x -
ln(1 + x) cancels, at the bound the oracle's own pin uses); given the same energy density
the face force is bit-exact. Whole substeps are compared at the tolerance of the other parity
tests, and the pool rises towards the magnet on the device as in the oracle.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu


def pool(grid=(64, 48), **kw):
    built = cases.ferro_pool(grid, **kw)
    p = Pair(grid, case=lambda g, **_: built[:4])
    p.gpu.set_magnets(built[4])
    p.g, p.mats = p.gpu.grid_params()
    p.ora.set_gparams(p.g, p.mats)
    return p


def test_magnet_stages_match():
    p = pool()
    try:
        assert p.g.caps & N.CAP_MAGNET and p.g.mag[0] == 1.0
        p.ora.begin_batch(6 / 240, 6)
        for k in range(6):
            p.ora.substep(0 if k == 0 else 1)
        p.sync_to_gpu("PARTICLES", "VEL", "COMP", "AUX", "GRHO", "CELL", "SURF", "SURF2", "PRES", "STEP", "MAGC")
        p.ora.begin_batch(1 / 240, 1)
        p.sync_to_gpu("STEP")
        p.both("STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 1)
        p.sync_to_gpu("PARTICLES", "ACC")
        p.both("GRID_PREP")
        ca, cb = p.f("CELL", np.float32, (p.N, 4))
        assert np.array_equal(ca[:, 3], cb[:, 3])  # the levels, with the magnetic resting rule
        ma, mb = p.f("MAGC", np.float32, (p.N, 2))
        assert np.array_equal(ma[:, 0], mb[:, 0])  # susceptibility
        p.sync_to_gpu("CELL", "COMP", "VEL", "COEF", "SURF")
        p.both("MAGNET", 0)
        ma, mb = p.f("MAGC", np.float32, (p.N, 2))
        assert np.count_nonzero(mb[:, 1]) > 100
        assert np.allclose(ma[:, 1], mb[:, 1], rtol=4e-4, atol=0.0)
        assert rel_l2(ma[:, 1], mb[:, 1]) <= 1e-5
        p.sync_to_gpu("MAGC", "VEL", "CELL", "COEF")
        p.both("MAGNET", 1)
        va, vb = p.f("VEL", np.float32, (p.N, 4))
        assert np.array_equal(va, vb), f"{np.count_nonzero(va != vb)} velocity words differ"
    finally:
        p.close()


def test_whole_magnetic_substeps_match():
    p = pool(strength=0.5, magnet_gap=5.0)
    try:
        steps = 24
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
