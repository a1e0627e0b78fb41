"""Parity of settling on the device with the oracle (settle.comp, the family occupancy of
grid_prep, the particle side of particles.comp:332-351): a mud column under clear water.

From identical inputs the family occupancy is bit-exact; the settling factors go through
exp (the diluent's viscosity) and are compared at rtol 1e-5; given the same factors the
particle pass is bit-exact. Whole substeps are compared at the tolerance of the other parity
tests, and the concentrate is conserved on the device as in the oracle.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu


def column(grid=(64, 48)):
    return Pair(grid, case=cases.mud_column, grain=200e-6, water_above=0.2)


def test_settle_stages_match():
    p = column()
    try:
        assert p.g.caps & N.CAP_SETTLE
        p.ora.begin_batch(6 / 240, 6)
        for k in range(6):
            p.ora.substep(0 if k == 0 else 1)
        n = p.ora.n_slots()
        p.sync_to_gpu("PARTICLES", "VEL", "COMP", "AUX", "GRHO", "CELL", "SURF", "SURF2", "PRES", "STEP")
        p.ora.begin_batch(1 / 240, 1)
        p.sync_to_gpu("STEP")
        p.both("STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 0)
        p.sync_to_gpu("PARTICLES", "ACC")
        p.both("GRID_PREP")
        oa, ob = p.f("FAMOCC", np.float32, p.N)
        assert np.count_nonzero(ob) > 100
        assert np.array_equal(oa, ob)
        p.sync_to_gpu("COMP", "CELL", "FAMOCC")
        p.both("SETTLE")
        da, db = p.f("DSETTLE", np.float32, (p.N, 2))
        assert np.count_nonzero(db) > 100
        assert np.allclose(da, db, rtol=1e-5, atol=0.0)
        # the particle pass from the same factors
        p.gpu.put_field("DSETTLE", db)
        p.sync_to_gpu("PARTICLES", "VEL", "STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 1)
        pa, pb = p.f("PARTICLES", N.PARTICLE_DTYPE, n)
        assert np.array_equal(pa["s"], pb["s"]) and np.array_equal(pa["kind"], pb["kind"])
        assert np.array_equal(pa["pos"], pb["pos"])
    finally:
        p.close()


def test_whole_settling_substeps_match():
    p = column()
    try:
        steps = 24
        p0 = p.gpu.download_particles()
        mud0 = (p0["kind"] & 255) == N.ROW_HONEY + 1
        total0 = float(np.sum(p0["s"][mud0], dtype=np.float64))
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
        assert rel_l2(pa["s"], pb["s"]) <= 1e-5
        mud = (pa["kind"] & 255) == N.ROW_HONEY + 1
        assert float(np.sum(pa["s"][mud], dtype=np.float64)) == pytest.approx(total0, rel=2e-4)
        assert np.all((pa["s"] >= 0) & (pa["s"] <= 1))
    finally:
        p.close()
