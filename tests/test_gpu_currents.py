"""Parity of currents and drains on the device with the oracle (particles.comp:578-592,
809-829): a water dam break with a current through the block and a drain in it. From
identical inputs, one particle pass removes the same particles and drives the same ones; the
drive passes through exp, which differs by ulps between the host libm and the device, so
velocities are compared at rtol 1e-6 and positions at 1e-6 rel-L2. Whole substeps are
compared at the tolerance of the other parity tests, with the same particles drained.
"""
import numpy as np
import pytest

from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu


def tank(grid=(96, 54)):
    p = Pair(grid, frac=(0.6, 0.6))
    W, H = grid
    currents = [N.Current.new((0.3 * W, 0.3 * H), 8.0, 90.0, 0.5), N.Current.drain((0.15 * W, 0.2 * H), 5.0, 1.0)]
    p.gpu.set_currents(currents)
    p.g, p.mats = p.gpu.grid_params()
    p.ora.set_gparams(p.g, p.mats)
    return p


def test_currents_in_the_tables():
    p = tank()
    try:
        assert p.g.ext[3] == 2.0 and p.g.caps & N.CAP_CURRENTS
        z = np.array(p.g.current_zone, np.float32)[:2]
        assert z[0, 3] == N.CURRENT_RATE and z[1, 3] == -N.DRAIN_RATE
        assert z[0, 2] == 8.0 and z[1, 2] == 5.0
        p.gpu.set_currents([])
        g, _ = p.gpu.grid_params()
        assert g.ext[3] == 0.0 and not g.caps & N.CAP_CURRENTS
    finally:
        p.close()


def test_particle_pass_with_currents_matches():
    p = tank()
    try:
        p.ora.begin_batch(4 / 240, 4)
        for k in range(4):
            p.ora.substep(0 if k == 0 else 1)
        p.sync_to_gpu("PARTICLES", "VEL", "COMP", "AUX", "GRHO", "CELL", "SURF", "SURF2", "PRES", "STEP")
        p.ora.begin_batch(1 / 240, 1)
        p.sync_to_gpu("STEP")
        p.both("STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 1)
        a = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        b = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        dead_a, dead_b = a["kind"] == N.KIND_DEAD, b["kind"] == N.KIND_DEAD
        assert dead_b.sum() > 5 and np.array_equal(dead_a, dead_b)
        live = ~dead_b
        assert np.allclose(a["vel"][live], b["vel"][live], rtol=1e-6, atol=1e-5)
        assert rel_l2(a["pos"][live], b["pos"][live]) <= 1e-6
    finally:
        p.close()


def test_whole_substeps_with_currents_match():
    p = tank()
    try:
        steps = 24
        p.gpu.record(steps, steps / 240)
        p.gpu.wait()
        p.ora.begin_batch(steps / 240, steps)
        for k in range(steps):
            p.ora.substep(1 if k > 0 else 0)
        pa = p.gpu.download_particles()
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = pb[pb["kind"] != N.KIND_DEAD]
        assert len(pb) < len(p.parts)  # the drain took some
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        assert np.array_equal(pa["id"], pb["id"])
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
    finally:
        p.close()
