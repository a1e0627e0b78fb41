"""Parity of reactions on the device with the oracle (particles.comp:740-770, the reaction
mask of grid_prep): in a mixed block of acid and lye the same particles turn into brine on
the device as in the oracle (the draw depends on the particle's id, the substep count and its
cell's mask, bit-exact on both sides), and a used-up reactant leaves the same particles and
ledger counts.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("into", [cases.ROW_BRINE, -1])
def test_reacting_mix_matches(into):
    p = Pair((64, 48), case=cases.reacting_mix, rate=12.0, into=into)
    try:
        assert p.g.caps & N.CAP_REACT
        steps = 12
        p.gpu.record(steps, steps / 240)
        p.gpu.wait()
        p.ora.begin_batch(steps / 240, steps)
        for k in range(steps):
            p.ora.substep(1 if k > 0 else 0)
        ma, mb = p.f("RMASK", np.uint32, p.N)
        assert np.array_equal(ma, mb)
        pa = p.gpu.download_particles()
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = pb[pb["kind"] != 0]
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        assert np.array_equal(pa["id"], pb["id"])
        assert np.array_equal(pa["kind"], pb["kind"])  # the same particles reacted
        changed = (pb["kind"] & 255) == cases.ROW_BRINE + 1
        assert into < 0 or changed.sum() > 50
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
        boiled, _, _ = p.gpu.download_ledger_counts()
        ob = p.ora.field("LEDGER", np.uint32, N.LEDGER_WORDS)[16:16 + N.MAX_MATERIALS]
        assert np.array_equal(boiled, ob)
    finally:
        p.close()
