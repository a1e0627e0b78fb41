"""Parity of evaporation on the device with the oracle (particles.comp:771-807): a volatile
pool loses the same particles on the device as in the oracle (the draw depends only on the
particle's id, the substep count and its cell, which are the same on both sides), the
survivors agree at the tolerance of the other parity tests, and the ledger counts agree.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu


def test_evaporating_pool_matches():
    p = Pair((64, 48), case=cases.volatile_pool, volatility=6.0)
    try:
        assert p.g.caps & N.CAP_VOLATILE
        steps = 24
        p.gpu.record(steps, steps / 240)
        p.gpu.wait()
        p.ora.begin_batch(steps / 240, steps)
        for k in range(steps):
            p.ora.substep(1 if k > 0 else 0)
        pa = p.gpu.download_particles()
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = pb[pb["kind"] != 0]
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        gone = len(p.parts) - len(pb)
        assert gone > 10
        assert np.array_equal(pa["id"], pb["id"])  # the same particles left
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
        boiled, _, _ = p.gpu.download_ledger_counts()
        ob = p.ora.field("LEDGER", np.uint32, N.LEDGER_WORDS)[16:16 + N.MAX_MATERIALS]
        assert np.array_equal(boiled, ob) and boiled[cases.ROW_VOLATILE] == gone
    finally:
        p.close()
