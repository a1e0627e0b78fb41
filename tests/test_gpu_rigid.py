"""Parity of the rigid bodies on the device with the oracle (rigid.comp, the stiff cells of
grid_prep.comp, the frozen and held particles of particles.comp): an ice block floating in
warm water (frozen cells, a thermal tank) and a block of a liquid stiffer than
rigid_viscosity falling in air (stiff cells).

From identical inputs the five passes are bit-exact: the labels (every cell's root), the
bodies' 64-bit sums (integer, compared root by root: slots are claimed in the order roots
arrive, which the sums do not depend on) and the face velocities they apply. The held
particles' gather, and whole substeps, are compared at the tolerances of the other parity
tests.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from test_gpu_parity import Pair, rel_l2

pytestmark = pytest.mark.gpu

CASES = {"ice": cases.ice_block, "stiff": cases.stiff_block}


def up_to_projection(p):
    """One substep of the oracle up to its projection, from the reset state; the device takes
    the oracle's fields, so that the rigid passes start from identical inputs."""
    p.ora.begin_batch(1 / 240, 1)
    p.sync_to_gpu("STEP")
    p.both("STEP")
    p.both("CLEAR_ACC")
    p.both("PARTICLES", 0)
    p.both("GRID_PREP")
    a, b = p.f("RH", np.float32, (p.N, 4))
    assert np.array_equal(a, b), "stiff flags"
    for st in ("COEF", "SURF", "VISC", "RHS", "PRESSURE", "PROJECT"):
        p.ora.stage(st)
    assert np.all(np.isfinite(p.ora.field("VEL", np.float32, (p.N, 4))))
    p.sync_to_gpu("VEL", "CELL", "RH", "THERM")


def sums_by_root(rl, words):
    out = {}
    roots = np.flatnonzero(np.arange(len(rl)) == rl[:, 0])
    for r in roots:
        b = rl[r, 1]
        if b < 0:
            continue
        base = 16 + N.BODY_WORDS * b
        out[int(r)] = words[base:base + N.BODY_WORDS].copy()
    return out


@pytest.mark.parametrize("kind", ["ice", "stiff"])
def test_rigid_passes_bitexact(kind):
    grid = (64, 48)
    p = Pair(grid, case=CASES[kind])
    try:
        assert p.g.caps & 4
        up_to_projection(p)
        for mode in (0, 1, 2):
            p.both("RIGID", mode)
        a, b = p.f("RL", np.int32, (p.N, 2))
        assert np.array_equal(a[:, 0], b[:, 0]), f"{np.count_nonzero(a[:, 0] != b[:, 0])} labels differ"
        assert (b[:, 0] >= 0).sum() > 20
        assert np.array_equal(a[:, 1] >= 0, b[:, 1] >= 0)
        p.both("RIGID", 3)
        wa, wb = p.f("RBODY", np.uint32, N.RIGID_WORDS)
        assert wa[0] == wb[0] > 0
        sa, sb = sums_by_root(a, wa), sums_by_root(b, wb)
        assert sa.keys() == sb.keys()
        for r in sb:
            assert np.array_equal(sa[r], sb[r]), f"body at root {r}: sums differ"
        # the device's slots are its own: hand it the oracle's table for the last pass
        p.sync_to_gpu("RL", "RBODY")
        p.both("RIGID", 4)
        va, vb = p.f("VEL", np.float32, (p.N, 4))
        assert np.array_equal(va, vb), f"{np.count_nonzero(va != vb)} velocity words differ"
    finally:
        p.close()


@pytest.mark.parametrize("kind", ["ice", "stiff"])
def test_held_particles_match(kind):
    grid = (64, 48)
    p = Pair(grid, case=CASES[kind])
    try:
        up_to_projection(p)
        for mode in range(5):
            p.ora.stage("RIGID", mode)
        p.sync_to_gpu("VEL", "RL", "RBODY", "PARTICLES", "COMP", "AUX", "GRHO", "CELL", "SURF", "SURF2", "DHEAT")
        p.ora.begin_batch(1 / 240, 1)
        p.sync_to_gpu("STEP")
        p.both("STEP")
        p.both("CLEAR_ACC")
        p.both("PARTICLES", 1)
        pa = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-6
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-6
        a, b = p.f("ACC", np.int32, p.N * p.stride)
        assert np.count_nonzero(a != b) <= 0.01 * a.size
    finally:
        p.close()


@pytest.mark.parametrize("kind", ["ice", "stiff"])
def test_whole_rigid_substeps_match(kind):
    grid = (64, 48)
    p = Pair(grid, case=CASES[kind])
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
        rl_a, rl_b = p.f("RL", np.int32, (p.N, 2))
        assert np.count_nonzero(rl_a[:, 0] != rl_b[:, 0]) <= 0.002 * p.N
    finally:
        p.close()
