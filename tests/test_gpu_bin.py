"""The binning of particles by cell (bin.hip: sort.comp and claim.comp with a defined order)
against the textbook: numpy's stable argsort of the same keys, and the oracle's spawn step.

Every comparison is bit for bit: the sorted records, the permutation, the claimed slots.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidSim
from oracle.binding import OracleSim

pytestmark = pytest.mark.gpu


def keys_of(parts, W, H, tiled):
    """cell_of (sort.comp:28-30): the truncated position clamped into the grid; with `tiled`
    the 8 x 8 tile first; dead slots after every live key."""
    x = np.clip(np.trunc(parts["pos"][:, 0]).astype(np.int64), 0, W - 1)
    y = np.clip(np.trunc(parts["pos"][:, 1]).astype(np.int64), 0, H - 1)
    if tiled:
        k = (((y >> 3) * ((W + 7) >> 3) + (x >> 3)) << 6) | ((y & 7) << 3) | (x & 7)
    else:
        k = y * W + x
    return np.where(parts["kind"] == N.KIND_DEAD, np.int64(1) << 40, k)


def random_particles(n, W, H, seed, dead_share=0.1, near_sorted=False):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, N.PARTICLE_DTYPE)
    if near_sorted:
        # what the substep loop sorts: particles already in order that moved a little
        c = np.sort(rng.integers(0, W * H, n))
        p["pos"][:, 0] = (c % W) + rng.random(n) * 1.5 - 0.25
        p["pos"][:, 1] = (c // W) + rng.random(n) * 1.5 - 0.25
    else:
        # some outside the grid: the key clamps them to the border cells
        p["pos"][:, 0] = rng.random(n) * (W + 4) - 2
        p["pos"][:, 1] = rng.random(n) * (H + 4) - 2
    p["vel"] = rng.standard_normal((n, 2))
    p["kind"] = 1
    p["kind"][rng.random(n) < dead_share] = N.KIND_DEAD
    p["id"] = np.arange(n, dtype=np.uint32)
    return p


def device_bin(parts, W, H, tiled):
    n = len(parts)
    src = N.DeviceBuffer.of(parts)
    dst = N.DeviceBuffer(max(n, 1) * 48)
    perm = N.DeviceBuffer(max(n, 1) * 4)
    try:
        N.bin_particles(src.ptr, n, W, H, tiled, dst.ptr, perm.ptr)
        return dst.download(N.PARTICLE_DTYPE, n), perm.download(np.uint32, n)
    finally:
        for b in (src, dst, perm):
            b.close()


@pytest.mark.parametrize("n,grid,tiled,near", [
    (1, (16, 16), 0, False), (4095, (64, 36), 0, False), (4096, (64, 36), 1, False), (4097, (64, 36), 1, True),
    (100_003, (384, 216), 0, False), (100_003, (384, 216), 1, True), (300_000, (2048, 1152), 1, False),
    (1_494_912, (1024, 576), 1, True)])
def test_bin_particles_is_the_stable_sort(n, grid, tiled, near):
    W, H = grid
    parts = random_particles(n, W, H, seed=n + tiled, near_sorted=near)
    got, perm = device_bin(parts, W, H, tiled)
    order = np.argsort(keys_of(parts, W, H, tiled), kind="stable")
    assert np.array_equal(perm, order.astype(np.uint32))
    assert got.tobytes() == parts[order].tobytes()


def test_bin_particles_refuses_bad_arguments():
    L = N.lib()
    assert L.sp_bin_particles(None, 0, 64, 36, 0, None, None, None) == N.OK
    assert L.sp_bin_particles(None, 10, 64, 36, 0, None, None, None) == N.E_INVALID
    assert L.sp_bin_particles(None, 0, 0, 36, 0, None, None, None) == N.E_INVALID
    assert L.sp_bin_particles(None, 0, 64, 36, 2, None, None, None) == N.E_INVALID
    buf = N.DeviceBuffer(48 * 8)
    try:
        assert L.sp_bin_particles(buf.ptr, 8, 64, 36, 0, buf.ptr + 48, None, None) == N.E_INVALID
        assert b"overlap" in L.sp_error_text()
    finally:
        buf.close()


@pytest.mark.parametrize("fresh_room", [0, 37])
def test_claim_fills_dead_slots_in_slot_order(fresh_room):
    """Spawns beyond the fresh slots go to the dead slots, the k-th dead slot taking the k-th
    of them (claim.comp:10-17 with a defined order), as in the oracle's spawn step."""
    W, H = 96, 54
    params, field, parts, _ = cases.dam_break((W, H), frac=(0.4, 0.6))
    n0 = len(parts)
    rng = np.random.default_rng(7)
    slots = parts.copy()
    slots["id"] = np.arange(n0, dtype=np.uint32)
    dead = rng.random(n0) < 0.2
    slots["kind"][dead] = N.KIND_DEAD
    cap = n0 + fresh_room
    sim = FluidSim((W, H), cap, params=params)
    try:
        sim.set_solids(field)
        sim.reset(parts)
        sim.put_field("PARTICLES", slots)
        # more spawns than fresh slots and dead slots together: the rest are dropped
        m = fresh_room + int(dead.sum()) + 50
        sp = parts[rng.integers(0, n0, m)].copy()
        sp["pos"] += 0.25
        sp["kind"] = 1
        sp["t"] = params.ambient_temp
        sp["s"] = 1.0
        accepted = sim.spawn(sp)
        assert accepted == m
        sim.record(0, 0.0)
        sim.wait()
        got = sim.field("PARTICLES", N.PARTICLE_DTYPE, cap)
        # what the handle queued: ids after the reset's, the fresh mark (FluidSim::spawn)
        want_sp = sp.copy()
        want_sp["id"] = np.arange(n0, n0 + m, dtype=np.uint32)
        want_sp["kind"] |= N.KIND_FRESH
        o = OracleSim((W, H), cap)
        o.reset(slots)
        o.spawn(want_sp)
        want = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
        assert o.n_slots() == cap
        assert got.tobytes() == want.tobytes()
        assert int((got["kind"] == N.KIND_DEAD).sum()) == 0
    finally:
        sim.close()


def test_spawn_speed_bounds_the_time_step():
    """A fast pour into a still tank: the record's CFL step takes the spawns' speed
    (ext_speed = max(tool speed, spawn speed), mod.rs:2300; step.comp:25), and the record
    after it no longer does (the spawn speed is cleared once recorded, mod.rs:2520)."""
    W, H = 64, 36
    params, field, parts, _ = cases.dam_break((W, H), frac=(0.4, 0.6))
    params.fixed_hz = 0.0
    sim = FluidSim((W, H), W * H * 4, params=params)
    try:
        sim.set_solids(field)
        sim.reset(parts)
        sim.record(1, 1 / 240)
        sim.wait()
        g, _ = sim.grid_params()
        speed = 900.0  # cells / s, far above anything in the resting block
        sp = parts[:20].copy()
        sp["pos"][:, 1] = H - 6.0
        sp["vel"] = (0.0, -speed)
        sp["t"] = params.ambient_temp
        sp["s"] = 1.0
        assert sim.spawn(sp) == 20
        sim.record(1, 1 / 240)
        sim.wait()
        bound = g.cfl / (speed + np.sqrt(5.0 * max(g.gravity, 0.0)) + 1e-6)
        assert sim.solver.dt_cfl <= bound * (1 + 1e-6)
        step = sim.field("STEP", N.STEP_DTYPE, 1)[0]
        assert step["dt"] <= bound * (1 + 1e-6)
        g2, _ = sim.grid_params()
        assert g2.ext_speed == 0.0
    finally:
        sim.close()
