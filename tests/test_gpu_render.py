"""The render-facing outputs (FluidSim::particle_buffer, sdf_buffer, cell_buffer,
composition_buffer, surface_buffer, material_buffer, resets, shares_a_group; mod.rs:
1608-1943): each device view holds what the tank's other outputs report (read back through
HIP from the view's address), the host copies equal the views, the views follow the tank
when it is evicted and restored (a new generation, the same contents), and the tank's stream
orders a reader after the step.
"""
import ctypes as C

import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidSim

pytestmark = pytest.mark.gpu


def hip():
    lib = C.CDLL("libamdhip64.so")
    lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    lib.hipMemcpy.restype = C.c_int
    lib.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    lib.hipMemcpyAsync.restype = C.c_int
    lib.hipStreamSynchronize.argtypes = [C.c_void_p]
    lib.hipStreamSynchronize.restype = C.c_int
    return lib


def read_view(v, stream=None):
    """The view's bytes, copied from its device address (on `stream` when given)."""
    out = np.zeros(v.bytes, np.uint8)
    if v.bytes:
        h = hip()
        if stream is None:
            assert h.hipMemcpy(out.ctypes.data, v.ptr, v.bytes, 2) == 0
        else:
            assert h.hipMemcpyAsync(out.ctypes.data, v.ptr, v.bytes, 2, stream) == 0
            assert h.hipStreamSynchronize(stream) == 0
    return out


def four(grid=(96, 54)):
    params, field, parts, rows = cases.four_liquids(grid)
    sim = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    sim.set_solids(field)
    sim.reset(parts)
    return sim, field, parts


def test_views_hold_what_the_tank_reports():
    grid = (96, 54)
    W, H = grid
    sim, field, parts = four(grid)
    try:
        sim.record(8, 8 / 240)
        sim.wait()
        n = len(parts)
        vel, cell = sim.download_grid()
        _, mats = sim.grid_params()
        want = {
            "PARTICLES": (sim.field("PARTICLES", N.PARTICLE_DTYPE, n).view(np.uint8), n, 48),
            "CELL": (cell.view(np.uint8).reshape(-1), W * H, 16),
            "VELOCITY": (vel.view(np.uint8).reshape(-1), W * H, 16),
            "COMPOSITION": (sim.field("COMP", np.float32, (W * H, 4)).view(np.uint8).reshape(-1), W * H, 16),
            "SURFACE": (sim.field("SPROP", np.float32, (W * H, 4)).view(np.uint8).reshape(-1), W * H, 16),
            "MATERIALS": (np.frombuffer(bytes(mats), np.uint8), N.MAX_MATERIALS, C.sizeof(N.GMat)),
            "SDF": (np.ascontiguousarray(field.sdf, np.float32).view(np.uint8).reshape(-1), (W + 1) * (H + 1), 4),
        }
        for which, (data, count, stride) in want.items():
            v = sim.render_view(which)
            assert v.ptr and v.device == 0 and v.count == count and v.stride == stride, which
            assert v.bytes == count * stride == data.size, which
            got = read_view(v)
            assert np.array_equal(got, data), which
            assert np.array_equal(sim.render_download(which), got), which
        # the particle view is the slot array: the live particles as downloaded
        live = sim.download_particles()
        slots = read_view(sim.render_view("PARTICLES")).view(N.PARTICLE_DTYPE)
        assert np.array_equal(slots[slots["kind"] != 0], live)
    finally:
        sim.close()


def test_views_follow_the_tank_through_eviction():
    sim, _, _ = four()
    try:
        sim.record(4, 4 / 240)
        sim.wait()
        before = {w: read_view(sim.render_view(w)) for w in ("PARTICLES", "CELL", "COMPOSITION")}
        g0 = sim.render_view("CELL").generation
        sim.evict()
        v = sim.render_view("CELL")  # restores the tank
        assert v.generation != g0
        for w, data in before.items():
            assert np.array_equal(read_view(sim.render_view(w)), data), w
        # ordered on the tank's stream: a read queued behind the step sees the step's result
        s = sim.stream()
        assert s
        sim.record(4, 4 / 240)
        got = read_view(sim.render_view("CELL"), stream=s)
        sim.wait()
        assert np.array_equal(got, read_view(sim.render_view("CELL")))
    finally:
        sim.close()


def test_resets_and_shared_groups():
    grid = (64, 36)
    sim, _, parts = four(grid)
    try:
        assert sim.resets() == 1
        assert sim.shares_a_group()  # water and honey are both aqueous
        params, field, water, _ = cases.dam_break(grid)
        sim.reset(water)
        assert sim.resets() == 2
        assert not sim.shares_a_group()
    finally:
        sim.close()
