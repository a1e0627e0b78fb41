"""The accessors of FluidSim that callers use besides the step (mod.rs:1946-1973, 3120-3150,
3342-3350, 3436-3462) through the C ABI: slots, dt_estimate, recorded_iters, substep_cost,
frame_blocking, profile_passes, the rigid labels and bodies, and the material table in the
reference's GpuMat layout (material.rs:508-549).
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidSim
from test_gpu_render import four, read_view

pytestmark = pytest.mark.gpu


def dam(grid=(96, 54)):
    params, field, parts, _ = cases.dam_break(grid)
    sim = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    sim.set_solids(field)
    sim.reset(parts)
    return sim, params, parts


def test_slots_dt_estimate_and_recorded_iterations():
    sim, params, parts = dam()
    try:
        assert sim.slots() == len(parts)
        sim.record(4, 4 / 240)
        sim.wait()
        assert sim.slots() == len(parts)
        assert sim.dt_estimate() == pytest.approx(1 / 240, rel=1e-6)  # fixed_hz 240
        rec = sim.recorded_iters()
        st = sim.solver
        # adapt (mod.rs:3100-3115): what the batch used, a margin of iters_margin, plus 2
        used = st.pressure_iters_max
        margin = max(params.iters_margin, 0.0)
        assert rec[0] == min(max(used + int(np.float32(used) * np.float32(margin)) + 2, 2), params.pressure_max_iters)
        assert rec[1] == rec[0]  # no viscous liquid: ready with what the first solve needs
        assert all(r > 0 for r in rec)
    finally:
        sim.close()


def test_substep_cost_and_frame_blocking():
    sim, params, _ = dam()
    try:
        assert sim.substep_cost() == 0.0
        covered = sim.frame_blocking(8, 8 / 240)
        assert covered == pytest.approx(8 / 240, rel=1e-5)
        c1 = sim.substep_cost()
        ph = sim.timings()
        assert c1 > 0.0
        # the batch's own device time per substep: at least the timed phases of a substep
        assert 0.5e-3 * ph[0] <= c1 <= 50e-3
        t0 = sim.sim_time()
        covered = sim.frame_blocking(4, 4 / 240)
        assert sim.sim_time() - t0 == pytest.approx(covered, rel=1e-9)
        c2 = sim.substep_cost()
        assert c2 > 0.0 and c2 != c1  # smoothed towards the new batch
        with pytest.raises(N.SproutError) as e:
            tool = N.ToolState(mode=9)  # no such tool
            sim.frame_blocking(1, 1 / 240, tool)
        assert e.value.code == N.E_INVALID
    finally:
        sim.close()


def test_profile_passes_runs_one_substep():
    sim, _, _ = dam()
    try:
        sim.record(2, 2 / 240)
        sim.wait()
        t0 = sim.sim_time()
        n0 = sim.solver.substeps
        passes = sim.profile_passes()
        names = [n for n, _ in passes]
        assert names == ["step_clear", "particles", "grid", "pressure", "project", "viscosity", "pressure2_project",
                         "total"]
        ms = dict(passes)
        assert all(v >= 0.0 for v in ms.values())
        assert ms["particles"] > 0.0 and ms["pressure"] > 0.0
        parts = sum(v for n, v in passes if n != "total")
        assert parts == pytest.approx(ms["total"], rel=1e-3, abs=1e-4)
        assert sim.solver.substeps == n0 + 1
        assert sim.sim_time() - t0 == pytest.approx(1 / 240, rel=1e-5)
    finally:
        sim.close()


def test_material_table_in_the_reference_layout():
    sim, field, parts = four()
    try:
        sim.record(1, 1 / 240)
        sim.wait()
        g, mats = sim.grid_params()
        t = sim.material_table()
        assert t.dtype.itemsize == 208
        for r in range(N.MAX_MATERIALS):
            m = mats[r]
            assert list(t["ia"][r]) == [m.group, m.diluent, m.chan_w, m.chan_s]
            assert list(t["ib"][r]) == [m.fam, m.rheo, m.temp_model, m.residue]
            assert np.array_equal(t["a"][r], np.float32([m.rho, m.ln_nu, m.sigma, m.mix_rate]))
            assert np.array_equal(t["b"][r], np.float32([m.cos_c, m.sin_c, m.grain, m.rho_solid]))
            assert np.array_equal(t["misc"][r], np.float32([m.freeze_shift, m.load, m.visc_exp, m.expansion]))
            assert np.array_equal(t["el"][r], np.float32([m.gr[1], m.gr[2], m.gr[3], m.residue_share]))
            assert np.array_equal(t["gr"][r][:2], np.float32([m.gr[0], m.fume]))
            assert np.array_equal(t["th"][r], np.float32(list(m.th)))
            assert np.array_equal(t["rx"][r], np.float32(list(m.rx)))
        # gpu_materials (mod.rs:2190-2207): an isothermal tank's relative density is the density
        # at the ambient temperature (the expansion folded in) over the reference water density
        p = sim.params
        for r in (N.ROW_WATER, N.ROW_MERCURY):
            m = p.mat[r]
            rho = m.density * min(max(1.0 - m.expansion * (p.ambient_temp - m.temp_ref), 0.7), 1.3)
            assert t["a"][r][0] == pytest.approx(rho / 998.0, rel=1e-6)  # SimParams::WATER_DENSITY
        # the device view holds the same table
        v = sim.render_view("GPUMAT")
        assert v.count == N.MAX_MATERIALS and v.stride == 208
        assert np.array_equal(read_view(v).view(N.GPUMAT_DTYPE), t)
    finally:
        sim.close()


def test_rigid_labels_and_bodies():
    grid = (64, 48)
    params, field, parts, _ = cases.ice_block(grid)
    sim = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    try:
        sim.set_solids(field)
        sim.reset(parts)
        sim.record(2, 2 / 240)
        sim.wait()
        lab = sim.download_labels().reshape(-1, 2)
        raw = sim.field("RL", np.int32, (grid[0] * grid[1], 2))
        assert np.array_equal(lab, raw)
        roots = np.flatnonzero(lab[:, 0] == np.arange(len(lab)))
        assert len(roots) >= 1  # the ice block
        bodies = sim.download_bodies()
        words = sim.field("RBODY", np.uint32, N.RIGID_WORDS)
        assert len(bodies) == min(int(words[0]), N.RIGID_BODIES)
        for b in range(len(bodies)):
            o = 16 + N.BODY_WORDS * b
            for q in range(7):
                v = int(words[o + 2 * q]) | int(words[o + 2 * q + 1]) << 32
                v = v - (1 << 64) if v >= 1 << 63 else v
                assert bodies[b, q] == v / 65536.0
        # a body's mass is its cells' (a frozen block of at least a few cells)
        slots = lab[roots, 1]
        assert np.all(slots >= 0)
        assert bodies[slots, 0].max() > 4.0
    finally:
        sim.close()
