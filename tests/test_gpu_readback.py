"""Read-backs of recorded batches taken up without blocking (FluidSim::poll_readbacks,
mod.rs:3030-3115) through the C ABI: sp_record returns with the batch in flight, sp_poll
takes up the batches that finished (simulated time, statistics, the next batches' recorded
iterations), and a tank driven with several frames in flight ends where the same tank
driven frame by frame (record, then wait) ends, bit for bit, while its solves converge.
"""
import time

import numpy as np
import pytest

import cases
from sprout_amd.fluid_sim import FluidSim

pytestmark = pytest.mark.gpu


def dam(grid=(128, 72)):
    params, field, parts, _ = cases.dam_break(grid)
    sim = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    sim.set_solids(field)
    sim.reset(parts)
    return sim


def test_frames_in_flight_end_where_frame_by_frame_ends():
    a, b = dam(), dam()
    try:
        frames = 24
        for _ in range(frames):
            a.record(4, 4 / 240)
            a.wait()
        for _ in range(frames):
            b.record(4, 4 / 240)  # at most four in flight: the fifth waits for the oldest
        b.wait()
        sa, sb = a.solver, b.solver
        print(f"\nframe by frame: rec {a.recorded_iters()}, in flight: rec {b.recorded_iters()}; "
              f"unconverged {sa.pressure_unconverged} / {sb.pressure_unconverged}", flush=True)
        assert sa.substeps == sb.substeps == 4 * frames
        assert a.sim_time() == pytest.approx(4 * frames / 240, rel=1e-6)
        assert b.sim_time() == a.sim_time()
        assert sa.pressure_unconverged == 0 and sb.pressure_unconverged == 0
        pa, pb = a.download_particles(), b.download_particles()
        assert pa.tobytes() == pb.tobytes()
        # both adapted from what the solves used; the tank in flight from older frames
        assert a.recorded_iters()[0] < 60 and b.recorded_iters()[0] < 60
    finally:
        a.close()
        b.close()


def test_poll_takes_up_finished_batches_without_waiting():
    sim = dam()
    try:
        assert sim.solver.substeps == 0
        sim.record(4, 4 / 240)
        rec0 = sim.recorded_iters()
        assert rec0[0] == sim.params.pressure_max_iters  # a new scene records the most
        t0 = time.perf_counter()
        while sim.solver.substeps < 4:  # sp_poll never blocks; the batch finishes on its own
            assert time.perf_counter() - t0 < 30.0
            time.sleep(0.001)
        assert sim.sim_time() == pytest.approx(4 / 240, rel=1e-6)
        assert sim.recorded_iters()[0] < rec0[0]  # adapted from the finished batch
        assert sim.substep_cost() > 0.0
        sim.wait()
        assert sim.solver.substeps == 4
    finally:
        sim.close()


def test_reset_takes_up_batches_in_flight_first():
    sim = dam()
    try:
        r0 = sim.resets()
        for _ in range(3):
            sim.record(4, 4 / 240)
        _, _, parts, _ = cases.dam_break((128, 72))
        sim.reset(parts)
        st = sim.solver
        assert st.substeps == 12  # counted towards the old scene
        assert sim.resets() == r0 + 1
        sim.record(4, 4 / 240)
        sim.wait()
        assert sim.solver.substeps == 16
        assert sim.sim_time() == pytest.approx(16 / 240, rel=1e-6)
        assert np.isfinite(sim.download_particles()["pos"]).all()
    finally:
        sim.close()
