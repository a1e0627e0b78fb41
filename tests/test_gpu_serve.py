"""The request server (sp_server_*) and tank eviction (sp_evict).

Each scene's result in the server is the result it gets stepped alone with the same
requests (one sp_record + sp_wait per request), bit for bit: requests arrive from another
thread while launches run, same-size scenes share launches, and a budget that holds about
half the scenes forces evictions to host memory and restores in between.
"""
import random
import threading
import time

import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidServer, FluidSim

pytestmark = pytest.mark.gpu

SPECS = [((64, 36), "dam", 0, (0.4, 0.6)), ((64, 36), "four", 1, None), ((96, 54), "dam", 0, (0.5, 0.5)),
         ((64, 36), "dam", 0, (0.6, 0.3)), ((96, 54), "four", 2, None), ((64, 36), "four", 3, None)]


def build(spec):
    grid, kind, seed, frac = spec
    if kind == "four":
        params, field, parts, _ = cases.four_liquids(grid, seed=seed, viscosity_tol=1e-5)
    else:
        params, field, parts, _ = cases.dam_break(grid, frac=frac)
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


def state(s):
    p = s.download_particles()
    vel, cell = s.download_grid()
    pres, _ = s.download_pressure()
    return p, vel, cell, pres


def assert_same(a, b, k):
    sa, sb = state(a), state(b)
    assert sa[0].tobytes() == sb[0].tobytes(), f"scene {k}: particles"
    for name, x, y in zip(("vel", "cell", "pres"), sa[1:], sb[1:]):
        assert np.array_equal(x, y), f"scene {k}: {name}"
    assert a.sim_time() == b.sim_time()
    for f in ("substeps", "pressure_iters_max", "viscosity_iters_max", "pressure_unconverged",
              "viscosity_unconverged", "pressure_residual_max", "viscosity_residual_max"):
        assert getattr(a.solver, f) == getattr(b.solver, f), f"scene {k}: {f}"


def test_evicted_tank_continues_bit_for_bit():
    a, b = build(SPECS[1]), build(SPECS[1])
    try:
        for s in (a, b):
            s.record(4, 4 / 240)
            s.wait()
        a.evict()
        assert not a.resident
        a.evict()  # evicting twice is harmless
        for s in (a, b):
            s.record(4, 4 / 240)
            s.wait()
        assert a.resident
        assert_same(a, b, 0)
    finally:
        a.close()
        b.close()


@pytest.mark.skip(reason="the request server is switched off: a served run stalled on the GPU "
                         "and its cause is not found")
def test_served_scenes_match_each_scene_alone():
    rng = random.Random(7)
    plan = [[rng.choice((1, 1, 2, 3)) for _ in range(6)] for _ in SPECS]
    served = [build(sp) for sp in SPECS]
    alone = [build(sp) for sp in SPECS]
    counts = [len(s.download_particles()) for s in alone]
    sizes = [N.tank_bytes(sp[0][0], sp[0][1], sp[0][0] * sp[0][1] * 4)[0] for sp in SPECS]
    budget = sum(sorted(sizes)[-3:])  # about half the scenes at once
    srv = FluidServer(budget, streams=3)
    try:
        ids = [srv.add(s) for s in served]
        assert ids == list(range(len(SPECS)))
        order = [(k, n) for k, reqs in enumerate(plan) for n in reqs]
        # arrival order: shuffled across scenes, in order within a scene
        pos = {k: 0 for k in range(len(SPECS))}
        arrivals = []
        pool = list(range(len(SPECS))) * 6
        rng.shuffle(pool)
        for k in pool:
            arrivals.append((k, plan[k][pos[k]]))
            pos[k] += 1
        assert sorted(arrivals) == sorted(order)
        tickets = []

        def feed():
            for k, n in arrivals:
                tickets.append(srv.submit(k, n, n / 240))
                time.sleep(rng.random() * 0.002)

        th = threading.Thread(target=feed)
        th.start()
        th.join()
        srv.drain()
        infos = [srv.wait(t) for t in tickets]
        assert all(i.status == 0 for i in infos)
        st = srv.stats()
        assert st["completed"] == len(arrivals) and st["failed"] == 0
        assert st["evictions"] > 0 and st["restores"] > 0
        assert st["resident_bytes"] <= budget
        assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        assert st["particle_steps"] == sum(counts[k] * n for k, n in arrivals)
        srv.close()
        # the same requests, each scene alone
        for k, s in enumerate(alone):
            for n in plan[k]:
                s.record(n, n / 240)
                s.wait()
        for k in range(len(SPECS)):
            assert_same(served[k], alone[k], k)
    finally:
        srv.close()
        for s in served + alone:
            s.close()


@pytest.mark.skip(reason="the request server is switched off: a served run stalled on the GPU "
                         "and its cause is not found")
def test_server_refuses_direct_use_of_its_scenes():
    s = build(SPECS[0])
    srv = FluidServer(1 << 34, streams=1)
    try:
        srv.add(s)
        with pytest.raises(N.SproutError) as e:
            s.record(1, 1 / 240)
        assert e.value.code == N.E_BUSY
        t = srv.submit(0, 2, 2 / 240)
        info = srv.wait(t)
        assert info.status == 0 and info.substeps == 2 and info.latency_ms > 0
        srv.close()
        s.record(1, 1 / 240)  # the caller's again
        s.wait()
        assert s.solver.substeps == 3
    finally:
        srv.close()
        s.close()
