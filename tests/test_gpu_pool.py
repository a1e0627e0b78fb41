"""The continuous-batching pool (sp_pool_*): each scene's result in the pool is the result it
gets stepped alone with the same requests (one sp_record + sp_wait per request), bit for
bit. Requests arrive from another thread while the test's serving loop runs rounds,
same-size scenes of one schedule share launches, requests of different lengths overlap
(scenes join and leave rounds), and a budget that holds about half the scenes forces
evictions to host memory and restores between rounds.
"""
import random
import threading
import time

import pytest

from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidPool
from test_gpu_residency import SPECS, assert_same, build

pytestmark = pytest.mark.gpu


def run_pool(pool, arrivals, rng, spacing_s):
    """Submits `arrivals` (scene, substeps) from a feeder thread while this thread runs
    rounds; returns the tickets in arrival order."""
    tickets = []
    done = threading.Event()

    def feed():
        try:
            for k, n in arrivals:
                tickets.append(pool.submit(k, n, n / 240))
                time.sleep(rng.random() * spacing_s)
        finally:
            done.set()

    th = threading.Thread(target=feed)
    th.start()
    while not done.is_set():
        pool.step(wait_ms=1.0)
    th.join()
    pool.drain()
    return tickets


def test_pooled_scenes_match_each_scene_alone():
    rng = random.Random(11)
    plan = [[rng.choice((1, 1, 2, 3)) for _ in range(6)] for _ in SPECS]
    pooled = [build(sp) for sp in SPECS]
    alone = [build(sp) for sp in SPECS]
    counts = [len(s.download_particles()) for s in alone]
    sizes = [N.tank_bytes(sp[0][0], sp[0][1], sp[0][0] * sp[0][1] * 4)[0] for sp in SPECS]
    budget = sum(sorted(sizes)[-3:])  # about half the scenes at once
    pool = FluidPool(budget, streams=3)
    try:
        ids = [pool.add(s) for s in pooled]
        assert ids == list(range(len(SPECS)))
        pos = {k: 0 for k in range(len(SPECS))}
        arrivals = []
        order = list(range(len(SPECS))) * 6
        rng.shuffle(order)
        for k in order:
            arrivals.append((k, plan[k][pos[k]]))
            pos[k] += 1
        print(f"pool: {len(arrivals)} requests over {len(SPECS)} scenes, budget {budget} bytes", flush=True)
        tickets = run_pool(pool, arrivals, rng, 0.002)
        infos = [pool.wait(t) for t in tickets]
        assert all(i.status == 0 for i in infos)
        for (k, n), i in zip(arrivals, infos):
            assert i.scene == k and i.substeps == n
            assert i.submit_ms <= i.launch_ms <= i.done_ms
        st = pool.stats()
        print("pool stats:", {k: st[k] for k in ("rounds", "launches", "tanks_per_launch", "evictions",
                                                   "restores", "p50_ms", "p99_ms")}, flush=True)
        assert st["completed"] == len(arrivals) and st["failed"] == 0
        assert st["evictions"] > 0 and st["restores"] > 0
        assert st["resident_bytes"] <= budget
        assert st["tanks_per_launch"] > 1.0  # same-key scenes shared launches
        assert st["rounds"] >= max(sum(p) for p in plan)
        assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
        assert st["particle_steps"] == sum(counts[k] * n for k, n in arrivals)
        pool.close()
        # the same requests, each scene alone
        for k, s in enumerate(alone):
            for n in plan[k]:
                s.record(n, n / 240)
                s.wait()
        for k in range(len(SPECS)):
            assert_same(pooled[k], alone[k], k)
    finally:
        pool.close()
        for s in pooled + alone:
            s.close()


def test_pool_owns_its_scenes_until_closed():
    s = build(SPECS[0])
    pool = FluidPool(1 << 34, streams=1)
    try:
        pool.add(s)
        with pytest.raises(N.SproutError) as e:
            s.record(1, 1 / 240)
        assert e.value.code == N.E_BUSY
        t = pool.submit(0, 2, 2 / 240)
        assert pool.poll(t) is None  # nothing runs until a round does
        info = pool.wait(t)  # runs the rounds itself
        assert info.status == 0 and info.substeps == 2 and info.tanks == 1 and info.latency_ms > 0
        assert pool.stats()["rounds"] == 2
        with pytest.raises(N.SproutError):
            pool.submit(1, 1, 1 / 240)  # no such scene
        pool.close()
        s.record(1, 1 / 240)  # the caller's again
        s.wait()
        assert s.solver.substeps == 3
    finally:
        pool.close()
        s.close()
