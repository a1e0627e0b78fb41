"""Benchmark of the MI355X fluid step at 1024 x 576: particle-steps/s and pressure-solve ms,
with p50/p99 step latency batched and served.

One "step" is one substep of the fluid step (step, particle pass, grid passes, viscosity,
the MGPCG pressure solves, projections) over the whole tank. A single tank is recorded as the
reference's frame loop records it: one record per 1/60 s frame of 4 substeps, each adapting
its recorded iterations from the read-backs finished so far (poll_readbacks,
mod.rs:3030-3115), up to four frames in flight; batches of tanks record 16 substeps, then
wait. Inputs are resident in device memory when the timed region starts.

Headline (`value`, config.workload):
* four_liquids_1024x576_mixed_mg (default): oil, water, honey and mercury in four seeded
  rectangles tiling 80 % x 80 % of a 1024 x 576 tank (about 1.49 M particles at 4 per cell),
  densities 800 / 1000 / 1300 / 13500 kg/m^3, viscosities log-uniform over 1e-6 .. 1e-1 m^2/s
  and tensions uniform over 0.02 .. 0.07 N/m from seed 0; cell 5 mm, fixed dt 1/240 s;
  pressure tolerance 1e-6 and viscosity tolerance 1e-5; the implicit viscosity with the
  second pressure solve; the pressure multigrid's coarse levels in BF16 (iterates,
  right-hand sides) and FP16 (matrices), FP32 level 0, smoothing and CG vectors, the level-0
  residual in compensated FP32, FP64 reductions (--mg-precision fp32: FP32 levels).
  Other headline workloads: --workload dam (one block of water), two_fluid (--grid 512x288:
  water under oil, a honey blob, a 10-cell floor strip at 350 K), mixed (--tanks tanks of
  uneven sizes as one batch).

Legs beside the headline (each its own key):
* fp32_single: the same tank with FP32 multigrid levels, the plain FP32 path.
* batched: --batch-tanks tanks of the headline workload (seeds 0 ..) as one batch; ratio to
  fp32_single.
* batched_uneven: 64 dam-break tanks of uneven sizes (128x72 .. 512x288 by seed, blocks of
  30-60 %, water or oil, BF16/FP16 multigrid levels) as one batch for 500 substeps.
* served: one-substep requests for 64 resident uneven tanks arriving as a Poisson process at
  each of --serve-rates, run by the continuous-batching pool (sp_pool_*); p50/p99 latency
  from submit to done and the aggregate particle-steps/s.
* served_large: 256 resident scenes up to 2048 x 1152 with two to six liquids (sizes drawn
  128x72 35 %, 256x144 30 %, 512x288 20 %, 1024x576 10 %, 2048x1152 5 %; layered liquids of
  the two-fluid tank's ranges), under an HBM budget of half their total so that scenes are
  evicted to host memory and restored; --serve-large-requests requests at each of
  --serve-large-rates.

`roofline` is for the first pressure solve of a substep (the MGPCG phase: multigrid build,
cg_scale, cg_init, V-cycles and CG iterations), the largest phase of a substep. Algorithmic
bytes per solve (FP32 sizes of the reference layout, each array element a pass must read or
write counted once): 213 B per grid cell for the set-up (mg_build 25, cg_scale 24, cg_init
52, the first V-cycle 112) and 216 B per active-tile cell per iteration (cg_apply 52,
cg_update 52, V-cycle 112); with BF16/FP16 coarse levels the V-cycle's 112 B is 56 (157 and
160). The iterations are the timed solves' mean, the active cells those of the last timed
substep's tile flags; the time is the phase's HIP-event time per substep on the tank's
stream. `roofline_particles` is the same figure for the particle pass (96 B per live
particle + 100 B per grid cell). `cpu_baseline` times the CPU oracle (a C restatement of the
same step) on every host core at once, one tank per core, on a bounded sample.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
CHUNK = 16  # substeps per record of a batch of tanks (record, then wait)
# A single tank is driven as the reference's frame loop drives it: a record per 1/60 s frame
# (4 substeps at 240 Hz), each taking up the read-backs finished so far without blocking, a
# few frames in flight, and a wait every SYNC substeps (the phase timings hold 64 substeps).
FRAME = 4
SYNC = 64
PMC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_dam_break_1024x576_pmc.md")


def pmc_traffic(W, H, mgp, active_cells, iters_mean):
    """HBM bytes of one pressure solve from the committed FETCH_SIZE + WRITE_SIZE summary of
    the same kernels at 1024 x 576 (each counter scaled by what it reads for a plain copy of
    known size): the set-up's bytes plus the measured bytes per active cell and CG iteration
    times this solve's active cells and mean iterations. None where none was measured."""
    if (W, H) != (1024, 576) or not mgp or not os.path.exists(PMC_SUMMARY):
        return None, None
    vals = {}
    for line in open(PMC_SUMMARY):
        if line.startswith("traffic_") and ":" in line:
            k, v = line.split(":", 1)
            vals[k.strip()] = float(v)
    try:
        setup = vals["traffic_setup_bytes_per_solve_mixed"]
        per = vals["traffic_bytes_per_active_cell_iteration_mixed"]
    except KeyError:
        return None, None
    return setup + per * active_cells * iters_mean, (
        "FETCH_SIZE + WRITE_SIZE of the pressure-solve kernels on the 1024x576 dam break "
        "(profiles/hbm_traffic_dam_break_1024x576_pmc.md): set-up bytes per solve plus bytes per "
        "active cell and iteration, scaled to this solve's active cells and mean iterations")


def workload(kind, grid, seed=0, mg_precision=0):
    import cases
    if kind == "four":
        out = cases.four_liquids(grid, seed=seed, viscosity_tol=1e-5)
    elif kind == "two_fluid":
        out = cases.two_fluid_heat(grid, seed=seed)
    else:
        out = cases.dam_break(grid, frac=(0.8, 0.8))
    out[0].mg_precision = mg_precision
    return out


def make_sim(grid, params, field, parts):
    from sprout_amd.fluid_sim import FluidSim
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


class Tanks:
    """Tanks set up from (grid, params, field, particles) specs, stepped alone or as one
    batch."""

    def __init__(self, specs):
        from sprout_amd.fluid_sim import FluidBatch
        self.sims = [make_sim(*sp) for sp in specs]
        self.batch = FluidBatch(self.sims) if len(self.sims) > 1 else None

    def chunk(self):
        return CHUNK if self.batch is not None else SYNC

    def step(self, k):
        if self.batch is not None:
            self.batch.record(k, k / 240.0)
            self.batch.wait()
        else:
            s = self.sims[0]
            done = 0
            while done < k:
                f = min(FRAME, k - done)
                s.record(f, f / 240.0)
                done += f
            s.wait()

    def run(self, n):
        done = 0
        while done < n:
            k = min(self.chunk(), n - done)
            self.step(k)
            done += k

    def close(self):
        if self.batch is not None:
            self.batch.close()
        for s in self.sims:
            s.close()


def tune(specs, substeps=48):
    """Set-up: the particle pass's launch shape is tuned (every shape gives bit-identical
    results) on a throwaway copy of the tanks, which share the tuning key (grid, tanks per
    launch, accumulator stride), so the measured tanks run W warm-up and K timed substeps
    and nothing else."""
    t = Tanks(specs)
    try:
        t.run(substeps)
    finally:
        t.close()


def pct(sorted_ms, q):
    return float(sorted_ms[min(max(int(np.ceil(q / 100 * len(sorted_ms))), 1), len(sorted_ms)) - 1])


def throughput_leg(specs, steps, warmup, latency_steps, what):
    """Aggregate particle-steps/s of `specs` as one batch (or alone) over `steps` substeps
    after `warmup`, and the latency of one substep of all of them, record + wait."""
    tune(specs)
    t = Tanks(specs)
    try:
        n_p = sum(s.stats().live for s in t.sims)
        t.run(warmup)
        t0 = time.perf_counter()
        t.run(steps)
        wall = time.perf_counter() - t0
        lat = []
        for _ in range(latency_steps):
            t1 = time.perf_counter()
            t.step(1)
            lat.append(1e3 * (time.perf_counter() - t1))
        lat = np.sort(np.array(lat)) if lat else np.zeros(1)
        out = {"what": what, "tanks": len(specs), "particles": int(n_p), "steps": steps, "warmup": warmup,
               "particle_steps_per_s": n_p * steps / wall, "ms_per_step": 1e3 * wall / steps,
               "step_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "requests": latency_steps,
                                   "request": f"1 substep of all {len(specs)} tank(s), record + wait"}}
        if t.batch is not None:
            out["launch_groups"] = t.batch.groups()
        return out
    finally:
        t.close()


def served_uneven(rates, seconds, n_scenes=64, seed=0, streams=4):
    """64 resident uneven tanks, every scene resident; `seconds` of arrivals per rate."""
    import cases
    from sprout_amd import _native as N
    specs = [(g, p, f, parts) for g, p, f, parts, _ in cases.mixed_tanks(n_scenes, seed)]
    sims = [make_sim(*sp) for sp in specs]
    budget = sum(N.tank_bytes(s.grid[0], s.grid[1], s.capacity)[0] for s in sims)
    by_rate, n_p = serve_rates_requests(sims, rates, [int(r * seconds) for r in rates], budget, seed, streams)
    return {"scenes": n_scenes, "particles": n_p, "seconds_per_rate": seconds, "streams": streams,
            "request": "1 substep (1/240 s) of one scene", "budget": "every scene resident",
            "scene_sizes": "128x72, 256x144, 384x216, 512x288 by seed (dam-break blocks of 30-60 %, water or oil)",
            "by_rate": by_rate}


def serve_rates_requests(sims, rates, requests, budget, seed, streams):
    """Poisson arrivals of one-substep requests for scenes drawn uniformly from `sims`, the
    given number at each rate, submitted by a feeder thread and run by the pool from this
    thread. The pool owns the scenes; they are closed at the end."""
    out = []
    from sprout_amd.fluid_sim import FluidPool
    counts = [s.stats().live for s in sims]
    pool = FluidPool(budget, streams=streams)
    try:
        for s in sims:
            pool.add(s)
        for _ in range(2):
            for k in range(len(sims)):
                pool.submit(k, 1, 1 / 240)
            pool.drain()
        rng = np.random.default_rng(seed + 1)
        for rate, n in zip(rates, requests):
            out.append(run_rate(pool, len(sims), rate, max(n, 1), rng))
        return out, int(sum(counts))
    finally:
        pool.close()
        for s in sims:
            s.close()


def run_rate(pool, n_scenes, rate, n, rng):
    times = np.cumsum(rng.exponential(1.0 / rate, n))
    scenes = rng.integers(0, n_scenes, n)
    pool.reset_stats()
    done = threading.Event()

    def feed():
        try:
            t0 = time.perf_counter()
            i = 0
            while i < n:
                now = time.perf_counter() - t0
                while i < n and times[i] <= now:
                    pool.submit(int(scenes[i]), 1, 1 / 240)
                    i += 1
                if i < n:
                    time.sleep(min(max(times[i] - now, 0.0), 0.001))
        finally:
            done.set()

    th = threading.Thread(target=feed)
    t0 = time.perf_counter()
    th.start()
    while not done.is_set():
        pool.step(wait_ms=0.5)
    th.join()
    pool.drain()
    wall = time.perf_counter() - t0
    st = pool.stats()
    print(f"served: {rate:.0f}/s offered, {st['completed']} requests, p50 {st['p50_ms']:.2f} ms, "
          f"p99 {st['p99_ms']:.2f} ms, {st['evictions']} evictions", file=sys.stderr, flush=True)
    return {
        "offered_rate_per_s": rate,
        "requests": int(st["completed"]),
        "failed": int(st["failed"]),
        "served_rate_per_s": st["completed"] / wall,
        "particle_steps_per_s": st["particle_steps"] / wall,
        "latency_ms": {"p50": st["p50_ms"], "p99": st["p99_ms"], "mean": st["mean_ms"], "max": st["max_ms"]},
        "rounds": int(st["rounds"]),
        "scenes_per_round": st["completed"] / max(st["rounds"], 1),
        "tanks_per_launch": st["tanks_per_launch"],
        "evictions": int(st["evictions"]),
        "restores": int(st["restores"]),
        "resident_bytes_end": int(st["resident_bytes"]),
    }


LARGE_SIZES = ((128, 72), (256, 144), (512, 288), (1024, 576), (2048, 1152))
LARGE_SHARE = (0.35, 0.3, 0.2, 0.1, 0.05)


def served_large(rates, requests, n_scenes=256, seed=0, streams=4, budget_share=0.5):
    """256 scenes up to 2048 x 1152, two to six layered liquids each (one seeded template
    per size and liquid count, each scene its own resident copy), under an HBM budget of
    `budget_share` of their total."""
    import cases
    from sprout_amd import _native as N
    rng = np.random.default_rng(seed)
    picks = [(LARGE_SIZES[int(rng.choice(len(LARGE_SIZES), p=LARGE_SHARE))], int(rng.integers(2, 7)))
             for _ in range(n_scenes)]
    templates = {}
    for g, k in sorted(set(picks)):
        params, field, parts, _ = cases.layered_liquids(g, k, seed=k)
        templates[(g, k)] = (g, params, field, parts)
    sims = [make_sim(*templates[pk]) for pk in picks]
    total = sum(N.tank_bytes(s.grid[0], s.grid[1], s.capacity)[0] for s in sims)
    budget = int(budget_share * total)
    by_rate, n_p = serve_rates_requests(sims, rates, [requests] * len(rates), budget, seed, streams)
    sizes = {}
    for g, _ in picks:
        sizes[f"{g[0]}x{g[1]}"] = sizes.get(f"{g[0]}x{g[1]}", 0) + 1
    return {"scenes": n_scenes, "particles": n_p, "scene_sizes": sizes,
            "liquids_per_scene": "2-6 (uniform by seed), layered, densities 900-1400 kg/m^3, viscosities 1e-6-0.09 m^2/s",
            "templates": len(templates), "requests_per_rate": requests, "streams": streams,
            "request": "1 substep (1/240 s) of one scene",
            "budget_bytes": budget, "resident_total_bytes": int(total),
            "budget": f"{budget_share:.0%} of the scenes' total: scenes are evicted to host memory and restored",
            "by_rate": by_rate}


def cpu_baseline(grid, tables, field, parts, seconds=12.0, threads=None):
    """The oracle on every host core at once (one independent copy of the tank per thread;
    the oracle is single-threaded C and the binding releases the interpreter lock), after
    one untimed substep each: substeps per thread sized from a first timed substep to take
    about `seconds`."""
    from oracle.binding import OracleSim

    W, H = grid
    g, mats = tables
    g.n_slots = len(parts)
    n_thr = threads or min(os.cpu_count() or 1, 16)

    def make():
        o = OracleSim(grid, W * H * 4)
        o.set_gparams(g, mats)
        o.set_solids(field)
        o.reset(parts)
        o.begin_batch(1 / 240, 1)
        o.substep(0)
        return o

    probe = make()
    probe.begin_batch(1 / 240, 1)
    t0 = time.perf_counter()
    probe.substep(1)
    one = max(time.perf_counter() - t0, 1e-3)
    del probe
    substeps = int(min(max(round(seconds / one), 2), 400))
    sims = [make() for _ in range(n_thr)]
    barrier = threading.Barrier(n_thr + 1)
    spent = [0.0] * n_thr

    def work(k):
        o = sims[k]
        o.begin_batch(substeps / 240, substeps)
        barrier.wait()
        t1 = time.perf_counter()
        for _ in range(substeps):
            o.substep(1)
        spent[k] = time.perf_counter() - t1

    ths = [threading.Thread(target=work, args=(k,)) for k in range(n_thr)]
    for th in ths:
        th.start()
    barrier.wait()
    t0 = time.perf_counter()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    return {
        "value": n_thr * len(parts) * substeps / wall,
        "unit": "particle-steps/s",
        "cores": n_thr,
        "kind": "port",
        "sample": f"{substeps} substeps of {n_thr} copies of the same {W}x{H} tank ({len(parts)} particles each), "
                  f"one per thread on {n_thr} host cores, after one untimed substep each; C oracle (CPU restatement "
                  f"of the reference step), {wall:.1f} s; one copy alone: {len(parts) / one:.4g} particle-steps/s",
    }


def pressure_roofline(sim, W, H, mgp, ms_solve, iters_mean):
    from sprout_amd import _native as N
    L = N.levels(W, H)
    n0 = ((W + 15) // 16) * ((H + 15) // 16)
    ptile = sim.field("PTILE", np.uint32, (L.tiles,))[:n0]
    tx = (W + 15) // 16
    act = 0
    for t in np.flatnonzero(ptile):
        x0, y0 = (t % tx) * 16, (t // tx) * 16
        act += min(16, W - x0) * min(16, H - y0)
    setup_b = 157.0 if mgp else 213.0
    iter_b = 160.0 if mgp else 216.0
    alg = setup_b * W * H + iter_b * iters_mean * act
    achieved = alg / (ms_solve * 1e-3) / 1e9
    return {
        "kernel": "pressure solve (MGPCG phase: mg_build, cg_scale, cg_init, V-cycles, CG iterations)",
        "bound": "hbm",
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": None,
        "launch_ms": ms_solve,
        "alg_bytes_per_launch": alg,
        "launch": "one first pressure solve of a substep (its kernels, HIP events on the tank's stream)",
        "iterations_mean": iters_mean,
        "active_cells": int(act),
        "grid_cells": W * H,
        "bytes_per_cell_setup": setup_b,
        "bytes_per_active_cell_iteration": iter_b,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--grid", default="1024x576")
    ap.add_argument("--workload", choices=["four", "dam", "mixed", "two_fluid"], default="four")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="about how long the CPU baseline sample runs")
    ap.add_argument("--tanks", type=int, default=1,
                    help="tanks of the headline advanced together in one batch (seeds 0 .. tanks-1)")
    ap.add_argument("--latency-steps", type=int, default=64,
                    help="single-substep step requests timed one by one after the throughput run")
    ap.add_argument("--mg-precision", choices=["fp32", "mixed"], default="mixed",
                    help="pressure multigrid storage: FP32, or BF16 iterates / FP16 coarse matrices")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the solves' comparison with FP64 solves of the same systems")
    ap.add_argument("--no-fp32-leg", action="store_true")
    ap.add_argument("--batch-tanks", type=int, default=8, help="tanks of the batched leg; 0: none")
    ap.add_argument("--uneven-steps", type=int, default=500, help="substeps of the 64 uneven tanks; 0: no such leg")
    ap.add_argument("--serve-rates", default="1000,2000,4000",
                    help="request rates (per s) for the served leg over 64 uneven tanks; empty: none")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-large-rates", default="1000,5000,20000",
                    help="request rates (per s) for the served leg over 256 scenes up to 2048x1152; empty: none")
    ap.add_argument("--serve-large-requests", type=int, default=1500)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "a batch of tanks shards by scene across GPUs, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import cases
    import fp64_ref
    from sprout_amd import _native as N

    W, H = (int(x) for x in args.grid.split("x"))
    mgp = 1 if args.mg_precision == "mixed" else 0
    if args.workload == "mixed":
        mgp = 1
        specs = [(g, p, f, parts) for g, p, f, parts, _ in cases.mixed_tanks(max(args.tanks, 1), 0)]
        rows = None
    else:
        specs = []
        for seed in range(max(args.tanks, 1)):
            params, field, parts, rows_s = workload(args.workload, (W, H), seed, mgp)
            specs.append(((W, H), params, field, parts))
            if seed == 0:
                rows = rows_s
    _, params, field, parts = specs[0]
    print(f"bench: tuning the particle pass on a copy of the {len(specs)} tank(s)", file=sys.stderr, flush=True)
    tune(specs)
    T = Tanks(specs)
    sim = T.sims[0]
    n_p = sum(len(sp[3]) for sp in specs)
    print(f"bench: {len(specs)} tank(s), {n_p} particles set up", file=sys.stderr, flush=True)
    T.run(args.warmup)
    it0 = int(sim.field("SOLVE", N.SOLVE_DTYPE, 1)[0]["iters_sum"])
    ker, phases = [], []
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(T.chunk(), args.steps - done)
        T.step(k)
        if T.batch is None:
            ker.append(sim.kernel_time(0))
            phases.append(sim.timings().copy()[:7])
        else:
            bt = T.batch.timings(0)
            ker.append((float(bt[7]), k))
            phases.append(bt[:7].copy())
        done += k
    wall = time.perf_counter() - t0
    it1 = int(sim.field("SOLVE", N.SOLVE_DTYPE, 1)[0]["iters_sum"])
    iters_mean = ((it1 - it0) & 0xFFFFFFFF) / args.steps
    lat = []
    for _ in range(args.latency_steps):
        t1 = time.perf_counter()
        T.step(1)
        lat.append(1e3 * (time.perf_counter() - t1))
    lat = np.sort(np.array(lat)) if lat else np.zeros(1)
    st = sim.solver
    ms_step = 1e3 * wall / args.steps
    value = n_p * args.steps / wall
    launches = sum(n for _, n in ker)
    k_ms = sum(ms * n for ms, n in ker) / max(launches, 1)
    ph = np.mean(np.array(phases), axis=0)
    group_tanks = T.batch.groups()[0] if T.batch is not None else 1
    g0 = [sp for sp in specs if sp[0] == specs[0][0]][:group_tanks]
    part_bytes = sum(96.0 * len(sp[3]) + 100.0 * sp[0][0] * sp[0][1] for sp in g0)
    part_gbs = part_bytes / (k_ms * 1e-3) / 1e9
    live = sum(s.stats().live for s in T.sims)
    name = {"four": "four_liquids", "dam": "dam_break", "mixed": "mixed", "two_fluid": "two_fluid_heat"}[args.workload]
    if args.workload == "mixed":
        wname = f"mixed_{len(specs)}_tanks_bf16mg"
        sizes = sorted({sp[0] for sp in specs})
        grid_desc = {f"{g[0]}x{g[1]}": sum(1 for sp in specs if sp[0] == g) for g in sizes}
    else:
        wname = f"{name}_{W}x{H}_" + ("fp32" if mgp == 0 else "mixed_mg") + (f"_batch{len(specs)}" if len(specs) > 1 else "")
        grid_desc = [W, H]
    config = {
        "workload": wname,
        "tanks": len(specs),
        "grid": grid_desc,
        "particles": n_p,
        "live_particles_end": int(live),
        "cell_size_m": float(params.cell_size),
        "dt": "1/240 s fixed",
        "pressure_tol": float(params.pressure_tol),
        "substeps_per_record": FRAME if len(specs) == 1 else CHUNK,
        "driving": ("a record per 4 substeps (a 1/60 s frame), up to 4 in flight, each adapting its recorded "
                    f"iterations from the read-backs finished so far; a wait every {SYNC} substeps"
                    if len(specs) == 1 else f"a batch record of {CHUNK} substeps, then a wait"),
        "mg_storage": "fp32" if mgp == 0 else "bf16 iterates and coarse rhs, fp16 coarse matrices (fp32 smoothing, "
                                              "compensated-fp32 level-0 residual)",
    }
    if args.workload == "four":
        config["seed"] = 0
        config["liquids"] = [{"row": r, "density": params.mat[r].density, "viscosity": params.mat[r].viscosity,
                              "tension": params.mat[r].tension} for r in rows]
        config["viscosity_tol"] = float(params.viscosity_tol)
    roof = pressure_roofline(sim, sim.grid[0], sim.grid[1], mgp, float(ph[3]), iters_mean)
    if T.batch is None:
        roof["traffic"], src = pmc_traffic(W, H, mgp, roof["active_cells"], iters_mean)
        if src:
            roof["traffic_source"] = src
    out = {
        "metric": "particle-steps/s & pressure-solve ms at 1024x576; p50/p99 step latency, batched",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": config,
        "step_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": float(lat[-1]), "requests": args.latency_steps,
                            "request": f"1 substep of all {len(specs)} tank(s), record + wait"},
        "pressure_solve_ms": float(ph[3]),
        "pressure_iters_mean": iters_mean,
        "pressure2_solve_ms": float(ph[6]),
        "viscosity_solve_ms": float(ph[5]),
        "pressure_iters_max": int(st.pressure_iters_max),
        "viscosity_iters_max": int(st.viscosity_iters_max),
        "pressure_unconverged": int(st.pressure_unconverged),
        "viscosity_unconverged": int(st.viscosity_unconverged),
        "pressure_residual_max": float(st.pressure_residual_max),
        "viscosity_residual_max": float(st.viscosity_residual_max),
        "phase_ms_per_substep": {"particles": float(ph[1]), "grid": float(ph[2]), "pressure": float(ph[3]),
                                  "project": float(ph[4]), "viscosity": float(ph[5]),
                                  "pressure2_project": float(ph[6]), "device_total": float(ph[0])},
        "roofline": roof,
        "roofline_particles": {"kernel": "k_particles", "bound": "hbm", "achieved": part_gbs, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": part_gbs / HBM_PEAK_GBS, "launch_ms": k_ms,
                               "alg_bytes_per_launch": part_bytes},
        "cpu_baseline": None,
    }
    if not args.no_accuracy and T.batch is None:
        acc = {"what": "each solve of one more substep of the tank against the exact FP64 solution of the same "
                       "system (pressure: sparse LU; viscosity: FP64 CG to 1e-13), beside an FP64 CG stopped at the "
                       "same tolerance"}
        g, _ = sim.grid_params()
        if g.visc_mode == 2:
            ve, K = fp64_ref.viscosity_accuracy(sim, float(g.tol[2]), 1)
            acc["viscosity"] = dict(ve, iterations=int(K["v_iters"]))
            sim.record(1, 1 / 240)
            sim.wait()
        errs, vel_err, K = fp64_ref.pressure_accuracy(sim, (float(params.pressure_tol), float(params.correction_tol)),
                                                      gather=1)
        acc["pressure"] = errs[0]
        acc["correction"] = errs[1]
        acc["pressure_velocity_rel_l2_vs_fp64"] = vel_err
        acc["pressure_iterations"] = [int(v) for v in K["iters"]]
        out["accuracy"] = acc
    shapes = N.autotune_shapes()
    out["autotune"] = {
        "what": "particle-pass launch shape per (grid, tanks per launch, accumulator stride), chosen by timing "
                "each candidate on the launches themselves (every shape gives bit-identical results); tuned during "
                "set-up on a throwaway copy of each leg's tanks",
        "candidates": [{"threads": t, "box_ints": b, "min_groups_per_cu": m} for t, b, m in shapes],
        "chosen": [{"grid": [e["grid_w"], e["grid_h"]], "tanks": e["tanks"], "stride": e["stride"],
                    "shape": shapes[e["best"]] if e["best"] >= 0 else None, "best_ms": e["best_ms"],
                    "default_ms": e["default_ms"]} for e in N.autotune_report()],
    }
    tables0 = sim.grid_params()
    if T.batch is not None:
        out["launch_groups"] = T.batch.groups()
    T.close()
    print(f"bench: headline {value:.4g} particle-steps/s, {ms_step:.3f} ms per substep", file=sys.stderr, flush=True)
    if args.workload != "mixed":
        if not args.no_fp32_leg:
            p32, f32, parts32, _ = workload(args.workload, (W, H), 0, 0)
            out["fp32_single"] = throughput_leg([((W, H), p32, f32, parts32)], 96, 32, 0,
                                                "the headline tank with FP32 multigrid levels alone")
        if args.batch_tanks > 1:
            bspecs = []
            for seed in range(args.batch_tanks):
                p, f, pp, _ = workload(args.workload, (W, H), seed, mgp)
                bspecs.append(((W, H), p, f, pp))
            out["batched"] = throughput_leg(bspecs, 64, 32, 16,
                                            f"{args.batch_tanks} tanks of the headline workload (seeds 0 ..) as one batch")
            if "fp32_single" in out:
                out["batched"]["vs_plain_fp32_single"] = (out["batched"]["particle_steps_per_s"] /
                                                          out["fp32_single"]["particle_steps_per_s"])
            print(f"bench: batched {out['batched']['particle_steps_per_s']:.4g}", file=sys.stderr, flush=True)
    if args.uneven_steps > 0:
        uspecs = [(g, p, f, parts_u) for g, p, f, parts_u, _ in cases.mixed_tanks(64, 0)]
        out["batched_uneven"] = throughput_leg(uspecs, args.uneven_steps, 32, 16,
                                               "64 dam-break tanks of uneven sizes (128x72 .. 512x288 by seed, blocks "
                                               "of 30-60 %, water or oil, bf16/fp16 multigrid levels) as one batch")
        if "fp32_single" in out:
            out["batched_uneven"]["vs_plain_fp32_single"] = (out["batched_uneven"]["particle_steps_per_s"] /
                                                             out["fp32_single"]["particle_steps_per_s"])
        print(f"bench: uneven {out['batched_uneven']['particle_steps_per_s']:.4g}", file=sys.stderr, flush=True)
    if args.serve_rates:
        out["served"] = served_uneven([float(r) for r in args.serve_rates.split(",")], args.serve_seconds)
    if args.serve_large_rates:
        out["served_large"] = served_large([float(r) for r in args.serve_large_rates.split(",")],
                                           args.serve_large_requests)
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(specs[0][0], tables0, field, parts, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
