"""Benchmark of the MI355X fluid step at 1024 x 576: particle-steps/s and pressure-solve ms.

One "step" is one substep of the fluid step (step, particle pass, grid passes, viscosity,
the MGPCG pressure solves, projections) over the whole tank, recorded in batches of 16
substeps as the reference's blocking helpers record them (advance_blocking,
mod.rs:3284-3312). Inputs are resident in device memory when the timed region starts.

Workloads (config.workload):
* four_liquids_1024x576_mixed_mg (default): oil, water, honey and mercury in four seeded
  rectangles tiling 80 % x 80 % of a 1024 x 576 tank (about 1.49 M particles at 4 per cell),
  densities 800 / 1000 / 1300 / 13500 kg/m^3, viscosities log-uniform over 1e-6 .. 1e-1 m^2/s
  and tensions uniform over 0.02 .. 0.07 N/m from seed 0; cell 5 mm, fixed dt 1/240 s;
  pressure tolerance 1e-6 and viscosity tolerance 1e-5 (the tightest the reference
  accepts); the implicit viscosity with the second pressure solve; the pressure multigrid's
  coarse levels in BF16 (iterates, right-hand sides) and FP16 (matrices), FP32 level 0,
  smoothing and CG vectors, FP64 reductions (--mg-precision fp32: FP32 levels).

Besides the single tank the line carries `fp32_single` (the same tank with FP32 multigrid
levels, the plain FP32 path) and `batched` (--batch-tanks tanks of the workload, seeds 0 ..,
advanced as one batch) with its ratio to the plain FP32 tank.
* dam_break_1024x576_fp32 (--workload dam): one block of water of the same size, explicit
  viscosity, one pressure solve.
* two_fluid_heat_512x288_fp32 (--workload two_fluid, --grid 512x288): water below, oil above,
  a honey-like blob dropped from the top centre, 1 cm cells, everything at 20 deg C on a floor
  strip held at 76.85 deg C (350 K): heat conduction, the implicit viscosity with the second
  pressure solve, and the rigid-body passes a thermal water tank runs.
* mixed_64_tanks_bf16mg (--workload mixed): 64 independent dam-break tanks advanced as one
  batch, grids drawn from 128x72, 256x144, 384x216 and 512x288 by seed, each block filling
  30-60 % of its tank, water or oil, pressure multigrid levels in BF16/FP16; tanks of one
  grid and schedule share every launch.

`served` (--serve-rates): step requests of one substep for scenes drawn uniformly from 64
resident mixed tanks, arriving as a Poisson process at each given rate, run by the
continuous-batching pool (sp_pool_*) from this process's serving loop; p50/p99 request
latency (submit to done) and the aggregate particle-steps/s of the served requests.

Prints one JSON line. `roofline` is for the particle pass (k_particles), the largest single
kernel of a substep: algorithmic bytes per launch = 96 B per live particle (its 48-byte
state read and written) + 100 B per grid cell (face velocities 16, composition 16, SDF 4,
accumulator read-modify-write 64), over its HIP-event-timed average launch. `cpu_baseline`
times the CPU oracle (a single-threaded C restatement of the same step) on a bounded sample
of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = {"dam": os.path.join(ROOT, "profiles", "k_particles_dam_break_1024x576_pmc.md"),
               "four": os.path.join(ROOT, "profiles", "k_particles_four_liquids_1024x576_pmc.md")}


def pmc_traffic(kind, W, H):
    """HBM bytes per k_particles launch from the committed FETCH_SIZE + WRITE_SIZE summary
    of this workload (each counter scaled by what it reads for a plain copy of known size),
    or None where none was measured."""
    path = PMC_SUMMARY.get(kind)
    if (W, H) != (1024, 576) or not path or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


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


def throughput_leg(kind, grid, n_tanks, mgp, steps, warmup, latency_steps, chunk=16):
    """A leg of the same workload: `n_tanks` tanks (seeds 0 ..) advanced as one batch (or the
    tank alone), multigrid storage `mgp`; aggregate particle-steps/s over `steps` substeps
    after `warmup`, and the latency of one substep of all of them, record + wait."""
    from sprout_amd.fluid_sim import FluidBatch, FluidSim

    W, H = grid
    sims = []
    for seed in range(n_tanks):
        params, field, parts, rows = workload(kind, grid, seed, mgp)
        sim = FluidSim(grid, W * H * 4, params=params)
        sim.set_solids(field)
        sim.reset(parts)
        sims.append(sim)
    n_p = sum(s.stats().live for s in sims)
    batch = FluidBatch(sims) if n_tanks > 1 else None

    def step(k):
        if batch is not None:
            batch.record(k, k / 240.0)
            batch.wait()
        else:
            sims[0].record(k, k / 240.0)
            sims[0].wait()

    for k0 in range(0, warmup, chunk):
        step(min(chunk, warmup - k0))
    t0 = time.perf_counter()
    for k0 in range(0, steps, chunk):
        step(min(chunk, steps - k0))
    wall = time.perf_counter() - t0
    lat = []
    for _ in range(latency_steps):
        t1 = time.perf_counter()
        step(1)
        lat.append(1e3 * (time.perf_counter() - t1))
    lat = np.sort(np.array(lat)) if lat else np.zeros(1)

    def pct(q):
        return float(lat[min(max(int(np.ceil(q / 100 * len(lat))), 1), len(lat)) - 1])
    out = {"tanks": n_tanks, "particles": int(n_p), "steps": steps, "warmup": warmup,
           "mg_storage": "fp32" if mgp == 0 else "bf16/fp16 coarse levels",
           "particle_steps_per_s": n_p * steps / wall, "ms_per_step": 1e3 * wall / steps,
           "step_latency_ms": {"p50": pct(50), "p99": pct(99), "requests": latency_steps,
                               "request": f"1 substep of all {n_tanks} tank(s), record + wait"}}
    if batch is not None:
        out["launch_groups"] = batch.groups()
        batch.close()
    for s in sims:
        s.close()
    return out


def served(rates, seconds, n_scenes=64, seed=0, streams=4):
    """The served leg: for each rate, Poisson arrivals of one-substep requests for scenes
    drawn uniformly from `n_scenes` resident mixed tanks over `seconds`, run by the pool
    from this thread while a feeder thread submits; then drained."""
    import threading
    import cases
    from sprout_amd import _native as N
    from sprout_amd.fluid_sim import FluidPool, FluidSim

    sims = []
    for grid, params, field, parts, rows in cases.mixed_tanks(n_scenes, seed):
        s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
        s.set_solids(field)
        s.reset(parts)
        sims.append(s)
    counts = [s.stats().live for s in sims]
    budget = sum(N.tank_bytes(s.grid[0], s.grid[1], s.capacity)[0] for s in sims)
    pool = FluidPool(budget, streams=streams)
    for s in sims:
        pool.add(s)
    # warm-up: every scene's first substeps (no gather after a reset, first sorts)
    for _ in range(3):
        for k in range(n_scenes):
            pool.submit(k, 1, 1 / 240)
        pool.drain()
    rng = np.random.default_rng(seed + 1)
    out = []
    for rate in rates:
        n = max(int(rate * seconds), 1)
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
        print(f"served: {rate:.0f}/s offered, {st['completed']} requests, p99 {st['p99_ms']:.2f} ms", file=sys.stderr,
              flush=True)
        out.append({
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
        })
    pool.close()
    for s in sims:
        s.close()
    return {"scenes": n_scenes, "particles": int(sum(counts)), "seconds_per_rate": seconds, "streams": streams,
            "request": "1 substep (1/240 s) of one scene", "budget": "every scene resident", "by_rate": out}


def cpu_baseline(grid, tables, field, parts, substeps):
    """The oracle on one host core: `substeps` substeps of the same tank from the same state,
    after one untimed substep (the first after a reset, without a gather). `tables`: the
    grid-unit tables the device tank derived (sp_sim grid_params)."""
    from oracle.binding import OracleSim

    W, H = grid
    g, mats = tables
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    o.begin_batch(1 / 240, 1)
    o.substep(0)
    if substeps <= 0:
        # about ten seconds of work: one substep measured first
        o.begin_batch(1 / 240, 1)
        t0 = time.perf_counter()
        o.substep(1)
        substeps = int(min(max(round(10.0 / max(time.perf_counter() - t0, 1e-3)), 4), 400))
    o.begin_batch(substeps / 240, substeps)
    t0 = time.perf_counter()
    for _ in range(substeps):
        o.substep(1)
    dt = time.perf_counter() - t0
    return {
        "value": len(parts) * substeps / dt,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{substeps} substeps of the same {W}x{H} tank ({len(parts)} particles) after the warm-up "
                  f"substeps, single-threaded C oracle, {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--grid", default="1024x576")
    ap.add_argument("--workload", choices=["four", "dam", "mixed", "two_fluid"], default="four")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=0,
                    help="substeps of the CPU baseline sample; 0: as many as take about ten seconds")
    ap.add_argument("--tanks", type=int, default=1,
                    help="tanks advanced together in one batch (seeds 0 .. tanks-1; value is their aggregate)")
    ap.add_argument("--latency-steps", type=int, default=64,
                    help="single-substep step requests timed one by one after the throughput run")
    ap.add_argument("--mg-precision", choices=["fp32", "mixed"], default="mixed",
                    help="pressure multigrid storage: FP32, or BF16 iterates / FP16 coarse matrices")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the pressure solve's comparison with an FP64 solve of the same system")
    ap.add_argument("--serve-rates", default="1000,2000,4000",
                    help="comma-separated request rates (per s) for the served leg over 64 mixed tanks; empty: none")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--batch-tanks", type=int, default=8,
                    help="tanks of the batched leg (the same workload, seeds 0 ..); 0: no batched leg")
    ap.add_argument("--no-fp32-leg", action="store_true",
                    help="skip the plain FP32 single-tank leg the batched leg is compared with")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "a tank shards by scene across GPUs, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    from sprout_amd.fluid_sim import FluidBatch, FluidSim

    W, H = (int(x) for x in args.grid.split("x"))
    tanks = []
    mgp = 1 if args.mg_precision == "mixed" else 0
    if args.workload == "mixed":
        import cases
        mgp = 1
        for grid, params, field, parts, rows in cases.mixed_tanks(max(args.tanks, 1), 0):
            sim = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
            sim.set_solids(field)
            sim.reset(parts)
            tanks.append((sim, params, field, parts, rows))
    else:
        for seed in range(max(args.tanks, 1)):
            params, field, parts, rows = workload(args.workload, (W, H), seed, mgp)
            sim = FluidSim((W, H), W * H * 4, params=params)
            sim.set_solids(field)
            sim.reset(parts)
            tanks.append((sim, params, field, parts, rows))
    sim, params, field, parts, rows = tanks[0]
    n_p = sum(len(t[3]) for t in tanks)
    batch = FluidBatch([t[0] for t in tanks]) if len(tanks) > 1 else None

    def step(k):
        if batch is not None:
            batch.record(k, k / 240.0)
            batch.wait()
        else:
            sim.record(k, k / 240.0)
            sim.wait()

    def run(n):
        done = 0
        while done < n:
            k = min(args.batch, n - done)
            step(k)
            done += k

    print(f"bench: {len(tanks)} tank(s), {n_p} particles set up", file=sys.stderr, flush=True)
    run(args.warmup)
    ker = []
    phases = []
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.batch, args.steps - done)
        step(k)
        if batch is None:
            ker.append(sim.kernel_time(0))
            phases.append(sim.timings().copy()[:7])
        else:
            bt = batch.timings(0)
            ker.append((float(bt[7]), k))
            phases.append(bt[:7].copy())
        done += k
    wall = time.perf_counter() - t0
    # step latency: one substep requested and waited for at a time (every tank of the batch)
    lat = []
    for _ in range(args.latency_steps):
        t1 = time.perf_counter()
        step(1)
        lat.append(1e3 * (time.perf_counter() - t1))
    lat = np.sort(np.array(lat)) if lat else np.zeros(1)

    def pct(p):
        return float(lat[min(max(int(np.ceil(p / 100 * len(lat))), 1), len(lat)) - 1])
    st = sim.solver
    ms_step = 1e3 * wall / args.steps
    value = n_p * args.steps / wall
    launches = sum(n for _, n in ker)
    k_ms = sum(ms * n for ms, n in ker) / max(launches, 1)
    ph = np.mean(np.array(phases), axis=0)
    # one particle-pass launch covers every tank of the (first) launch group
    group_tanks = batch.groups()[0] if batch is not None else 1
    if args.workload == "mixed":
        g0 = [t for t in tanks if t[0].grid == tanks[0][0].grid and t[4] == tanks[0][4]][:group_tanks]
        alg_bytes = sum(96.0 * len(t[3]) + 100.0 * t[0].grid[0] * t[0].grid[1] for t in g0)
    else:
        alg_bytes = group_tanks * (96.0 * len(parts) + 100.0 * W * H)
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    live = sum(t[0].stats().live for t in tanks)
    name = {"four": "four_liquids", "dam": "dam_break", "mixed": "mixed", "two_fluid": "two_fluid_heat"}[args.workload]
    if args.workload == "mixed":
        wname = f"mixed_{len(tanks)}_tanks_bf16mg"
        sizes = sorted({t[0].grid for t in tanks})
        grid_desc = {f"{g[0]}x{g[1]}": sum(1 for t in tanks if t[0].grid == g) for g in sizes}
    else:
        wname = f"{name}_{W}x{H}_" + ("fp32" if mgp == 0 else "mixed_mg") + (f"_batch{len(tanks)}" if len(tanks) > 1 else "")
        grid_desc = [W, H]
    config = {
        "workload": wname,
        "tanks": len(tanks),
        "grid": grid_desc,
        "particles": n_p,
        "live_particles_end": int(live),
        "cell_size_m": float(params.cell_size),
        "dt": "1/240 s fixed",
        "pressure_tol": float(params.pressure_tol),
        "substeps_per_batch": args.batch,
        "mg_storage": "fp32" if mgp == 0 else "bf16 iterates and coarse rhs, fp16 coarse matrices (fp32 smoothing)",
    }
    if args.workload == "four":
        config["seed"] = 0
        config["liquids"] = [{"row": r, "density": params.mat[r].density, "viscosity": params.mat[r].viscosity,
                              "tension": params.mat[r].tension} for r in rows]
        config["viscosity_tol"] = float(params.viscosity_tol)
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
        "step_latency_ms": {"p50": pct(50), "p99": pct(99), "max": float(lat[-1]), "requests": args.latency_steps,
                            "request": f"1 substep of all {len(tanks)} tank(s), record + wait"},
        "pressure_solve_ms": float(ph[3]),
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
        "roofline": {
            "kernel": "k_particles",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": pmc_traffic(args.workload, W, H),
            "launch_ms": k_ms,
            "alg_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": None,
    }
    if not args.no_accuracy:
        import fp64_ref
        errs, vel_err, K = fp64_ref.pressure_accuracy(sim, float(params.pressure_tol), gather=1)
        out["accuracy"] = {
            "what": "one more substep of tank 0 up to its first pressure solve, then that solve and its projection, "
                    "against the exact FP64 solution of the same system (sparse LU)",
            "pressure_err_vs_fp64": errs[0]["err_vs_fp64"] if errs[0] else None,
            "pressure_true_residual": errs[0]["true_residual"] if errs[0] else None,
            "pressure_fp64_cg_same_tol_err": errs[0]["fp64_cg_same_tol_err"] if errs[0] else None,
            "correction_err_vs_fp64": errs[1]["err_vs_fp64"] if errs[1] else None,
            "correction_true_residual": errs[1]["true_residual"] if errs[1] else None,
            "velocity_rel_l2_vs_fp64": vel_err,
            "iterations": [int(v) for v in K["iters"]],
        }
    from sprout_amd import _native as N
    shapes = N.autotune_shapes()
    out["autotune"] = {
        "what": "particle-pass launch shape per (grid, tanks per launch, accumulator stride), chosen by timing "
                "each candidate on the launches themselves (every shape gives bit-identical results)",
        "candidates": [{"threads": t, "box_ints": b, "min_groups_per_cu": m} for t, b, m in shapes],
        "chosen": [{"grid": [e["grid_w"], e["grid_h"]], "tanks": e["tanks"], "stride": e["stride"],
                    "shape": shapes[e["best"]] if e["best"] >= 0 else None, "best_ms": e["best_ms"],
                    "default_ms": e["default_ms"]} for e in N.autotune_report()],
    }
    tables0 = tanks[0][0].grid_params()
    if batch is not None:
        out["launch_groups"] = batch.groups()
        batch.close()
    for t in tanks:
        t[0].close()
    if args.workload != "mixed":
        if not args.no_fp32_leg:
            out["fp32_single"] = throughput_leg(args.workload, (W, H), 1, 0, 96, 32, 0)
        if args.batch_tanks > 1:
            out["batched"] = throughput_leg(args.workload, (W, H), args.batch_tanks, mgp, 64, 32, 16)
            if "fp32_single" in out:
                out["batched"]["vs_plain_fp32_single"] = (out["batched"]["particle_steps_per_s"] /
                                                          out["fp32_single"]["particle_steps_per_s"])
    if args.serve_rates:
        out["served"] = served([float(r) for r in args.serve_rates.split(",")], args.serve_seconds)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tanks[0][0].grid, tables0, field, parts, args.cpu_substeps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
