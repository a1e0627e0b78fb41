"""Benchmark of the MI355X fluid step: particle-steps/s and pressure-solve ms at 1024 x 576.

One "step" is one substep of the fluid step (step, particle pass, grid passes, MGPCG
pressure solve, projection) over the whole tank, recorded in batches of 16 substeps as the
reference's blocking helpers record them (advance_blocking, mod.rs:3284-3312). Inputs are
resident in device memory when the timed region starts.

Workload (named in config.workload): a 1024 x 576 dam break in FP32 — the grid of the
metric, with the plain single-liquid physics of the reference's CPU-runnable dam-break case
(water, cell 5 mm, fixed dt 1/240 s, pressure tolerance 1e-6, 3-cell solid frame) and a
block of 80 % x 80 % of the tank, about 1.49 M particles at 4 per cell (the particle count
of the four-liquid 1024 x 576 case, whose viscous liquids need the implicit viscosity solve
this build does not carry yet).

Prints one JSON line. `roofline` is for the particle pass (k_particles), the kernel with the
largest share of the substep: algorithmic bytes per launch = 96 B per live particle (its
48-byte state read and written) + 100 B per grid cell (face velocities 16, composition 16,
SDF 4, accumulator read-modify-write 64), over its HIP-event-timed average launch.
`cpu_baseline` times the CPU oracle (a single-threaded C restatement of the same step) on a
bounded sample of the same workload.
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
PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_particles_dam_break_1024x576_pmc.md")


def pmc_traffic(W, H, n_p):
    """HBM bytes per k_particles launch from the committed FETCH_SIZE + WRITE_SIZE summary
    (each counter scaled by what it reads for a plain copy of known size). It is quoted only
    for the workload it was measured on; any other grid or particle count gets null."""
    if (W, H, n_p) != (1024, 576, 1494912) or not os.path.exists(PMC_SUMMARY):
        return None
    for line in open(PMC_SUMMARY):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def workload(grid=(1024, 576), frac=(0.8, 0.8)):
    import cases
    return cases.dam_break(grid, frac=frac)


def cpu_baseline(grid, frac, params, field, parts, rows, substeps=4):
    """The oracle on the host's cores: `substeps` substeps of the same tank from the same
    state (the first one without a gather, as after a reset)."""
    from sprout_amd import _native as N
    from oracle.binding import OracleSim

    W, H = grid
    g, mats = N.derive_grid_params(params, W, H, rows)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    o.begin_batch(1 / 240, 1)
    o.substep(0)
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
        "sample": f"{substeps} substeps of the same {W}x{H} tank ({len(parts)} particles) after one warm-up "
                  f"substep, single-threaded C oracle, {dt:.1f} s",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=240)
    ap.add_argument("--warmup", type=int, default=32)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--grid", default="1024x576")
    ap.add_argument("--frac", default="0.8x0.8")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=24)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}); "
              "see DESIGN.md for how a tank would shard", file=sys.stderr)
        sys.exit(2)

    from sprout_amd.fluid_sim import FluidSim

    W, H = (int(x) for x in args.grid.split("x"))
    frac = tuple(float(x) for x in args.frac.split("x"))
    params, field, parts, rows = workload((W, H), frac)
    n_p = len(parts)
    sim = FluidSim((W, H), W * H * 4, params=params)
    sim.set_solids(field)
    sim.reset(parts)

    def run(n):
        done = 0
        while done < n:
            k = min(args.batch, n - done)
            sim.record(k, k / 240.0)
            sim.wait()
            done += k

    run(args.warmup)
    ker = []
    phases = []
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        k = min(args.batch, args.steps - done)
        sim.record(k, k / 240.0)
        sim.wait()
        ker.append(sim.kernel_time(0))
        phases.append(sim.timings().copy())
        done += k
    wall = time.perf_counter() - t0
    st = sim.solver
    ms_step = 1e3 * wall / args.steps
    value = n_p * args.steps / wall
    launches = sum(n for _, n in ker)
    k_ms = sum(ms * n for ms, n in ker) / max(launches, 1)
    ph = np.mean(np.array(phases), axis=0)
    alg_bytes = 96.0 * n_p + 100.0 * W * H
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    live = sim.stats().live
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
        "config": {
            "workload": f"dam_break_{W}x{H}_fp32",
            "grid": [W, H],
            "particles": n_p,
            "live_particles_end": int(live),
            "block": list(frac),
            "cell_size_m": 0.005,
            "dt": "1/240 s fixed",
            "pressure_tol": 1e-6,
            "substeps_per_batch": args.batch,
        },
        "pressure_solve_ms": float(ph[3]),
        "pressure_iters_max": int(st.pressure_iters_max),
        "pressure_unconverged": int(st.pressure_unconverged),
        "pressure_residual_max": float(st.pressure_residual_max),
        "phase_ms_per_substep": {"particles": float(ph[1]), "grid": float(ph[2]), "pressure": float(ph[3]),
                                  "project": float(ph[4]), "device_total": float(ph[0])},
        "roofline": {
            "kernel": "k_particles",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": pmc_traffic(W, H, n_p),
            "launch_ms": k_ms,
            "alg_bytes_per_launch": alg_bytes,
        },
        "cpu_baseline": None,
    }
    sim.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline((W, H), frac, params, field, parts, rows, args.cpu_substeps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
