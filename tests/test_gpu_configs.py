"""The five benchmark workloads at their stated sizes on the device, and the reference's own
validation checks (src/validate.rs) on the HIP path.

* Dam break, 384 x 216, 76,153 particles: after one substep the grid velocities and the
  pressure against the oracle (rel-L2 1e-5), then the whole 100-substep horizon against the
  oracle particle by particle (by id).
* Two fluids with heat, 512 x 288 (water under oil, a honey blob, a 10-cell floor strip at
  350 K): 20 substeps against the oracle (positions, temperatures, the heat ledger), then on
  to the 200-substep horizon against the oracle.
* Four liquids, 1024 x 576 (about 1.5 M particles), mixed-precision multigrid: two substeps
  against the oracle (which solves in FP32), then a viscosity solve and a pressure solve
  each against the exact FP64 solution of its system.
* 64 dam-break tanks of uneven sizes (128 x 72 .. 512 x 288, water or oil, BF16 multigrid
  levels) advanced as one batch for 500 substeps; two tanks of every size bit-identical to
  themselves stepped alone.
* A continuous-batching pool of scenes up to 2048 x 1152 with two to six liquids, under an
  HBM budget that forces evictions; every scene bit-identical to itself stepped alone.
* validate.rs on the device: a resting pool is hydrostatic (validate.rs:125-141), and the
  discrete divergence after the projection is at the solver's tolerance (rhs.comp:49,
  project.comp:26-57).

The device and the oracle reduce the solves' FP64 sums in different orders, so their last
bits differ; one substep is held to rel-L2 1e-5 and longer horizons to the bars stated per
test, which leave room for that difference to grow with the flow.
"""
import random

import numpy as np
import pytest

import cases
import fp64_ref
from sprout_amd import _native as N
from sprout_amd import scene as sc
from sprout_amd.fluid_sim import FluidBatch, FluidPool, FluidSim
from oracle.binding import OracleSim
from test_gpu_parity import Pair, first_substep_prefix, rel_l2
from test_gpu_pool import run_pool
from test_gpu_residency import assert_same
from test_oracle import discrete_divergence

pytestmark = pytest.mark.gpu


def by_id(p):
    return p[np.argsort(p["id"], kind="stable")]


def device_and_oracle(grid, case, **kw):
    params, field, parts, _ = case(grid, **kw)
    W, H = grid
    sim = FluidSim(grid, W * H * 4, params=params)
    sim.set_solids(field)
    sim.reset(parts)
    g, mats = sim.grid_params()
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return sim, o, params, field, parts


def advance(sim, o, n, first):
    """n substeps as one batch on both sides (the first substep after a reset has no
    gather)."""
    sim.record(n, n / 240)
    o.begin_batch(n / 240, n)
    for k in range(n):
        o.substep(0 if (first and k == 0) else 1)
    sim.wait()


def particles(sim, o):
    a = by_id(sim.download_particles())
    b = by_id(o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots()))
    b = b[b["kind"] != 0] if len(b) != len(a) else b
    assert np.array_equal(a["id"], b["id"])
    return a, b


def ledger(x):
    L = x.field("LEDGER", np.uint32, N.LEDGER_WORDS)[:16].astype(np.uint64)
    return [float((int(L[2 * e]) | int(L[2 * e + 1]) << 32) / 65536.0) for e in range(4)]


# ---- 384 x 216 dam break, 100 substeps -----------------------------------------------------

def test_dam_break_384x216_one_substep_then_one_hundred():
    grid = (384, 216)
    W, H = grid
    sim, o, params, _, parts = device_and_oracle(grid, cases.dam_break)
    try:
        assert len(parts) == 76153
        advance(sim, o, 1, True)
        va, _ = sim.download_grid()
        vb = o.field("VEL", np.float32, (W * H, 4))
        ev = fp64_ref.live_faces_rel_l2(va, vb)
        pa, dt = sim.download_pressure()
        pb = o.field("PRES", np.float32, (W * H, 2))[:, 0].reshape(H, W)
        ep = rel_l2(pa, pb)
        print(f"dam break 384x216, 1 substep: grid velocity rel-L2 {ev:.3g}, pressure {ep:.3g}", flush=True)
        assert ev <= 1e-5 and ep <= 1e-5
        a, b = particles(sim, o)
        assert rel_l2(a["pos"], b["pos"]) <= 1e-6
        advance(sim, o, 99, False)
        a, b = particles(sim, o)
        err, verr = rel_l2(a["pos"], b["pos"]), rel_l2(a["vel"], b["vel"])
        print(f"dam break 384x216, 100 substeps: positions rel-L2 {err:.3g}, velocities {verr:.3g}", flush=True)
        assert err <= 1e-4 and verr <= 1e-2
        assert np.all((a["pos"][:, 0] > 0) & (a["pos"][:, 0] < W) & (a["pos"][:, 1] > 0) & (a["pos"][:, 1] < H))
        assert abs(sim.sim_time() - 100 / 240) < 1e-6
        st = sim.solver
        print(f"  solver: {st.substeps} substeps, pressure iterations max {st.pressure_iters_max}, "
              f"unconverged {st.pressure_unconverged}, residual max {st.pressure_residual_max:.3g}", flush=True)
        assert st.substeps == 100
    finally:
        sim.close()


# ---- 512 x 288 two fluids with heat, 200 substeps ------------------------------------------

def test_two_fluids_with_heat_512x288():
    grid = (512, 288)
    sim, o, params, _, parts = device_and_oracle(grid, cases.two_fluid_heat)
    try:
        g, _ = sim.grid_params()
        assert g.caps & N.CAP_THERMAL
        assert g.visc_mode == 2  # the honey's viscosity takes the implicit solve
        advance(sim, o, 20, True)
        a, b = particles(sim, o)
        err = rel_l2(a["pos"], b["pos"])
        terr = rel_l2(a["t"] + 273.15, b["t"] + 273.15)
        la, lb = ledger(sim), ledger(o)
        print(f"two fluids 512x288, 20 substeps: positions rel-L2 {err:.3g}, temperatures {terr:.3g}, "
              f"ledger {la} vs {lb}", flush=True)
        assert err <= 1e-5 and terr <= 1e-6
        assert la[2] > 0.0  # heat came in from the strip
        for x, y in zip(la, lb):
            assert abs(x - y) <= 1e-4 * max(abs(y), 1.0)
        # The falling blob stirs the two layers. The solves' last-bit differences stay at
        # 1e-7 until a threshold rule of the particle pass (a crowded cell's jitter, a wall
        # contact) decides one particle differently; that particle's neighbourhood then
        # drifts apart. Printed every 20 substeps: the rel-L2 and how many particles moved
        # apart by more than 0.01 cell.
        curve = [(err, 0)]
        for k in range(9):
            advance(sim, o, 20, False)
            a, b = particles(sim, o)
            far = int((np.abs(a["pos"] - b["pos"]).max(1) > 0.01).sum())
            curve.append((rel_l2(a["pos"], b["pos"]), far))
        terr = rel_l2(a["t"] + 273.15, b["t"] + 273.15)
        la, lb = ledger(sim), ledger(o)
        print(f"two fluids 512x288, every 20 substeps to 200 (rel-L2, particles > 0.01 cell apart of {len(a)}): "
              f"{[(float(f'{e:.2g}'), n) for e, n in curve]}, temperatures {terr:.3g}, ledger {la} vs {lb}",
              flush=True)
        assert curve[-1][0] <= 1e-2 and terr <= 1e-5
        # Past the divergence of single particles the stirred region is compared by what does
        # not depend on which particle went where: each liquid's centroid and kinetic energy.
        for kind in np.unique(b["kind"]):
            ma, mb = a["kind"] == kind, b["kind"] == kind
            assert ma.sum() == mb.sum()
            ca, cb = a["pos"][ma].astype(np.float64).mean(0), b["pos"][mb].astype(np.float64).mean(0)
            ka = float((a["vel"][ma].astype(np.float64) ** 2).sum())
            kb = float((b["vel"][mb].astype(np.float64) ** 2).sum())
            print(f"  liquid {kind}: centroid {ca} vs {cb}, kinetic energy ratio {ka / max(kb, 1e-30):.6f}", flush=True)
            assert np.abs(ca - cb).max() <= 0.25
            assert abs(ka - kb) <= 0.05 * kb + 1e-6 * ma.sum()
        for x, y in zip(la, lb):
            assert abs(x - y) <= 1e-3 * max(abs(y), 1.0)
    finally:
        sim.close()


# ---- 1024 x 576 four liquids, mixed-precision multigrid ------------------------------------

def test_four_liquids_1024x576_mixed_against_the_oracle_and_fp64():
    grid = (1024, 576)
    sim, o, params, field, parts = device_and_oracle(grid, cases.four_liquids, viscosity_tol=1e-5)
    sim.close()  # the oracle keeps the FP32 tables; the device solves with BF16/FP16 storage
    params.mg_precision = 1
    W, H = grid
    sim = FluidSim(grid, W * H * 4, params=params)
    try:
        sim.set_solids(field)
        sim.reset(parts)
        assert len(parts) > 1_400_000
        g, _ = sim.grid_params()
        assert g.visc_mode == 2 and g.mg_half
        advance(sim, o, 2, True)
        a, b = particles(sim, o)
        err, verr = rel_l2(a["pos"], b["pos"]), rel_l2(a["vel"], b["vel"])
        print(f"four liquids 1024x576, 2 substeps, mixed device vs FP32 oracle: positions rel-L2 {err:.3g}, "
              f"velocities {verr:.3g}", flush=True)
        assert err <= 1e-5 and verr <= 1e-3
        st = sim.solver
        assert st.pressure_unconverged == 0 and st.viscosity_unconverged == 0
        ve, K = fp64_ref.viscosity_accuracy(sim, float(g.tol[2]), 1)
        print(f"  viscosity solve vs FP64: {ve}, iterations {K['v_iters']}", flush=True)
        assert K["v_done"] == 1
        assert ve["vel_err_vs_fp64"] <= max(1e-5, ve["fp64_cg_same_tol_vel_err"])
        assert ve["true_residual"] <= 3 * float(g.tol[2])
        sim.record(1, 1 / 240)
        sim.wait()
        pe, vel_err, K = fp64_ref.pressure_accuracy(sim, params.pressure_tol, 1)
        print(f"  pressure solve vs FP64: {pe[0]}, projected face velocities rel-L2 {vel_err:.3g}, "
              f"iterations {K['iters']}", flush=True)
        assert K["done"][0] == 1
        assert np.sqrt(K["rr"][0] / K["bb"][0]) <= params.pressure_tol * (1 + 1e-6)
        # The compensated level-0 residual takes the error against the exact solution to
        # about 1e-7 (an FP32 CG's own is about 1e-5 here); the true residual of an FP32
        # iterate cannot go below eps |A| |x| / |b|, and the projected velocities are held
        # to 1e-5 or to the change one ulp of the pressure makes, whichever is larger.
        assert pe[0]["err_vs_fp64"] <= 1e-6
        assert pe[0]["true_residual"] <= max(2 * pe[0]["fp32_floor"], 4 * params.pressure_tol)
        assert vel_err <= max(1e-5, 1.5 * pe[0]["vel_floor_1ulp"])
    finally:
        sim.close()


# ---- 64 tanks of uneven sizes in one batch, 500 substeps -----------------------------------

def build_tank(spec):
    grid, params, field, parts, _ = spec
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


def test_sixty_four_uneven_tanks_in_one_batch_500_substeps():
    specs = cases.mixed_tanks(64, seed=0, mg_precision=1)
    assert {sp[0] for sp in specs} == set(cases.MIXED_SIZES)
    chosen = []
    for size in cases.MIXED_SIZES:
        chosen += [k for k, sp in enumerate(specs) if sp[0] == size][:2]
    tanks = [build_tank(sp) for sp in specs]
    batch = FluidBatch(tanks)
    try:
        for k in range(25):
            batch.record(20, 20 / 240)
            batch.wait()
            if k == 0:
                print(f"64 tanks: launch groups {batch.groups()}", flush=True)
        for k, s in enumerate(tanks):
            assert s.solver.substeps == 500, k
            p = s.download_particles()
            assert len(p) == len(specs[k][3])
            assert np.all(np.isfinite(p["pos"])) and np.all(np.isfinite(p["vel"]))
        for k in chosen:
            alone = build_tank(specs[k])
            try:
                for _ in range(25):
                    alone.record(20, 20 / 240)
                    alone.wait()
                assert_same(tanks[k], alone, k)
            finally:
                alone.close()
    finally:
        batch.close()
        for s in tanks:
            s.close()


# ---- served scenes up to 2048 x 1152 with two to six liquids -------------------------------

SERVED = [((2048, 1152), 6, 0), ((1024, 576), 2, 1), ((1024, 576), 5, 2), ((512, 288), 3, 3),
          ((512, 288), 4, 4), ((512, 288), 6, 5), ((256, 144), 2, 6), ((256, 144), 3, 7)]


def build_served(spec):
    grid, k, seed = spec
    params, field, parts, _ = cases.layered_liquids(grid, k, seed=seed)
    s = FluidSim(grid, grid[0] * grid[1] * 4, params=params)
    s.set_solids(field)
    s.reset(parts)
    return s


def test_served_scenes_up_to_2048x1152_under_an_hbm_budget():
    rng = random.Random(5)
    plan = [[rng.choice((1, 1, 2)) for _ in range(3)] for _ in SERVED]
    sizes = [N.tank_bytes(sp[0][0], sp[0][1], sp[0][0] * sp[0][1] * 4)[0] for sp in SERVED]
    budget = max(sizes) + sorted(sizes)[-2]  # the largest scene and one other at once
    assert budget < sum(sizes)
    pooled = [build_served(sp) for sp in SERVED]
    pool = FluidPool(budget, streams=3)
    try:
        for s in pooled:
            pool.add(s)
        pos = {k: 0 for k in range(len(SERVED))}
        order = list(range(len(SERVED))) * 3
        rng.shuffle(order)
        arrivals = []
        for k in order:
            arrivals.append((k, plan[k][pos[k]]))
            pos[k] += 1
        tickets = run_pool(pool, arrivals, rng, 0.003)
        infos = [pool.wait(t) for t in tickets]
        assert all(i.status == 0 for i in infos)
        st = pool.stats()
        print("served pool:", {k: st[k] for k in ("rounds", "launches", "tanks_per_launch", "evictions", "restores",
                                                  "p50_ms", "p99_ms")}, flush=True)
        assert st["completed"] == len(arrivals) and st["failed"] == 0
        assert st["evictions"] > 0 and st["restores"] > 0
        assert st["resident_bytes"] <= budget
        pool.close()
        for k, spec in enumerate(SERVED):
            alone = build_served(spec)
            try:
                for n in plan[k]:
                    alone.record(n, n / 240)
                    alone.wait()
                assert_same(pooled[k], alone, k)
            finally:
                alone.close()
    finally:
        pool.close()
        for s in pooled:
            s.close()


# ---- validate.rs on the device -------------------------------------------------------------

def test_hydrostatic_pool_on_the_device():
    """validate.rs:125-141: water 100 cells deep on the 512 x 288 grid, ten seconds: no
    particle lost, the surface (99.5th percentile of the heights, less the wall) within 2 %
    of 100 cells, the rms speed under 1 % of sqrt(g h); and below the surface the pressure
    is rho g depth (the pressure field holds p dt)."""
    grid = (512, 288)
    W, H = grid
    params = N.default_params()
    scene = sc.Scene(grid)
    wall = float(scene.wall)
    parts = cases.settle(sc.fill_rect(scene, (0.0, 0.0, float(W), 100.0 + wall), N.ROW_WATER), params)
    sim = FluidSim(grid, W * H * 4, params=params)
    try:
        sim.set_solids(scene.rasterize())
        sim.reset(parts)
        sim.advance_blocking(10.0)
        g, mats = sim.grid_params()
        p = sim.download_particles()
        assert len(p) == len(parts)
        top = float(np.percentile(p["pos"][:, 1], 99.5)) - wall
        v_ref = np.sqrt(g.gravity * 100.0)
        rms = float(np.sqrt(np.mean((p["vel"].astype(np.float64) ** 2).sum(1))))
        print(f"hydrostatic pool: {len(p)} particles, surface {top:.2f} cells, rms speed / sqrt(g h) "
              f"{rms / v_ref:.5f}, {sim.solver.substeps} substeps", flush=True)
        assert abs(top / 100.0 - 1.0) < 0.02
        assert rms / v_ref < 0.01
        pres, dt = sim.download_pressure()
        pres = pres / dt
        surface = top + wall
        ys = np.arange(int(wall) + 2, int(surface) - 4)
        want = mats[0].rho * g.gravity * (surface - (ys + 0.5))
        for cx in (W // 4, W // 2, 3 * W // 4):
            assert np.abs(pres[ys, cx] - want).max() <= 0.05 * want.max(), cx
    finally:
        sim.close()


@pytest.mark.parametrize("grid", [(384, 216), (1024, 576)])
def test_divergence_after_projection_on_the_device(grid):
    """rhs.comp:49 after project.comp: the discrete divergence over the pressure unknowns is
    at the solver's tolerance relative to what the solve started from."""
    W, H = grid
    p = Pair(grid)
    try:
        first_substep_prefix(p)
        for s in ("GRID_PREP", "COEF", "SURF", "RHS"):
            p.gpu.stage(s)
        solid = p.field.solid.reshape(H, W, 4)
        unk = p.gpu.field("MGD", np.float32, (p.L.total,))[: W * H].reshape(H, W) > 0
        d0 = discrete_divergence(p.gpu.field("VEL", np.float32, (W * H, 4)).reshape(H, W, 4), solid, W, H)
        p.gpu.stage("PRESSURE")
        p.gpu.stage("PROJECT")
        d1 = discrete_divergence(p.gpu.field("VEL", np.float32, (W * H, 4)).reshape(H, W, 4), solid, W, H)
        # what is left is the solve's true residual, which for an FP32 pressure is bounded
        # below by eps |A| |x| (the FP32 floor), not by the tolerance
        A, unk_a = fp64_ref.pressure_matrix(p.gpu.field("MGA", np.float32, (p.L.total, 4)),
                                            p.gpu.field("MGD", np.float32, (p.L.total,)), W, H)
        x = p.gpu.field("PRES", np.float32, (W * H, 2))[:, 0].astype(np.float64)[unk_a]
        floor = np.finfo(np.float32).eps * np.linalg.norm(abs(A) @ np.abs(x))
        n0, n1 = np.linalg.norm(d0[unk]), np.linalg.norm(d1[unk])
        print(f"{W}x{H}: divergence {n0:.4g} -> {n1:.4g} (FP32 floor {floor:.4g})", flush=True)
        assert n0 > 1.0
        assert n1 <= max(2e-5 * n0, 2 * floor)
    finally:
        p.close()
