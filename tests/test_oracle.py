"""Pins the CPU oracle (oracle/oracle.c) against what fixes its answers without the
reference binary: closed forms, invariants, brute-force linear algebra on small grids, and
the reference's own unit-test values. These run on the CPU (no GPU)."""
import numpy as np
import pytest
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import cases
from sprout_amd import _native as N
from sprout_amd import scene as sc
from oracle.binding import OracleSim


def make(grid, frac=(0.4, 0.6), **kw):
    W, H = grid
    params, field, parts, rows = cases.dam_break(grid, frac=frac, **kw)
    g, mats = N.derive_grid_params(params, W, H, rows)
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return o, g, mats, parts, field


def bspline_ref(r):
    """Quadratic B-spline by distance (the textbook form, not the shader's offset form)."""
    r = abs(r)
    if r < 0.5:
        return 0.75 - r * r
    if r < 1.5:
        return 0.5 * (1.5 - r) ** 2
    return 0.0


# ---- scatter (particles.comp deposit) -------------------------------------------------

def test_scatter_single_particle_matches_textbook_bspline():
    """One particle: every node quantum of every stagger equals 2^20 x the textbook
    B-spline weight of its distance to the node, to rounding. A wrong stagger or base
    index moves the weights by half a cell and fails."""
    W, H = 32, 24
    o, g, mats, parts, field = make((W, H))
    rng = np.random.default_rng(7)
    for trial in range(20):
        p = parts[:1].copy()
        p["pos"] = rng.uniform(8, 20, 2).astype(np.float32)
        p["vel"] = 0
        p["c"] = 0
        o.reset(p)
        o.begin_batch(1 / 240, 1)
        o.stage("STEP")
        o.stage("CLEAR_ACC")
        o.stage("PARTICLES", 0)
        acc = o.field("ACC", np.int32, W * H * g.acc_stride).reshape(H, W, g.acc_stride)
        x, y = float(p["pos"][0][0]), float(p["pos"][0][1])
        rho = mats[0].rho
        for ch, ox, oy, scale in ((0, 0.0, 0.5, rho), (2, 0.5, 0.0, rho), (4, 0.5, 0.5, 1.0)):
            want = np.zeros((H, W))
            for j in range(H):
                for i in range(W):
                    want[j, i] = bspline_ref(x - (i + ox)) * bspline_ref(y - (j + oy)) * scale * 2 ** 20
            got = acc[..., ch].astype(np.float64)
            assert np.abs(got - want).max() <= 1.0, (trial, ch)
            assert 4 <= np.count_nonzero(got) <= 9


def test_scatter_conserves_weight_and_zero_momentum():
    o, g, mats, parts, field = make((64, 36))
    o.begin_batch(1 / 240, 1)
    o.stage("STEP")
    o.stage("CLEAR_ACC")
    o.stage("PARTICLES", 0)
    acc = o.field("ACC", np.int32, 64 * 36 * g.acc_stride).reshape(-1, g.acc_stride).astype(np.int64)
    n = len(parts)
    # partition of unity: sum of the 9 rounded quanta is within 4.5 of 2^20 per particle
    assert abs(acc[:, 4].sum() - n * 2 ** 20) <= 4.5 * n
    assert abs(acc[:, 0].sum() - n * mats[0].rho * 2 ** 20) <= 4.5 * n + 1e-3 * n
    assert np.all(acc[:, 1] == 0) and np.all(acc[:, 3] == 0)


def test_grid_prep_face_velocity_is_momentum_over_mass():
    """Particles moving uniformly (C = 0): every face with enough mass reads that velocity,
    v minus one step of gravity (grid_prep.comp:29-43)."""
    W, H = 48, 32
    o, g, mats, parts, field = make((W, H))
    p = parts.copy()
    p["vel"] = [3.5, -2.0]
    o.reset(p)
    o.begin_batch(1 / 240, 1)
    for s in ("STEP", "CLEAR_ACC"):
        o.stage(s)
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    vel = o.field("VEL", np.float32, (W * H, 4))
    step = o.field("STEP", N.STEP_DTYPE, 1)[0]
    acc = o.field("ACC", np.int32, W * H * g.acc_stride).reshape(-1, g.acc_stride)
    # faces holding at least half a cell of liquid: the 2^22 momentum quantum is then
    # below 2e-4 of the momentum (smaller faces are fixed-point noise by design)
    live_u = (np.abs(vel[:, 2]) < 1e29) & (acc[:, 0] >= 2 ** 19)
    live_v = (np.abs(vel[:, 3]) < 1e29) & (acc[:, 2] >= 2 ** 19)
    assert live_u.sum() > 100 and live_v.sum() > 100
    np.testing.assert_allclose(vel[live_u, 0], 3.5, rtol=2e-4)
    np.testing.assert_allclose(vel[live_v, 1], -2.0 - g.gravity * step["dt"], rtol=2e-4)


def test_gather_reproduces_linear_field():
    """APIC gather of an affine face field (all faces live): the particle takes the field's
    value and C its gradient (particles.comp:165-222)."""
    W, H = 40, 40
    o, g, mats, parts, field = make((W, H))
    a, b, c0, d, e, f0 = 3.0, -1.5, 2.0, 0.75, 2.25, -4.0
    j, i = np.mgrid[0:H, 0:W].astype(np.float64)
    vel = np.zeros((H, W, 4), np.float32)
    vel[..., 0] = a * i + b * (j + 0.5) + c0          # u at (i, j + 0.5)
    vel[..., 1] = d * (i + 0.5) + e * j + f0          # v at (i + 0.5, j)
    o.begin_batch(1e-6, 1)
    o.stage("STEP")
    o.put_field("VEL", vel.reshape(-1, 4))
    comp = np.zeros((H * W, 4), np.float32)
    comp[:, 3] = 1.0  # rest occupancy around every particle
    o.put_field("COMP", comp)
    rng = np.random.default_rng(1)
    p = parts[:64].copy()
    p["pos"] = rng.uniform(12, 28, (64, 2)).astype(np.float32)
    o.reset(p)
    step = o.field("STEP", N.STEP_DTYPE, 1)
    step["dt"] = 1e-6
    o.put_field("STEP", step)
    o.stage("PARTICLES", 1)
    q = o.field("PARTICLES", N.PARTICLE_DTYPE, 64)
    x, y = p["pos"][:, 0].astype(np.float64), p["pos"][:, 1].astype(np.float64)
    np.testing.assert_allclose(q["vel"][:, 0], a * x + b * y + c0, rtol=2e-5, atol=2e-4)
    np.testing.assert_allclose(q["vel"][:, 1], d * x + e * y + f0, rtol=2e-5, atol=2e-4)
    lim = g.apic_clamp / 1e-6
    assert lim > 10
    np.testing.assert_allclose(q["c"][:, 0], a, atol=2e-3)
    np.testing.assert_allclose(q["c"][:, 1], b, atol=2e-3)
    np.testing.assert_allclose(q["c"][:, 2], d, atol=2e-3)
    np.testing.assert_allclose(q["c"][:, 3], e, atol=2e-3)


# ---- pressure assembly and solve --------------------------------------------------------

def assemble(o, W, H):
    """The pressure system up to rhs.comp for the seeded block, first substep."""
    o.begin_batch(1 / 240, 1)
    for s in ("STEP", "CLEAR_ACC"):
        o.stage(s)
    o.stage("PARTICLES", 0)
    for s in ("GRID_PREP", "COEF", "SURF", "RHS"):
        o.stage(s)


def level0_matrix(o, W, H):
    L = N.levels(W, H)
    mga = o.field("MGA", np.float32, (L.total, 4))[: W * H].astype(np.float64)
    mgd = o.field("MGD", np.float32, (L.total,))[: W * H].astype(np.float64)
    return dense_level(mga, mgd, W, H)


def dense_level(mga, mgd, W, H):
    n = W * H
    rows, cols, vals = [], [], []
    unk = mgd > 0
    for k in np.flatnonzero(unk):
        x, y = k % W, k // W
        rows.append(k)
        cols.append(k)
        vals.append(mgd[k])
        for d, (dx, dy) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            nx, ny = x + dx, y + dy
            if 0 <= nx < W and 0 <= ny < H and unk[ny * W + nx]:
                rows.append(k)
                cols.append(ny * W + nx)
                vals.append(-mga[k, d])
    return sps.csr_matrix((vals, (rows, cols)), shape=(n, n)), unk


def test_pressure_matrix_symmetric_and_diagonally_dominant():
    W, H = 64, 36
    o, *_ = make((W, H))
    assemble(o, W, H)
    A, unk = level0_matrix(o, W, H)
    assert unk.sum() > 300
    assert abs(A - A.T).max() == 0.0
    d = A.diagonal()[unk]
    off = np.asarray(abs(A).sum(axis=1)).ravel()[unk] - d
    assert np.all(d >= off * (1 - 1e-6))


def test_galerkin_coarse_levels_are_PtAP():
    W, H = 64, 36
    o, *_ = make((W, H))
    assemble(o, W, H)
    o.stage("MG_BUILD")
    L = N.levels(W, H)
    mga = o.field("MGA", np.float32, (L.total, 4)).astype(np.float64)
    mgd = o.field("MGD", np.float32, (L.total,)).astype(np.float64)
    for l in range(L.n - 1):
        wf, hf = L.dims[l][0], L.dims[l][1]
        wc, hc = L.dims[l + 1][0], L.dims[l + 1][1]
        of, oc = L.offs[l], L.offs[l + 1]
        Af, unk = dense_level(mga[of:of + wf * hf], mgd[of:of + wf * hf], wf, hf)
        Ac, unkc = dense_level(mga[oc:oc + wc * hc], mgd[oc:oc + wc * hc], wc, hc)
        # prolongation: coarse value to its 2 x 2 children that are unknowns
        kf = np.flatnonzero(unk)
        xf, yf = kf % wf, kf // wf
        P = sps.csr_matrix((np.ones(len(kf)), (kf, (yf // 2) * wc + xf // 2)), shape=(wf * hf, wc * hc))
        G = (P.T @ Af @ P).toarray()
        keep = np.diag(G) > 1e-7 * max(1.0, np.abs(G).max())
        Gd = np.where(np.outer(keep, keep), G, 0.0)
        diff = np.abs(Gd - Ac.toarray()).max()
        assert diff <= 1e-4 * max(1.0, np.abs(G).max()), (l, diff)


def vcycle(o, r, L):
    full = np.zeros((L.total, 2), np.float32)
    full[: len(r)] = r
    o.put_field("MGR", full)
    o.stage("VCYCLE")
    return o.field("MGZ", np.float32, (L.total, 2))[: len(r)].astype(np.float64)


def test_vcycle_is_symmetric_positive_definite():
    """CG needs a symmetric positive-definite preconditioner (mg_up.comp:7-10)."""
    W, H = 64, 36
    o, *_ = make((W, H))
    assemble(o, W, H)
    o.stage("MG_BUILD")
    L = N.levels(W, H)
    mgd = o.field("MGD", np.float32, (L.total,))[: W * H]
    unk = mgd > 0
    rng = np.random.default_rng(5)
    for _ in range(3):
        r1 = np.zeros((W * H, 2), np.float32)
        r2 = np.zeros((W * H, 2), np.float32)
        r1[unk] = rng.standard_normal((unk.sum(), 2))
        r2[unk] = rng.standard_normal((unk.sum(), 2))
        z1 = vcycle(o, r1, L)
        z2 = vcycle(o, r2, L)
        for k in range(2):
            a, b = r2[:, k] @ z1[:, k], r1[:, k] @ z2[:, k]
            assert abs(a - b) <= 1e-4 * max(abs(a), abs(b), 1e-30)
            assert r1[:, k] @ z1[:, k] > 0


def test_mgpcg_solves_the_system():
    """The CG iterate satisfies |b - A x| <= tol |b| and agrees with a direct solve."""
    W, H = 64, 36
    o, g, *_ = make((W, H))
    assemble(o, W, H)
    A, unk = level0_matrix(o, W, H)
    b = o.field("PRHS", np.float32, (W * H, 4))[:, :2].astype(np.float64)
    o.stage("PRESSURE")
    x = o.field("PRES", np.float32, (W * H, 2)).astype(np.float64)
    K = o.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
    assert K["done"].tolist() == [1, 1]
    idx = np.flatnonzero(unk)
    Au = A[idx][:, idx]
    for k, tol in ((0, g.tol[0]), (1, g.tol[1])):
        bu, xu = b[idx, k], x[idx, k]
        res = np.linalg.norm(bu - Au @ xu) / np.linalg.norm(bu)
        # the recursive residual the solve stops on (reduce_tail.glsl:94) is below tol; the
        # true residual of the FP32-stored iterate is bounded by that plus the FP32 floor
        # eps |A| |x| / |b| of representing x at all
        assert np.sqrt(K["rr"][k] / K["bb"][k]) <= tol
        floor = 2.0 ** -24 * np.linalg.norm(abs(Au) @ abs(xu)) / np.linalg.norm(bu)
        assert res <= 3 * tol + 10 * floor, (k, res, floor)
        if k == 0:
            xd = spla.spsolve(Au.tocsc(), bu)
            assert np.linalg.norm(xu - xd) / np.linalg.norm(xd) <= 1e-3


def discrete_divergence(vel, solid, W, H):
    """rhs.comp:49: aR u(i+1) - aL u(i) + aT v(i+W) - aB v(i), per cell."""
    u = vel[..., 0].astype(np.float64)
    v = vel[..., 1].astype(np.float64)
    aL, aB = solid[..., 0].astype(np.float64), solid[..., 1].astype(np.float64)
    div = np.zeros((H, W))
    div[1:-1, 1:-1] = (aL[1:-1, 2:] * u[1:-1, 2:] - aL[1:-1, 1:-1] * u[1:-1, 1:-1]
                       + aB[2:, 1:-1] * v[2:, 1:-1] - aB[1:-1, 1:-1] * v[1:-1, 1:-1])
    return div


def test_divergence_after_projection():
    """After project.comp the discrete divergence of every pressure unknown (rhs.comp:49)
    is at the solver's tolerance relative to the divergence it started from."""
    W, H = 64, 36
    o, g, mats, parts, field = make((W, H))
    assemble(o, W, H)
    L = N.levels(W, H)
    unk = o.field("MGD", np.float32, (L.total,))[: W * H].reshape(H, W) > 0
    solid = field.solid.reshape(H, W, 4)
    d0 = discrete_divergence(o.field("VEL", np.float32, (W * H, 4)).reshape(H, W, 4), solid, W, H)
    o.stage("PRESSURE")
    o.stage("PROJECT")
    d1 = discrete_divergence(o.field("VEL", np.float32, (W * H, 4)).reshape(H, W, 4), solid, W, H)
    assert np.linalg.norm(d0[unk]) > 1.0
    assert np.linalg.norm(d1[unk]) <= 2e-5 * np.linalg.norm(d0[unk])


def test_hydrostatic_pool():
    """A resting pool (validate.rs:125-141): after half a second the pressure below the
    surface is rho g depth (p dt / dt in grid units), the liquid stays near rest."""
    W, H = 48, 40
    o, g, mats, parts, field = make((W, H), frac=(1.0, 0.5))
    steps = 120
    o.begin_batch(steps / 240, steps)
    for k in range(steps):
        o.substep(1 if k else 0)
    p = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    assert np.sqrt((p["vel"] ** 2).sum(1)).max() < 0.05 * np.sqrt(g.gravity * 17)
    step = o.field("STEP", N.STEP_DTYPE, 1)[0]
    pres = o.field("PRES", np.float32, (W * H, 2))[:, 0].reshape(H, W) / step["dt"]
    top = p["pos"][:, 1].max()
    cx = W // 2
    ys = np.arange(5, int(top) - 4)
    depth = top - (ys + 0.5)
    want = mats[0].rho * g.gravity * depth
    got = pres[ys, cx]
    assert np.abs(got - want).max() <= 0.05 * want.max()


# ---- bookkeeping kernels ----------------------------------------------------------------

def test_sort_is_stable_by_cell():
    o, g, mats, parts, field = make((64, 36))
    o.begin_batch(5 / 240, 5)
    for k in range(5):
        o.substep(1 if k else 0)
    before = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    rng = np.random.default_rng(2)
    before = before[rng.permutation(len(before))]
    o.put_field("PARTICLES", before)
    o.sort()
    after = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
    cell = np.clip(before["pos"][:, 1].astype(np.int64), 0, 35) * 64 + np.clip(before["pos"][:, 0].astype(np.int64), 0, 63)
    order = np.argsort(cell, kind="stable")
    assert np.array_equal(after["id"], before["id"][order])


def test_spawn_fills_fresh_then_dead_slots_in_order():
    """The record's spawn step (mod.rs:2485-2503 + claim.comp:10-17): fresh slots take the
    first pending particles, then every dead slot, in slot order, takes the next one; no live
    particle is overwritten and what finds no slot is dropped."""
    n0, cap = 200, 230
    o = OracleSim((32, 24), cap)
    rng = np.random.default_rng(5)
    slots = np.zeros(n0, N.PARTICLE_DTYPE)
    slots["kind"] = 1
    slots["id"] = np.arange(n0)
    dead = np.sort(rng.choice(n0, 40, replace=False))
    slots["kind"][dead] = N.KIND_DEAD
    o.put_field("PARTICLES", slots)
    for m in (10, 30 + 25, 30 + 40 + 9):
        o.put_field("PARTICLES", slots)
        sp = np.zeros(m, N.PARTICLE_DTYPE)
        sp["kind"] = 2
        sp["id"] = 1000 + np.arange(m)
        o.spawn(sp)
        got = o.field("PARTICLES", N.PARTICLE_DTYPE, o.n_slots())
        fresh = min(m, cap - n0)
        assert o.n_slots() == n0 + fresh
        assert np.array_equal(got["id"][n0:], 1000 + np.arange(fresh))
        live = np.setdiff1d(np.arange(n0), dead)
        assert np.array_equal(got["id"][live], live)  # live particles untouched
        claimed = min(m - fresh, len(dead))
        assert np.array_equal(got["id"][dead[:claimed]], 1000 + fresh + np.arange(claimed))
        assert np.all(got["kind"][dead[claimed:]] == N.KIND_DEAD)


def test_cell_counts_and_stats():
    o, g, mats, parts, field = make((64, 36))
    got = o.cell_counts(3, 3, 30, 20).reshape(20, 30)
    x = np.floor(parts["pos"][:, 0]).astype(int) - 3
    y = np.floor(parts["pos"][:, 1]).astype(int) - 3
    m = (x >= 0) & (y >= 0) & (x < 30) & (y < 20)
    want = np.zeros((20, 30), np.uint32)
    np.add.at(want, (y[m], x[m]), 1)
    assert np.array_equal(got, want)
    live, count, conc = o.stats()
    assert live == len(parts) and count[0] == len(parts) and conc[0] == 256 * len(parts)


def test_step_time_is_compensated():
    o, g, *_ = make((32, 24))
    n = 1000
    o.begin_batch(n / 240, n)
    for _ in range(n):
        o.stage("STEP")
    s = o.field("STEP", N.STEP_DTYPE, 1)[0]
    t = float(s["advanced"]) - float(np.array(s["pad0"], np.uint32).view(np.float32))
    assert abs(t - n * float(np.float32(1 / 240))) < 1e-6
    assert s["substeps"] == n


# ---- host-side tables (params.rs tests) ---------------------------------------------------

def test_grid_units_follow_the_cell_size():
    """params.rs:1423-1437."""
    p = N.default_params()
    g, mats = N.derive_grid_params(p, 64, 64, [0])
    assert abs(g.gravity - 981.0) < 1e-3
    assert abs(np.exp(mats[0].ln_nu) - 0.01) < 1e-6
    assert abs(mats[0].sigma - 72.0 / 0.998) < 0.01
    p.cell_size = 0.001
    g2, mats2 = N.derive_grid_params(p, 64, 64, [0])
    assert abs(g2.gravity / 981.0 - 10.0) < 1e-3
    assert abs(mats2[0].sigma / mats[0].sigma - 1000.0) < 0.1


def test_tank_layout_channels():
    """Tank::layout (mod.rs:271-331): water first (remainder), honey as the primary family
    with a concentration channel, oil and mercury with weight channels."""
    p = N.default_params()
    g, mats = N.derive_grid_params(p, 64, 64, [N.ROW_HONEY, N.ROW_OIL, N.ROW_MERCURY])
    assert list(g.present[: g.n_present]) == [0, 1, 2, 8]
    assert (mats[0].chan_w, mats[1].chan_w, mats[1].chan_s, mats[1].fam) == (-1, -1, 5, 0)
    assert (mats[2].chan_w, mats[8].chan_w) == (6, 7)
    assert g.acc_stride == 8 and g.n_fam == 1 and g.primary == 1
    assert g.group_mask == 0b111


def test_levels_match_the_reference_hierarchy():
    """Levels::new (mod.rs:727-753) at 1024 x 576: ten levels, the tail from 32 x 18."""
    L = N.levels(1024, 576)
    dims = [tuple(L.dims[k]) for k in range(L.n)]
    assert dims == [(1024, 576), (512, 288), (256, 144), (128, 72), (64, 36), (32, 18), (16, 9), (8, 5), (4, 3), (2, 2)]
    assert L.tail == 5


def test_fill_rect_count_and_walls():
    """maps.rs:557-563: a honey fill of a 64 x 64 tank up to y = 23 holds 58 x 20 x 4."""
    scene = sc.Scene((64, 64))
    out = sc.fill_rect(scene, (0.0, 0.0, 64.0, 23.0), N.ROW_HONEY)
    assert len(out) == 58 * 20 * 4
    assert np.all(out["s"] == N.SPAWN_S)
    assert np.all((out["pos"][:, 0] > sc.WALL) & (out["pos"][:, 0] < 64 - sc.WALL) & (out["pos"][:, 1] > sc.WALL))


def test_rng_stream_is_the_xorshift_sequence():
    """The lane-parallel stream (jumps by a GF(2) matrix power) equals xorshift32 stepped one
    value at a time (maps.rs:456-465), across lane boundaries, and maps.rs's first values
    from the seed of the first fill."""
    def scalar(seed, n):
        r = sc.Rng(seed)
        return np.array([r.next() for _ in range(n)], np.float32)

    for seed, n in [(0x9E3779B9 ^ 1, 3), (12345, 4095), (0xDEADBEEF, 4097), (0x9E3779B9 ^ (7 * 747796405 + 1), 12289)]:
        assert np.array_equal(sc._rng_stream(seed, n), scalar(seed, n)), (seed, n)
    # xorshift32 from 1: 270369, 67634689, 2647435461 (Marsaglia 2003), then >> 8 / 2^24
    want = np.array([270369 >> 8, 67634689 >> 8, 2647435461 >> 8], np.float64) / (1 << 24)
    assert np.array_equal(sc._rng_stream(1, 3), want.astype(np.float32))


def test_config1_dam_break_count():
    """384 x 216, block x < 0.4 W, y < 0.6 H with the 3-cell frame: 301 x 253 particles."""
    scene, parts = sc.dam_break((384, 216))
    assert len(parts) == 301 * 253


def test_rasterize_frame():
    f = sc.Scene((40, 30)).rasterize()
    solid = f.solid.reshape(30, 40, 4)
    assert np.all(solid[:2, :, 3] == 0) and np.all(solid[5:25, 5:35, 3] == 1)
    assert np.all(solid[5:25, 5:35, :2] == 1)
    region = f.region.reshape(30, 40)
    assert np.count_nonzero(region & sc.REGION_VENT) == 1


def test_grid_tables_scale_with_their_units():
    """The grid-unit tables (gpu_params / gpu_materials, mod.rs:2179-2330) by dimensional
    analysis rather than re-derivation: with cells a quarter the size, gravity (cells/s^2)
    grows 4x, kinematic viscosity (cells^2/s) 16x, tension over water's density (cells^3/s^2)
    64x, a yield stress or shear modulus over density (cells^2/s^2) 16x, a grain diameter
    (cells) 4x and a conductivity (cells^2/s) 16x, while relative densities and heat
    capacities stay; a liquid 1.5x as dense has 1.5x the relative density and volumetric heat
    capacity; a viscosity 10x as large adds ln 10 to ln nu; the tolerances and step factors pass
    through unchanged."""
    rows = [N.ROW_WATER, N.ROW_HONEY, N.ROW_OIL, N.ROW_MERCURY]

    def tables(dx, edit=None):
        p = N.default_params()
        p.cell_size = dx
        h = p.mat[N.ROW_HONEY]
        h.yield_stress, h.shear_modulus, h.relax_time, h.grain = 3.0, 40.0, 0.5, 2e-4
        for (row, field), v in (edit or {}).items():
            setattr(p.mat[row], field, v)
        return p, *N.derive_grid_params(p, 96, 64, rows)

    p, g, m = tables(0.004)
    _, g4, m4 = tables(0.001)
    assert g4.gravity / g.gravity == pytest.approx(4.0, rel=1e-6)
    assert g.gravity == pytest.approx(9.81 / 0.004, rel=1e-6)
    for r in rows:
        assert np.exp(m4[r].ln_nu - m[r].ln_nu) == pytest.approx(16.0, rel=1e-5)
        assert m4[r].sigma / m[r].sigma == pytest.approx(64.0, rel=1e-5)
        assert m4[r].th[1] / m[r].th[1] == pytest.approx(16.0, rel=1e-5)
        assert m4[r].rho == m[r].rho and m4[r].th[0] == m[r].th[0]
        assert (m4[r].cos_c, m4[r].sin_c) == (m[r].cos_c, m[r].sin_c)
    h = N.ROW_HONEY
    assert m4[h].rheo_p[0] / m[h].rheo_p[0] == pytest.approx(16.0, rel=1e-5)
    assert m4[h].gr[1] / m[h].gr[1] == pytest.approx(16.0, rel=1e-5)
    assert m4[h].grain / m[h].grain == pytest.approx(4.0, rel=1e-5)
    assert m[h].grain == pytest.approx(2e-4 / 0.004, rel=1e-6)
    # absolute values at 4 mm cells (SI over the cell): tension over water's density, yield
    # stress and modulus over water's density
    assert m[N.ROW_WATER].sigma == pytest.approx(p.mat[N.ROW_WATER].tension / (998.0 * 0.004 ** 3), rel=1e-5)
    assert m[h].rheo_p[0] == pytest.approx(3.0 / (998.0 * 0.004 ** 2), rel=1e-5)
    assert m[h].gr[1] == pytest.approx(40.0 / (998.0 * 0.004 ** 2), rel=1e-5)
    o = N.ROW_OIL
    _, _, md = tables(0.004, {(o, "density"): 1.5 * p.mat[o].density})
    assert md[o].rho / m[o].rho == pytest.approx(1.5, rel=1e-6)
    assert md[o].th[0] / m[o].th[0] == pytest.approx(1.5, rel=1e-6)
    _, _, mv = tables(0.004, {(o, "viscosity"): 10.0 * p.mat[o].viscosity})
    assert mv[o].ln_nu - m[o].ln_nu == pytest.approx(np.log(10.0), abs=2e-6)
    assert (g.tol[0], g.tol[1], g.tol[2]) == (np.float32(p.pressure_tol), np.float32(p.correction_tol),
                                               np.float32(p.viscosity_tol))
    assert g.cfl == np.float32(p.cfl) and g.sep_tau == np.float32(p.droplet_cells)


def test_capillary_step_limit_is_the_capillary_time():
    """The longest substep a tank allows (dt_capillary, params.rs:962-982) is half the capillary
    time sqrt(rho dx^3 / (2 pi sigma)) of its worst surface (largest sigma / rho, a pair's
    interfacial tension over the lighter liquid), in SI, where that is below dt_max."""
    p = N.default_params()
    p.cell_size = 5e-4
    g, _ = N.derive_grid_params(p, 64, 64, [N.ROW_WATER])
    w = p.mat[N.ROW_WATER]
    cap = 0.5 * np.sqrt(w.density * p.cell_size ** 3 / (2 * np.pi * w.tension))
    assert cap < p.dt_max
    assert g.dt_max == pytest.approx(cap, rel=1e-5)
    # water and oil: the pair's tension over the oil's density may be the worst surface
    g2, _ = N.derive_grid_params(p, 64, 64, [N.ROW_WATER, N.ROW_OIL])
    o = p.mat[N.ROW_OIL]
    worst = max(w.tension / w.density, o.tension / o.density,
                p.pair_tension[0][o.group] / min(w.density, o.density))
    assert g2.dt_max == pytest.approx(0.5 * np.sqrt(p.cell_size ** 3 / (2 * np.pi * worst)), rel=1e-5)
