"""Pins the oracle's implicit viscosity solve (visc_*.comp, vmg*.comp, rhs_visc.comp,
project_visc.comp restated in oracle/oracle.c) against what fixes it without the reference:

* the system's definition (visc.glsl:1-18): backward Euler on the full stress tensor,
  A = M + G^T W G with G the strain samples (D_xx, D_yy at cell centres, D_xy at nodes with
  mirrored no-slip faces), assembled here as a sparse matrix and compared with the matrix
  the oracle stores for its multigrid (visc_diag.comp couplings) and its Jacobi diagonal;
* Galerkin coarsening P^T A P by brute force (vmg_build.comp);
* a symmetric positive-definite V-cycle (vmg.glsl:13-21);
* the CG result against a direct sparse solve of the same system;
* the strain of a rigid rotation is zero (the stress tensor does not damp rotation);
* after the second projection (project_visc.comp) the velocity is divergence-free to the
  solver tolerance, and a resting viscous pool stays at rest (mod.rs:2671-2681).
These run on the CPU."""
import numpy as np
import scipy.sparse as sps
import scipy.sparse.linalg as spla

import cases
from sprout_amd import _native as N
from oracle.binding import OracleSim


def make_viscous(grid, frac=(0.5, 0.6), viscosity=1e-2, viscosity_tol=1e-5, **kw):
    W, H = grid
    params, field, parts, rows = cases.dam_break(grid, frac=frac, viscosity=viscosity)
    params.viscosity_tol = viscosity_tol
    for k, v in kw.items():
        setattr(params, k, v)
    g, mats = N.derive_grid_params(params, W, H, rows)
    assert g.visc_mode == 2
    g.n_slots = len(parts)
    o = OracleSim(grid, W * H * 4)
    o.set_gparams(g, mats)
    o.set_solids(field)
    o.reset(parts)
    return o, g, params


def to_split(o, vel_xy=None):
    """The first substep up to the viscosity split; vel_xy replaces the face velocities."""
    o.begin_batch(1 / 240, 1)
    for s in ("STEP", "CLEAR_ACC"):
        o.stage(s)
    o.stage("PARTICLES", 0)
    o.stage("GRID_PREP")
    if vel_xy is not None:
        W, H = o.grid
        vel = o.field("VEL", np.float32, (W * H, 4))
        vel[:, :2] = vel_xy
        o.put_field("VEL", vel)
    o.stage("COEF")
    o.stage("SURF")
    o.stage("VISC_SPLIT")


def faces(o):
    W, H = o.grid
    vel = o.field("VEL", np.float32, (W * H, 4))
    st = vel[:, 2:4]
    live = np.abs(st) < 1e29
    closed = st < -1e29
    return vel, live, closed


def active_cells(o):
    W, H = o.grid
    tiles = o.field("VTILES", np.uint32, ((W + 15) // 16) * ((H + 15) // 16))
    x, y = np.meshgrid(np.arange(W), np.arange(H))
    on = tiles[(y >> 4) * ((W + 15) // 16) + (x >> 4)].ravel() != 0
    interior = ((x >= 1) & (y >= 1) & (x < W - 1) & (y < H - 1)).ravel()
    return on, interior


def strain_operator(o, vw):
    """G (samples x 2N faces, u faces first) and the sample weights, from the definition:
    D_xx = u(i+1) - u(i), D_yy = v(i+W) - v(i) at cells with a centre weight; D_xy =
    ((u_a - u_b) + (v_a - v_b)) / 2 at nodes with a node weight, a closed face taking minus its
    live partner (no-slip wall half a cell away). Faces that are not live contribute nothing."""
    W, H = o.grid
    n = W * H
    _, live, closed = faces(o)
    rows, cols, vals, wts = [], [], [], []
    r = 0

    def add(row, comp, i, s):
        if 0 <= i < n and live[i, comp]:
            rows.append(row)
            cols.append(comp * n + i)
            vals.append(s)

    for i in np.flatnonzero(vw[:, 0] > 0):
        add(r, 0, i + 1, 1.0)
        add(r, 0, i, -1.0)
        wts.append(vw[i, 0])
        r += 1
        add(r, 1, i + W, 1.0)
        add(r, 1, i, -1.0)
        wts.append(vw[i, 0])
        r += 1
    for i in np.flatnonzero(vw[:, 1] > 0):
        for comp, a, b in ((0, i, i - W), (1, i, i - 1)):
            for f, partner, s in ((a, b, 0.5), (b, a, -0.5)):
                if live[f, comp]:
                    add(r, comp, f, s)
                elif closed[f, comp] and live[partner, comp]:
                    add(r, comp, partner, -s)
        wts.append(vw[i, 1])
        r += 1
    G = sps.csr_matrix((vals, (rows, cols)), shape=(r, 2 * n))
    return G, np.array(wts, np.float64)


def unknown_faces(o):
    W, H = o.grid
    n = W * H
    _, live, _ = faces(o)
    on, interior = active_cells(o)
    unk = np.concatenate([live[:, 0] & on & interior, live[:, 1] & on & interior])
    return unk


def definition_matrix(o):
    W, H = o.grid
    n = W * H
    vw = o.field("VW", np.float32, (n, 4)).astype(np.float64)
    coef = o.field("COEF", np.float32, (n, 4)).astype(np.float64)
    G, w = strain_operator(o, vw)
    M = sps.diags(np.concatenate([coef[:, 2], coef[:, 3]]))
    K = (G.T @ sps.diags(w) @ G).tocsr()
    return M, K


def stored_matrix(vma, vmb, vmd, Wl, Hl):
    """The level matrix as the multigrid stores it: vma (u-u left, u-u below, v-v left,
    v-v below), vmb (u to v of the cell, of the left cell, of the cell above, of the upper-left
    cell), vmd the diagonals; each coupling stored once."""
    n = Wl * Hl
    rows, cols, vals = [], [], []

    def put(a, b, v):
        if v != 0.0:
            rows.extend((a, b))
            cols.extend((b, a))
            vals.extend((v, v))

    for k in range(n):
        x, y = k % Wl, k // Wl
        a, b, d = vma[k], vmb[k], vmd[k]
        if d[0] != 0.0:
            rows.append(k), cols.append(k), vals.append(d[0])
        if d[1] != 0.0:
            rows.append(n + k), cols.append(n + k), vals.append(d[1])
        if x > 0:
            put(k, k - 1, a[0])
            put(n + k, n + k - 1, a[2])
            put(k, n + k - 1, b[1])
        if y > 0:
            put(k, k - Wl, a[1])
            put(n + k, n + k - Wl, a[3])
        put(k, n + k, b[0])
        if y < Hl - 1:
            put(k, n + k + Wl, b[2])
            if x > 0:
                put(k, n + k + Wl - 1, b[3])
    return sps.csr_matrix((vals, (rows, cols)), shape=(2 * n, 2 * n))


def level_fields(o, l):
    L = N.levels(*o.grid)
    off, Wl, Hl = L.offs[l], L.dims[l][0], L.dims[l][1]
    vma = o.field("VMA", np.float32, (L.total, 4))[off:off + Wl * Hl].astype(np.float64)
    vmb = o.field("VMB", np.float32, (L.total, 4))[off:off + Wl * Hl].astype(np.float64)
    vmd = o.field("VMD", np.float32, (L.total, 2))[off:off + Wl * Hl].astype(np.float64)
    return vma, vmb, vmd, Wl, Hl


def random_flow(W, H, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((W * H, 2)).astype(np.float32) * 20.0


GRID = (64, 36)


def test_viscosity_matrix_is_mass_plus_stress_operator():
    """The multigrid's level-0 matrix and the Jacobi diagonal equal M + G^T W G on the
    unknown faces; the matrix is symmetric and has no coupling to any other face."""
    o, g, _ = make_viscous(GRID)
    W, H = GRID
    to_split(o, random_flow(W, H, 1))
    unk = unknown_faces(o)
    assert unk.sum() > 500
    M, K = definition_matrix(o)
    A_def = (M + K).tocsr()[unk][:, unk].toarray()
    vma, vmb, vmd, _, _ = level_fields(o, 0)
    A_st = stored_matrix(vma, vmb, vmd, W, H)
    assert abs(A_st - A_st.T).max() == 0.0
    A_all = A_st.toarray()
    A_u = A_all[unk][:, unk]
    scale = np.abs(A_def).max()
    assert np.abs(A_u - A_def).max() <= 1e-6 * scale
    # nothing couples an unknown face to a face that is not one
    assert np.abs(A_all[unk][:, ~unk]).max() == 0.0
    vw = o.field("VW", np.float32, (W * H, 4)).astype(np.float64)
    diag = np.concatenate([vw[:, 2], vw[:, 3]])[unk]
    np.testing.assert_allclose(diag, np.diag(A_def), rtol=1e-6)
    assert np.all(np.linalg.eigvalsh(A_def) > 0)


def test_viscosity_galerkin_coarse_level_is_PtAP():
    o, g, _ = make_viscous(GRID)
    W, H = GRID
    to_split(o, random_flow(W, H, 2))
    vma, vmb, vmd, _, _ = level_fields(o, 0)
    A0 = stored_matrix(vma, vmb, vmd, W, H)
    cma, cmb, cmd, Wc, Hc = level_fields(o, 1)
    A1 = stored_matrix(cma, cmb, cmd, Wc, Hc).toarray()
    n0, n1 = W * H, Wc * Hc
    k = np.arange(n0)
    parent = (k // W // 2) * Wc + (k % W) // 2
    P = sps.csr_matrix((np.ones(2 * n0), (np.concatenate([k, n0 + k]), np.concatenate([parent, n1 + parent]))),
                       shape=(2 * n0, 2 * n1))
    G = (P.T @ A0 @ P).toarray()
    assert np.abs(G).max() > 0
    assert np.abs(G - A1).max() <= 1e-5 * np.abs(G).max()


def vcycle(o, r):
    W, H = o.grid
    o.put_field("VR", r.astype(np.float32))
    o.stage("VISC_VCYCLE")
    L = N.levels(W, H)
    return o.field("VMZ", np.float32, (L.total, 2))[: W * H].astype(np.float64)


def test_viscosity_vcycle_is_symmetric_positive_definite():
    """Conjugate gradients needs a symmetric positive-definite preconditioner: pre-smoothing
    red then black, post-smoothing the mirror image (vmg.glsl:13-21)."""
    o, g, _ = make_viscous(GRID)
    W, H = GRID
    to_split(o, random_flow(W, H, 3))
    unk = unknown_faces(o).reshape(2, -1).T
    rng = np.random.default_rng(4)
    for _ in range(3):
        r1 = np.where(unk, rng.standard_normal((W * H, 2)), 0.0)
        r2 = np.where(unk, rng.standard_normal((W * H, 2)), 0.0)
        z1 = vcycle(o, r1)
        z2 = vcycle(o, r2)
        a, b = np.sum(r2 * z1), np.sum(r1 * z2)
        assert abs(a - b) <= 1e-4 * max(abs(a), abs(b))
        assert np.sum(r1 * z1) > 0
        assert np.all(z1[~unk] == 0)


def solve_check(o, W, H):
    n = W * H
    vel0, live, _ = faces(o)
    ustar = np.concatenate([np.where(live[:, 0], vel0[:, 0], 0.0), np.where(live[:, 1], vel0[:, 1], 0.0)])
    unk = unknown_faces(o)
    M, K = definition_matrix(o)
    A = (M + K).tocsr()
    Auu = A[unk][:, unk]
    known = ~unk
    b = M.diagonal()[unk] * ustar[unk] - K[unk][:, known] @ ustar[known]
    x_direct = spla.spsolve(Auu.tocsc(), b)
    o.stage("VISC_SOLVE", 0)
    vel1, _, _ = faces(o)
    x = np.concatenate([vel1[:, 0], vel1[:, 1]]).astype(np.float64)
    r0 = b - Auu @ ustar[unk]
    r = b - Auu @ x[unk]
    return x, x_direct, ustar, unk, r, r0, vel0, vel1


def test_viscosity_cg_matches_direct_solve():
    """With the multigrid preconditioner: the new face velocities solve (M + G^T W G) x =
    M u* with the non-unknown faces held at u*, to the solver tolerance, and agree with a
    direct solve. Faces that are not unknowns keep their velocity."""
    W, H = GRID
    o, g, _ = make_viscous(GRID)
    to_split(o, random_flow(W, H, 5))
    x, x_direct, ustar, unk, r, r0, vel0, vel1 = solve_check(o, W, H)
    K = o.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
    assert K["v_done"] == 1 and K["v_iters"] > 2
    tol = float(g.tol[2])
    assert np.linalg.norm(r) <= 3 * tol * np.linalg.norm(r0) + 1e-6 * np.linalg.norm(ustar)
    err = np.linalg.norm(x[unk] - x_direct) / np.linalg.norm(x_direct)
    assert err < 1e-3, err
    np.testing.assert_array_equal(np.concatenate([vel1[:, 0], vel1[:, 1]])[~unk],
                                  np.concatenate([vel0[:, 0], vel0[:, 1]])[~unk])


def test_viscosity_jacobi_cg_matches_direct_solve():
    """The same system with the Jacobi preconditioner (viscosity_precond 0)."""
    W, H = GRID
    o, g, _ = make_viscous(GRID, viscosity_precond=0)
    to_split(o, random_flow(W, H, 6))
    x, x_direct, ustar, unk, r, r0, _, _ = solve_check(o, W, H)
    K = o.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
    assert K["v_done"] == 1
    tol = float(g.tol[2])
    assert np.linalg.norm(r) <= 3 * tol * np.linalg.norm(r0) + 1e-6 * np.linalg.norm(ustar)
    assert np.linalg.norm(x[unk] - x_direct) / np.linalg.norm(x_direct) < 1e-3


def test_rigid_rotation_has_no_strain():
    """D = (grad u + grad u^T) / 2 of u = omega (-(y - yc), x - xc) vanishes wherever every
    face of a strain sample is live (a swapped sign in D_xy gives -2 omega there)."""
    W, H = GRID
    o, g, _ = make_viscous(GRID, frac=(1.0, 1.0))
    omega, xc, yc = 3.0, 30.0, 17.0
    x, y = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    u = -omega * (y + 0.5 - yc)
    v = omega * (x + 0.5 - xc)
    to_split(o, np.stack([u.ravel(), v.ravel()], -1).astype(np.float32))
    _, live, _ = faces(o)
    vw = o.field("VW", np.float32, (W * H, 4))
    vs = o.field("VS", np.float32, (W * H, 4)).astype(np.float64)
    i = np.arange(W * H)
    inner = (i % W > 1) & (i % W < W - 2) & (i // W > 1) & (i // W < H - 2)
    cells = inner & (vw[:, 0] > 0) & live[np.minimum(i + 1, W * H - 1), 0] & live[:, 0] & live[:, 1] \
        & live[np.minimum(i + W, W * H - 1), 1]
    nodes = inner & (vw[:, 1] > 0) & live[:, 0] & live[i - W, 0] & live[:, 1] & live[i - 1, 1]
    assert cells.sum() > 200 and nodes.sum() > 200
    mag = omega * max(W, H)
    assert np.abs(vs[cells, :2]).max() <= 1e-5 * mag * vw[cells, 0].max()
    assert np.abs(vs[nodes, 2]).max() <= 1e-5 * mag * vw[nodes, 1].max()


def run_substeps(o, n):
    o.begin_batch(n / 240, n)
    for k in range(n):
        o.substep(0 if k == 0 else 1)


def test_second_projection_leaves_no_divergence():
    """After the viscosity and the second pressure solve (rhs_visc.comp, project_visc.comp)
    the discrete divergence of every pressure unknown (rhs.comp:49) is the second solve's
    residual: at most its tolerance times its right-hand side, to FP32 rounding."""
    W, H = GRID
    o, g, _ = make_viscous(GRID)
    run_substeps(o, 3)
    K = o.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
    assert K["p2_unconv"] == 0 and K["v_unconv"] == 0 and K["p2_iters_last"] > 0
    L = N.levels(W, H)
    mgd = o.field("MGD", np.float32, (L.total,))[: W * H]
    vel = o.field("VEL", np.float32, (W * H, 4))[:, :2].astype(np.float64)
    _, field, _, _ = cases.dam_break(GRID, frac=(0.5, 0.6), viscosity=1e-2)
    sol = np.asarray(field.solid, np.float64).reshape(W * H, 4)
    i = np.flatnonzero(mgd > 0)
    div = sol[i + 1, 0] * vel[i + 1, 0] - sol[i, 0] * vel[i, 0] + sol[i + W, 1] * vel[i + W, 1] - sol[i, 1] * vel[i, 1]
    b = o.field("PRHS", np.float32, (W * H, 4))[i, 0].astype(np.float64)
    scale = np.abs(vel).max()
    assert np.linalg.norm(div) <= 3 * float(g.tol[0]) * np.linalg.norm(b) + 2e-6 * scale * np.sqrt(len(i)), \
        (np.linalg.norm(div), np.linalg.norm(b))


def pool_speeds(projection, n=12):
    """RMS and largest face speed below the surface rows of a resting viscous layer."""
    W, H = GRID
    o, g, _ = make_viscous(GRID, frac=(1.0, 0.5), viscous_projection=projection)
    run_substeps(o, n)
    vel, live, _ = faces(o)
    cell = o.field("CELL", np.float32, (W * H, 4))
    liquid = cell[:, 3] > 0
    ys = np.arange(W * H) // W
    inner = liquid & (ys < ys[liquid].max() - 2)
    u = np.concatenate([vel[inner & live[:, 0], 0], vel[inner & live[:, 1], 1]]).astype(np.float64)
    return np.sqrt(np.mean(u * u)), np.abs(u).max(), float(g.gravity) / 240.0


def test_resting_viscous_pool_stays_at_rest():
    """A viscous layer at rest keeps still when the first pressure solve takes out gravity
    before the viscosity sees the velocity; with one solve after the viscosity the no-slip
    walls hold back the pull of gravity and the pool circulates (mod.rs:2671-2681)."""
    rms2, max2, gdt = pool_speeds(1)
    rms1, max1, _ = pool_speeds(0)
    assert max2 < 0.05 * gdt, (max2, gdt)
    assert rms2 < 0.1 * rms1, (rms2, rms1)


def test_fp64_matrix_free_viscosity_system_is_the_definition():
    """The FP64 matrix-free viscosity system the GPU tests compare each device solve with
    (fp64_ref.ViscSystem) is the definition's A = M + G^T W G on the unknown faces, and its
    right-hand side from the device's strain field is -(G^T W G u*) there; the oracle's
    solve agrees with its FP64 solution to the solver tolerance."""
    import fp64_ref

    W, H = GRID
    o, g, _ = make_viscous(GRID)
    to_split(o, random_flow(W, H, 7))
    n = W * H
    vel0, live, _ = faces(o)
    sysm = fp64_ref.ViscSystem(vel0, o.field("VW", np.float32, (n, 4)), o.field("COEF", np.float32, (n, 4)),
                               o.field("VTILES", np.uint32, ((W + 15) // 16) * ((H + 15) // 16)), W, H)
    unk = unknown_faces(o)
    assert np.array_equal(sysm.unk, unk)
    M, K = definition_matrix(o)
    A = (M + K).tocsr()[unk][:, unk]
    rng = np.random.default_rng(0)
    for _ in range(3):
        x = rng.standard_normal(int(unk.sum()))
        np.testing.assert_allclose(sysm.apply(x), A @ x, rtol=1e-12, atol=1e-9 * np.abs(A @ x).max())
    np.testing.assert_allclose(sysm.diag, A.diagonal(), rtol=1e-6)
    ustar = np.concatenate([np.where(live[:, 0], vel0[:, 0], 0.0), np.where(live[:, 1], vel0[:, 1], 0.0)])
    b = sysm.rhs(o.field("VS", np.float32, (n, 4)))
    want = -(K @ ustar)[unk]
    assert np.linalg.norm(b - want) <= 1e-6 * np.linalg.norm(want)
    o.stage("VISC_SOLVE", 0)
    e = fp64_ref.viscosity_errors(sysm, o.field("VS", np.float32, (n, 4)), o.field("VEL", np.float32, (n, 4)),
                                  float(g.tol[2]))
    assert e["true_residual"] <= 3 * float(g.tol[2]) + 1e-6
    # at least as close as an FP64 CG stopped at the same tolerance
    assert e["err_vs_fp64"] <= max(e["fp64_cg_same_tol_err"], 1e-6), e
    assert e["vel_err_vs_fp64"] <= max(e["fp64_cg_same_tol_vel_err"], 1e-6), e
