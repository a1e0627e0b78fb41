"""FP64 references for the pressure solve: the level-0 system the device solves, assembled
from the device's own matrix (mga couplings L, R, B, T and mgd diagonals, rhs.comp's
layout: (A x)_i = D_i x_i - sum_d w_d x_nb(d) over neighbours that are unknowns), solved
exactly (sparse LU) and by an FP64 conjugate-gradient at a given relative tolerance.

Used by the GPU tests and the bench's accuracy leg to report each solve's error against
FP64 (the solution of the FP32 solver, in FP32 or mixed-precision multigrid storage, against
the exact FP64 solution of the same system, next to what an FP64 CG stopped at the same
tolerance reaches).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

NB = ((-1, 0), (1, 0), (0, -1), (0, 1))


def pressure_matrix(mga, mgd, W, H):
    """(A, unknown mask) of level 0: A over the cells with mgd > 0, FP64 of the FP32 entries."""
    N = W * H
    d = np.asarray(mgd[:N], np.float64)
    a = np.asarray(mga[:N], np.float64).reshape(N, 4)
    unk = d > 0.0
    n = int(unk.sum())
    idx = np.full(N, -1, np.int64)
    idx[unk] = np.arange(n)
    ys, xs = np.divmod(np.arange(N), W)
    rows, cols, vals = [idx[unk]], [idx[unk]], [d[unk]]
    for k, (dx, dy) in enumerate(NB):
        nx, ny = xs + dx, ys + dy
        ok = unk & (nx >= 0) & (nx < W) & (ny >= 0) & (ny < H)
        nb = np.where(ok, ny * W + nx, 0)
        ok &= unk[nb]
        rows.append(idx[ok])
        cols.append(idx[nb[ok]])
        vals.append(-a[ok, k])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return A, unk


def rel(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b))


def solve_errors(A, unk, prhs, pres, tol):
    """Per potential (pressure*dt, density correction): the device solution's error against
    the exact FP64 solution, its true relative residual ||b - A x|| / ||b||, and the error
    of an FP64 CG (Jacobi-preconditioned, from zero) stopped at relative residual `tol`.
    Entries are None for a potential whose right-hand side is zero. Also returns the exact
    solutions as an (N, 2) FP64 array."""
    out = []
    M = sp.diags(1.0 / A.diagonal())
    Ac = A.tocsc()
    x64 = np.zeros((len(unk), 2), np.float64)
    for comp in (0, 1):
        b = np.asarray(prhs[:, comp], np.float64)[unk]
        x = np.asarray(pres[:, comp], np.float64)[unk]
        if not np.any(b):
            out.append(None)
            continue
        xe = sla.spsolve(Ac, b)
        x64[unk, comp] = xe
        xcg, _ = sla.cg(A, b, rtol=tol, M=M, maxiter=100000)
        out.append({"err_vs_fp64": rel(x, xe), "true_residual": rel(A @ x, b),
                    "fp64_cg_same_tol_err": rel(xcg, xe)})
    return out, x64


def live_faces_rel_l2(vel_a, vel_b):
    """rel-L2 of the face velocities (xy) over the faces that hold a velocity in both."""
    a = np.asarray(vel_a, np.float64).reshape(-1, 4)[:, :2]
    b = np.asarray(vel_b, np.float64).reshape(-1, 4)[:, :2]
    ok = (np.abs(a) < 1e29) & (np.abs(b) < 1e29)
    return rel(a[ok], b[ok])


# The stages of one substep up to the pressure right-hand side (record_substep, mod.rs:
# 2615-2708): step, accumulator clear, particle pass, grid passes with the surface tension,
# the viscosity before the solve, assembly.
PREFIX = ("STEP", "CLEAR_ACC", "PARTICLES", "GRID_PREP", "SURF", "VISC", "RHS")


def pressure_accuracy(sim, tol, gather):
    """One substep's stages up to the right-hand side on the device from the tank's current
    state, then its pressure solve and projection; returns the solve's errors against FP64
    (solve_errors), the projected face velocities' rel-L2 against those the exact FP64
    pressure gives from the same velocities, and the solve record."""
    from sprout_amd import _native as N
    W, H = sim.grid
    n = W * H
    L = N.levels(W, H)
    # a batch of one substep of 1/240 s (the step words record() writes, mod.rs:2521-2526)
    S = sim.field("STEP", N.STEP_DTYPE, 1)
    S["frame_left"] = 1.0 / 240.0
    S["batch_steps"] = 1
    S["batch_done"] = 0
    sim.put_field("STEP", S)
    for st in PREFIX:
        sim.stage(st, gather if st == "PARTICLES" else 0)
    mga = sim.field("MGA", np.float32, (L.total, 4))
    mgd = sim.field("MGD", np.float32, (L.total,))
    prhs = sim.field("PRHS", np.float32, (n, 4))
    A, unk = pressure_matrix(mga, mgd, W, H)
    sim.stage("PRESSURE")
    pres = sim.field("PRES", np.float32, (n, 2))
    K = sim.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
    vel_before = sim.field("VEL", np.float32, (n, 4))
    sim.stage("PROJECT")
    vel_dev = sim.field("VEL", np.float32, (n, 4))
    errs, x64 = solve_errors(A, unk, prhs, pres, tol)
    sim.put_field("VEL", vel_before)
    sim.put_field("PRES", x64.astype(np.float32))
    sim.stage("PROJECT")
    vel_ref = sim.field("VEL", np.float32, (n, 4))
    return errs, live_faces_rel_l2(vel_dev, vel_ref), K
