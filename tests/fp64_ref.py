"""FP64 references for the pressure and viscosity solves.

Pressure: the level-0 system the device solves, assembled
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
    of an FP64 CG (Jacobi-preconditioned, from zero) stopped at the potential's own relative
    residual tolerance (`tol`: one for both, or a pair: pressure_tol, correction_tol). Entries
    are None for a potential whose right-hand side is zero. Also returns the exact solutions
    as an (N, 2) FP64 array."""
    tols = tuple(tol) if isinstance(tol, (tuple, list)) else (tol, tol)
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
        xcg, _ = sla.cg(A, b, rtol=tols[comp], M=M, maxiter=100000)
        # the FP32 floor of the true residual: eps32 || |A| |x| || / ||b||
        floor = float(np.finfo(np.float32).eps * np.linalg.norm(abs(A) @ np.abs(x)) / np.linalg.norm(b))
        out.append({"tol": float(tols[comp]), "err_vs_fp64": rel(x, xe), "true_residual": rel(A @ x, b),
                    "fp32_floor": floor, "fp64_cg_same_tol_err": rel(xcg, xe)})
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
    (solve_errors, with "vel_floor_1ulp": the face velocities' rel-L2 change when the exact
    pressure, rounded to FP32, moves by one ulp), the projected face velocities' rel-L2
    against those the exact FP64 pressure gives from the same velocities, and the solve
    record."""
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
    x32 = x64.astype(np.float32)
    sim.put_field("PRES", x32)
    sim.stage("PROJECT")
    vel_ref = sim.field("VEL", np.float32, (n, 4))
    # the floor of any FP32 pressure: the projection of the exact pressure moved by one ulp
    # up or down at random (a seeded draw) against that of the exact pressure
    rng = np.random.default_rng(0)
    up = rng.random(x32.shape) < 0.5
    nudged = np.where(up, np.nextafter(x32, np.float32(np.inf)), np.nextafter(x32, np.float32(-np.inf)))
    sim.put_field("VEL", vel_before)
    sim.put_field("PRES", np.where(x64 != 0.0, nudged, x32).astype(np.float32))
    sim.stage("PROJECT")
    floor = live_faces_rel_l2(sim.field("VEL", np.float32, (n, 4)), vel_ref)
    for e in errs:
        if e is not None:
            e["vel_floor_1ulp"] = floor
    return errs, live_faces_rel_l2(vel_dev, vel_ref), K


# ---- the implicit viscosity solve ----------------------------------------------------------

def _at(a, k):
    """a[i + k] over the flat cell index i, zero outside the array (the shaders' flat
    neighbour indexing: i +- 1 crosses rows at the tank's edges, where the weights are 0)."""
    out = np.zeros_like(a)
    if k > 0:
        out[:-k] = a[k:]
    elif k < 0:
        out[-k:] = a[:k]
    else:
        out[:] = a
    return out


class ViscSystem:
    """The implicit viscosity system as the device assembles it (visc_diag.comp:75-117,
    visc_iter.comp, visc_apply.comp), in FP64 over the FP32 fields: unknowns are the live
    faces (u of a cell: its left face, v: its bottom face) of interior cells in active tiles;
    A p = C p + G^T W G p with C the face masses (coef.zw), G the strain of a face field
    (exx = u(i+1) - u(i), eyy = v(i+W) - v(i) at cells weighted by vw.x; the shear at cell
    corners weighted by vw.y, with a closed face mirroring its live neighbour) and G^T its
    divergence back onto the faces (closed neighbours' corners at full weight, open ones at
    half); the right-hand side is -G^T W G u* (the strain vs of the velocity before the
    solve). Matrix-free: A is applied by the stencil, never stored."""

    def __init__(self, vel, vw, coef, vtiles, W, H):
        self.W, self.H = W, H
        n = W * H
        vel = np.asarray(vel, np.float32).reshape(n, 4)
        self.vw = np.asarray(vw, np.float64).reshape(n, 4)
        self.coef = np.asarray(coef, np.float64).reshape(n, 4)
        self.live_u = np.abs(vel[:, 2]) < 1e29
        self.live_v = np.abs(vel[:, 3]) < 1e29
        self.closed_u = vel[:, 2] < -1e29
        self.closed_v = vel[:, 3] < -1e29
        ys, xs = np.divmod(np.arange(n), W)
        tiles = (ys >> 4) * ((W + 15) >> 4) + (xs >> 4)
        on = np.asarray(vtiles)[tiles] != 0
        interior = (xs >= 1) & (ys >= 1) & (xs < W - 1) & (ys < H - 1)
        self.cells = on & interior
        self.unk = np.concatenate([self.cells & self.live_u, self.cells & self.live_v])
        self.u_star = np.concatenate([vel[:, 0], vel[:, 1]]).astype(np.float64)
        self.diag = np.concatenate([self.vw[:, 2], self.vw[:, 3]])[self.unk]

    def strain(self, pu, pv):
        """(exx, eyy, exy) per cell of the face field (pu, pv), as strain_of_dir."""
        W = self.W
        wx, wy = self.vw[:, 0], self.vw[:, 1]
        PU = np.where(self.live_u, pu, 0.0)
        PV = np.where(self.live_v, pv, 0.0)
        ox = np.where(wx > 0, wx * (_at(PU, 1) - PU), 0.0)
        oy = np.where(wx > 0, wx * (_at(PV, W) - PV), 0.0)
        ua, ub, va, vb = PU, _at(PU, -W), PV, _at(PV, -1)
        ua = np.where(self.closed_u & _at(self.live_u, -W), -ub, ua)
        ub = np.where(_at(self.closed_u, -W) & self.live_u, -ua, ub)
        va = np.where(self.closed_v & _at(self.live_v, -1), -vb, va)
        vb = np.where(_at(self.closed_v, -1) & self.live_v, -va, vb)
        oz = np.where(wy > 0, wy * 0.5 * ((ua - ub) + (va - vb)), 0.0)
        return ox, oy, oz

    def gts(self, ox, oy, oz):
        """The strain's divergence onto the live faces, as gts_of."""
        W = self.W
        c1 = np.where(_at(self.closed_u, -W), 1.0, 0.5)
        c2 = np.where(_at(self.closed_u, W), -1.0, -0.5)
        gx = np.where(self.live_u, _at(ox, -1) - ox + oz * c1 + _at(oz, W) * c2, 0.0)
        d1 = np.where(_at(self.closed_v, -1), 1.0, 0.5)
        d2 = np.where(_at(self.closed_v, 1), -1.0, -0.5)
        gy = np.where(self.live_v, _at(oy, -W) - oy + oz * d1 + _at(oz, 1) * d2, 0.0)
        return gx, gy

    def apply(self, x):
        """A x for x over the unknowns."""
        n = self.W * self.H
        full = np.zeros(2 * n)
        full[self.unk] = x
        gx, gy = self.gts(*self.strain(full[:n], full[n:]))
        q = np.concatenate([self.coef[:, 2] * full[:n] + gx, self.coef[:, 3] * full[n:] + gy])
        return q[self.unk]

    def rhs(self, vs):
        """-G^T (W G u*) from the strain vs the device holds."""
        vs = np.asarray(vs, np.float64).reshape(-1, 4)
        gx, gy = self.gts(vs[:, 0], vs[:, 1], vs[:, 2])
        return -np.concatenate([gx, gy])[self.unk]

    def operator(self):
        m = int(self.unk.sum())
        return sla.LinearOperator((m, m), matvec=self.apply, dtype=np.float64)


def _cg(system, b, rtol, maxiter=200000):
    M = sla.LinearOperator(system.operator().shape, matvec=lambda r: r / system.diag, dtype=np.float64)
    x, info = sla.cg(system.operator(), b, rtol=rtol, M=M, maxiter=maxiter)
    assert info == 0, f"FP64 CG did not reach {rtol}"
    return x


def viscosity_errors(system, vs, vel_after, tol):
    """The device's viscosity solve against FP64: the increment it made to the unknown faces
    against the FP64 solution (CG to 1e-13, effectively exact), the velocities' rel-L2
    against u* + that solution, the true relative residual, and the error of an FP64 CG
    from zero stopped at relative residual `tol`."""
    b = system.rhs(vs)
    after = np.asarray(vel_after, np.float32).reshape(-1, 4)
    u_dev = np.concatenate([after[:, 0], after[:, 1]]).astype(np.float64)[system.unk]
    u0 = system.u_star[system.unk]
    d_dev = u_dev - u0
    nb = np.linalg.norm(b)
    if nb == 0.0:
        return {"unknowns": int(system.unk.sum()), "rhs_zero": True}
    d64 = _cg(system, b, 1e-13)
    dcg = _cg(system, b, tol)
    return {"unknowns": int(system.unk.sum()), "err_vs_fp64": rel(d_dev, d64),
            "vel_err_vs_fp64": rel(u_dev, u0 + d64), "true_residual": rel(system.apply(d_dev), b),
            "fp64_cg_same_tol_err": rel(dcg, d64), "fp64_cg_same_tol_vel_err": rel(u0 + dcg, u0 + d64)}


def viscosity_accuracy(sim, tol, gather):
    """One substep's stages on the device up to its implicit viscosity solve (after the
    first projection when the tank projects twice, record_substep mod.rs:2615-2708), the
    solve, and its errors against FP64 (viscosity_errors) with the solve record."""
    from sprout_amd import _native as N
    W, H = sim.grid
    n = W * H
    g, _ = sim.grid_params()
    assert g.visc_mode == 2, "the tank has no implicit viscosity solve"
    S = sim.field("STEP", N.STEP_DTYPE, 1)
    S["frame_left"] = 1.0 / 240.0
    S["batch_steps"] = 1
    S["batch_done"] = 0
    sim.put_field("STEP", S)
    for st in ("STEP", "CLEAR_ACC", "PARTICLES", "GRID_PREP", "SURF", "VISC_SPLIT"):
        sim.stage(st, gather if st == "PARTICLES" else 0)
    restrain = 0
    if g.v_twice:
        for st in ("RHS", "PRESSURE", "PROJECT"):
            sim.stage(st)
        restrain = 1
    vel0 = sim.field("VEL", np.float32, (n, 4))
    sim.stage("VISC_SOLVE", restrain)
    vel1 = sim.field("VEL", np.float32, (n, 4))
    n0 = ((W + 15) // 16) * ((H + 15) // 16)
    system = ViscSystem(vel0, sim.field("VW", np.float32, (n, 4)), sim.field("COEF", np.float32, (n, 4)),
                        sim.field("VTILES", np.uint32, n0), W, H)
    K = sim.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
    return viscosity_errors(system, sim.field("VS", np.float32, (n, 4)), vel1, tol), K
