"""Parity of the HIP implicit viscosity (visc.hip) and the second pressure solve with the
CPU oracle.

Per-cell stages (weights, tile flags, explicit step outside the active tiles, the Jacobi
diagonal, the multigrid matrices of every level, one V-cycle, the second right-hand side
and projection) are bit-exact: same operation order, no FMA contraction on either side.
The solves reduce their scalars in FP64 on both sides in a different order and are compared
within rel-L2 1e-5 (FP32), iteration counts within one.
"""
import numpy as np
import pytest

from sprout_amd import _native as N
from test_gpu_parity import Pair, first_substep_prefix, rel_l2

pytestmark = pytest.mark.gpu

VISCOUS = dict(frac=(0.5, 0.6), viscosity=1e-2)


def viscous_pair(grid=(64, 36), **opts):
    o = {"viscosity_tol": 1e-5}
    o.update(opts)
    return Pair(grid, opts=o, **VISCOUS)


def to_split(p, vel_xy=None):
    """Both sides at the same state up to the viscosity split, the oracle's fields copied to
    the device so the split starts from identical inputs."""
    first_substep_prefix(p)
    p.ora.stage("GRID_PREP")
    if vel_xy is not None:
        vel = p.ora.field("VEL", np.float32, (p.N, 4))
        vel[:, :2] = vel_xy
        p.ora.put_field("VEL", vel)
    p.ora.stage("COEF")
    p.ora.stage("SURF")
    p.sync_to_gpu("STEP", "VEL", "CELL", "COMP", "COEF", "SURF", "PJUMP", "SPROP")
    p.both("VISC_SPLIT")


def n0(p):
    return ((p.grid[0] + 15) // 16) * ((p.grid[1] + 15) // 16)


def random_flow(p, seed):
    return (np.random.default_rng(seed).standard_normal((p.N, 2)) * 20.0).astype(np.float32)


@pytest.mark.parametrize("grid,opts", [((64, 36), {}), ((64, 36), {"viscosity_mg_sweeps": 1}),
                                       ((128, 72), {"viscosity_mg_levels": 3})])
def test_visc_split_bitexact(grid, opts):
    """Weights, strain, tile flags, explicit step in the inactive tiles, Jacobi diagonal,
    warm start and the multigrid's matrices of every level it uses (two below the grid at
    128 x 72, where the hierarchy's tail starts at level 2)."""
    p = viscous_pair(grid, **opts)
    try:
        assert p.g.visc_mode == 2 and p.g.v_twice == 1
        to_split(p, random_flow(p, 1))
        ta, tb = p.f("VTILES", np.uint32, n0(p))
        assert np.array_equal(ta, tb) and 0 < ta.sum() < len(ta)
        for name, shape in [("VW", (p.N, 4)), ("VS", (p.N, 4)), ("VEL", (p.N, 4)), ("VD", (p.N, 2))]:
            a, b = p.f(name, np.float32, shape)
            assert np.array_equal(a, b), name
        ta, tb = p.f("VTILE", np.uint32, p.L.tiles)
        assert np.array_equal(ta, tb)
        for name, k in [("VMA", 4), ("VMB", 4), ("VMD", 2)]:
            a, b = p.f(name, np.float32, (p.L.total, k))
            assert np.array_equal(a, b), name
            assert np.abs(a[: p.N]).max() > 0, name
    finally:
        p.close()


def test_visc_vcycle_bitexact():
    """One V-cycle of the viscosity multigrid on the same residual (vmg_down, the coarse
    levels, visc_mgz's way up and A z)."""
    p = viscous_pair()
    try:
        to_split(p, random_flow(p, 2))
        vw = p.ora.field("VW", np.float32, (p.N, 4))
        vmd = p.ora.field("VMD", np.float32, (p.L.total, 2))[: p.N]
        r = np.where(vmd > 0, np.random.default_rng(3).standard_normal((p.N, 2)), 0.0).astype(np.float32)
        assert (vmd > 0).sum() > 500 and vw[:, 2].min() > 0
        p.gpu.put_field("VR", r)
        p.ora.put_field("VR", r)
        p.both("VISC_VCYCLE")
        for name in ("VMZ", "VMX"):
            a, b = p.f(name, np.float32, (p.L.total, 2))
            assert np.array_equal(a[: p.N], b[: p.N]), name
        a, b = p.f("VAZ", np.float32, (p.N, 2))
        assert np.array_equal(a, b)
    finally:
        p.close()


@pytest.mark.parametrize("opts", [{}, {"viscosity_precond": 0}, {"viscosity_mg_sweeps": 1}, {"viscosity_warm": 0}])
def test_visc_solve_matches(opts):
    """The whole viscosity solve from the same split: velocities within rel-L2 1e-5,
    iterations within one, both converged."""
    p = viscous_pair(**opts)
    try:
        to_split(p, random_flow(p, 4))
        p.both("VISC_SOLVE", 0)
        a, b = p.f("VEL", np.float32, (p.N, 4))
        live = (np.abs(b[:, 2:]) < 1e29)
        assert rel_l2(a[:, :2][live], b[:, :2][live]) <= 1e-5
        ka, kb = p.f("SOLVE", N.SOLVE_DTYPE, 1)
        assert ka[0]["v_done"] == kb[0]["v_done"] == 1
        assert kb[0]["v_iters"] > 2
        assert abs(int(ka[0]["v_iters"]) - int(kb[0]["v_iters"])) <= 1
        a, b = p.f("VD", np.float32, (p.N, 2))
        assert rel_l2(a, b) <= 1e-5
    finally:
        p.close()


def test_second_projection_matches():
    """rhs_visc (bit-exact), the second pressure solve (rel-L2 1e-5, iterations within one)
    and project_visc (bit-exact on the same pressure), from the oracle's state after the
    viscosity of the second substep."""
    p = viscous_pair()
    try:
        p.ora.begin_batch(2 / 240, 2)
        p.ora.substep(0)
        for s in ("STEP", "CLEAR_ACC"):
            p.ora.stage(s)
        p.ora.stage("PARTICLES", 1)
        for s in ("GRID_PREP", "COEF", "SURF", "VISC", "RHS", "PRESSURE", "PROJECT"):
            p.ora.stage(s)
        p.ora.stage("VISC_SOLVE", 1)
        p.sync_to_gpu("STEP", "VEL", "CELL", "COEF", "PRHS", "PRES", "MGA", "MGD", "PTILE", "P2KEEP")
        p.both("RHS_VISC")
        a, b = p.f("PRHS", np.float32, (p.N, 4))
        assert np.array_equal(a, b)
        a, b = p.f("PRES", np.float32, (p.N, 2))
        assert np.array_equal(a, b)
        p.both("PRESSURE2")
        a, b = p.f("PRES", np.float32, (p.N, 2))
        assert rel_l2(a[:, 0], b[:, 0]) <= 1e-5
        assert np.array_equal(a[:, 1], b[:, 1])  # the density correction is not solved again
        ka, kb = p.f("SOLVE", N.SOLVE_DTYPE, 1)
        assert ka[0]["done"][0] == 1 and kb[0]["p2_iters_last"] > 0
        assert abs(int(ka[0]["p2_iters_last"]) - int(kb[0]["p2_iters_last"])) <= 1
        p.sync_to_gpu("PRES")
        p.both("PROJECT_VISC")
        a, b = p.f("VEL", np.float32, (p.N, 4))
        assert np.array_equal(a, b)
        a, b = p.f("P2KEEP", np.float32, (p.N,))
        assert np.array_equal(a, b)
    finally:
        p.close()


@pytest.mark.parametrize("opts", [{}, {"viscosity_precond": 0}, {"viscous_projection": 0}])
def test_whole_viscous_substeps_match(opts):
    """Six full substeps of a viscous dam break (the warm starts of the later ones included):
    particles by id within rel-L2 1e-5 (positions) / 1e-4 (velocities), both solves
    converged every substep."""
    p = Pair((96, 54), opts=dict({"viscosity_tol": 1e-5}, **opts), **VISCOUS)
    try:
        steps = 6
        p.gpu.record(steps, steps / 240)
        p.gpu.wait()
        p.ora.begin_batch(steps / 240, steps)
        for k in range(steps):
            p.ora.substep(1 if k > 0 else 0)
        pa = p.gpu.download_particles()
        pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        pa = pa[np.argsort(pa["id"])]
        pb = pb[np.argsort(pb["id"])]
        assert np.array_equal(pa["id"], pb["id"])
        assert rel_l2(pa["pos"], pb["pos"]) <= 1e-5
        assert rel_l2(pa["vel"], pb["vel"]) <= 1e-4
        st = p.gpu.solver
        assert st.viscosity_unconverged == 0 and st.pressure_unconverged == 0
        assert st.viscosity_iters_max > 0
        kb = p.ora.field("SOLVE", N.SOLVE_DTYPE, 1)[0]
        assert abs(int(st.viscosity_iters_max) - int(kb["v_it_max"])) <= 1
    finally:
        p.close()
