"""Parity of the HIP fluid step with the CPU oracle, stage by stage and over whole substeps.

Integer and index work (scatter accumulators, tile flags, sort permutation, cell counts)
must be bit-exact. Float stages without transcendental functions are bit-exact too (same
operation order, no FMA contraction on either side); fields that pass through exp() are
compared to within a few ulp. The pressure solve reduces in FP64 on both sides, in a
different order, and is compared within rel-L2 1e-5.
"""
import numpy as np
import pytest

import cases
from sprout_amd import _native as N
from sprout_amd.fluid_sim import FluidSim
from oracle.binding import OracleSim

pytestmark = pytest.mark.gpu

F4 = np.float32


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


class Pair:
    """The same tank on the device and in the oracle."""

    def __init__(self, grid, frac=(0.4, 0.6), opts=None, case=None, **kw):
        self.grid = grid
        W, H = grid
        if case is None:
            self.params, self.field, self.parts, rows = cases.dam_break(grid, frac=frac, **kw)
        else:
            self.params, self.field, self.parts, rows = case(grid, **kw)
        for k, v in (opts or {}).items():
            setattr(self.params, k, v)
        cap = W * H * 4
        self.gpu = FluidSim(grid, cap, params=self.params)
        self.gpu.set_solids(self.field)
        self.gpu.reset(self.parts)
        self.g, self.mats = self.gpu.grid_params()
        self.ora = OracleSim(grid, cap)
        self.ora.set_gparams(self.g, self.mats)
        self.ora.set_solids(self.field)
        self.ora.reset(self.parts)
        self.N = W * H
        self.stride = self.g.acc_stride
        self.L = N.levels(W, H)

    def both(self, stage, arg=0):
        self.gpu.stage(stage, arg)
        self.ora.stage(stage, arg)

    def f(self, name, dtype, shape):
        return self.gpu.field(name, dtype, shape), self.ora.field(name, dtype, shape)

    def sync_to_gpu(self, *names):
        """Copy oracle fields into the device state so a stage starts from identical inputs."""
        shapes = {"VEL": (self.N, 4), "CELL": (self.N, 4), "COMP": (self.N, 4), "COEF": (self.N, 4),
                  "SURF": (self.N, 4), "PJUMP": (self.N, 2), "PRHS": (self.N, 4), "PRES": (self.N, 2),
                  "MGA": (self.L.total, 4), "MGD": (self.L.total,), "SPROP": (self.N, 4),
                  "VEL0": (self.N, 2), "PTILE": (self.L.tiles,), "AUX": (self.N, 4), "GRHO": (self.N, 4),
                  "SURF2": (self.N, 4), "GBASE": (self.N, 4), "VW": (self.N, 4), "VS": (self.N, 4),
                  "P2KEEP": (self.N,), "THERM": (self.N, 4), "DHEAT": (self.N, 4), "SOLID_T": (self.N,),
                  "ESTR": (self.N, 4), "PSTRESS": (self.ora.n_slots(), 4),
                  "FAMOCC": (self.N * max(self.g.n_fam, 1),), "DSETTLE": (self.N * max(self.g.n_fam, 1), 2),
                  "MAGC": (self.N, 2)}
        for v in ("VR", "VP", "VP2", "VQ", "VD", "VR2", "VQ2", "VAZ"):
            shapes[v] = (self.N, 2)
        for v in ("VMA", "VMB"):
            shapes[v] = (self.L.total, 4)
        for v in ("VMD", "VMR", "VMX", "VMZ"):
            shapes[v] = (self.L.total, 2)
        n0 = ((self.grid[0] + 15) // 16) * ((self.grid[1] + 15) // 16)
        for n in names:
            if n == "STEP":
                self.gpu.put_field("STEP", self.ora.field("STEP", N.STEP_DTYPE, 1))
            elif n == "SOLVE":
                self.gpu.put_field("SOLVE", self.ora.field("SOLVE", N.SOLVE_DTYPE, 1))
            elif n == "ACC":
                self.gpu.put_field("ACC", self.ora.field("ACC", np.int32, self.N * self.stride))
            elif n == "PARTICLES":
                self.gpu.put_field("PARTICLES", self.ora.field("PARTICLES", N.PARTICLE_DTYPE, self.ora.n_slots()))
            elif n == "HCELL":
                self.gpu.put_field("HCELL", self.ora.field("HCELL", np.int32, 2 * self.N))
            elif n == "LEDGER":
                self.gpu.put_field("LEDGER", self.ora.field("LEDGER", np.uint32, N.LEDGER_WORDS))
            elif n == "RL":
                self.gpu.put_field("RL", self.ora.field("RL", np.int32, 2 * self.N))
            elif n == "RBODY":
                self.gpu.put_field("RBODY", self.ora.field("RBODY", np.uint32, N.RIGID_WORDS))
            elif n == "RH":
                self.gpu.put_field("RH", self.ora.field("RH", np.float32, (self.N, 4)))
            elif n in ("PTILE", "VTILE"):
                self.gpu.put_field(n, self.ora.field(n, np.uint32, self.L.tiles))
            elif n == "VTILES":
                self.gpu.put_field(n, self.ora.field(n, np.uint32, n0))
            else:
                dt = np.float32
                self.gpu.put_field(n, self.ora.field(n, dt, shapes[n]))

    def close(self):
        self.gpu.close()


@pytest.fixture(scope="module")
def small():
    p = Pair((64, 36))
    yield p
    p.close()


def first_substep_prefix(p):
    p.ora.begin_batch(1 / 240, 1)
    p.both("STEP")
    p.both("CLEAR_ACC")
    p.both("PARTICLES", 0)


def test_step_state_bitexact(small):
    p = Pair((64, 36))
    try:
        p.both("STEP")
        a, b = p.f("STEP", N.STEP_DTYPE, 1)
        assert a.tobytes() == b.tobytes()
    finally:
        p.close()


def test_scatter_first_substep_bitexact():
    """P2G of the seeded block: every accumulator word equal (particles.comp:997-1101)."""
    for grid in [(64, 36), (96, 54), (128, 72)]:
        p = Pair(grid)
        try:
            first_substep_prefix(p)
            a, b = p.f("ACC", np.int32, p.N * p.stride)
            assert np.array_equal(a, b), f"{grid}: {np.count_nonzero(a != b)} words differ"
            assert a.reshape(-1, p.stride)[:, 4].sum() > 0
        finally:
            p.close()


def test_grid_prep_and_assembly_bitexact(small):
    p = small
    p.gpu.reset(p.parts)
    p.ora.reset(p.parts)
    first_substep_prefix(p)
    p.both("GRID_PREP")
    for name in ["VEL", "COMP", "SPROP"]:
        a, b = p.f(name, np.float32, (p.N, 4))
        assert np.array_equal(a, b), name
    a, b = p.f("CELL", np.float32, (p.N, 4))
    assert np.array_equal(a[:, [0, 1, 3]], b[:, [0, 1, 3]])
    np.testing.assert_allclose(a[:, 2], b[:, 2], rtol=4e-7, atol=0)
    ra, rb = p.f("REGION_AIR", np.uint32, 1)
    assert ra[0] == rb[0]
    p.sync_to_gpu("CELL")
    p.both("COEF")
    a, b = p.f("COEF", np.float32, (p.N, 4))
    assert np.array_equal(a, b)
    p.both("SURF")
    a, b = p.f("SURF", np.float32, (p.N, 4))
    assert np.array_equal(a, b)
    a, b = p.f("PJUMP", np.float32, (p.N, 2))
    assert np.array_equal(a, b)
    p.both("RHS")
    a, b = p.f("MGA", np.float32, (p.L.total, 4))
    assert np.array_equal(a[: p.N], b[: p.N])
    a, b = p.f("MGD", np.float32, (p.L.total,))
    assert np.array_equal(a[: p.N], b[: p.N])
    a, b = p.f("PRHS", np.float32, (p.N, 4))
    assert np.array_equal(a[:, 0], b[:, 0])
    np.testing.assert_allclose(a[:, 1], b[:, 1], rtol=4e-6, atol=1e-30)
    ta, tb = p.f("PTILE", np.uint32, (p.L.tiles,))
    n0 = ((p.grid[0] + 15) // 16) * ((p.grid[1] + 15) // 16)
    assert np.array_equal(ta[:n0], tb[:n0])


def test_pressure_solve_matches_oracle(small):
    p = small
    p.gpu.reset(p.parts)
    p.ora.reset(p.parts)
    first_substep_prefix(p)
    for s in ["GRID_PREP", "COEF", "SURF", "RHS"]:
        p.ora.stage(s)
    p.sync_to_gpu("STEP", "VEL", "CELL", "COMP", "COEF", "SURF", "PJUMP", "PRHS", "MGA", "MGD", "PRES", "PTILE", "SPROP")
    p.both("MG_BUILD")
    a, b = p.f("MGD", np.float32, (p.L.total,))
    assert np.array_equal(a, b)
    a, b = p.f("MGA", np.float32, (p.L.total, 4))
    assert np.array_equal(a, b)
    ta, tb = p.f("PTILE", np.uint32, (p.L.tiles,))
    assert np.array_equal(ta, tb)
    p.sync_to_gpu("PRES", "PTILE")
    p.both("PRESSURE")
    a, b = p.f("PRES", np.float32, (p.N, 2))
    assert rel_l2(a[:, 0], b[:, 0]) <= 1e-5
    assert rel_l2(a[:, 1], b[:, 1]) <= 1e-5
    ka, kb = p.f("SOLVE", N.SOLVE_DTYPE, 1)
    assert ka[0]["done"].tolist() == [1, 1]
    assert abs(int(ka[0]["iters"].max()) - int(kb[0]["iters"].max())) <= 1


def test_vcycle_bitexact(small):
    """One V-cycle on the same residual: every level's arithmetic in the same order."""
    p = small
    p.gpu.reset(p.parts)
    p.ora.reset(p.parts)
    first_substep_prefix(p)
    for s in ["GRID_PREP", "COEF", "SURF", "RHS", "MG_BUILD"]:
        p.ora.stage(s)
    p.sync_to_gpu("MGA", "MGD", "PTILE")
    rng = np.random.default_rng(3)
    r = np.zeros((p.L.total, 2), np.float32)
    r[: p.N] = rng.standard_normal((p.N, 2)).astype(np.float32)
    mgd = p.ora.field("MGD", np.float32, (p.L.total,))
    r[: p.N][mgd[: p.N] <= 0] = 0
    p.gpu.put_field("MGR", r)
    p.ora.put_field("MGR", r)
    p.both("VCYCLE")
    a, b = p.f("MGZ", np.float32, (p.L.total, 2))
    unk = mgd[: p.N] > 0
    assert np.array_equal(a[: p.N][unk], b[: p.N][unk])


def test_project_bitexact(small):
    p = small
    p.gpu.reset(p.parts)
    p.ora.reset(p.parts)
    first_substep_prefix(p)
    for s in ["GRID_PREP", "COEF", "SURF", "RHS", "PRESSURE"]:
        p.ora.stage(s)
    p.sync_to_gpu("STEP", "VEL", "CELL", "COEF", "PJUMP", "PRES", "MGD", "PRHS")
    p.both("PROJECT")
    a, b = p.f("VEL", np.float32, (p.N, 4))
    assert np.array_equal(a, b)
    sa, sb = p.f("STEP", N.STEP_DTYPE, 1)
    assert sa["umax_grid"][0] == sb["umax_grid"][0]


def test_gather_advect_bitexact(small):
    """G2P + advection + collision + scatter on identical grid fields (second substep)."""
    p = small
    p.gpu.reset(p.parts)
    p.ora.reset(p.parts)
    p.ora.begin_batch(2 / 240, 2)
    p.ora.substep(0)
    p.ora.stage("STEP")
    p.ora.stage("CLEAR_ACC")
    p.sync_to_gpu("STEP", "VEL", "COMP", "CELL", "SURF", "PARTICLES")
    p.gpu.stage("CLEAR_ACC")
    p.both("PARTICLES", 1)
    pa = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, len(p.parts))
    pb = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
    assert pa.tobytes() == pb.tobytes()
    a, b = p.f("ACC", np.int32, p.N * p.stride)
    assert np.array_equal(a, b)
    sa, sb = p.f("STEP", N.STEP_DTYPE, 1)
    assert sa["umax_part"][0] == sb["umax_part"][0]


def test_whole_substeps_match(small):
    """Ten full substeps from the same seeded block: particles by id, grid velocity and
    pressure within rel-L2 1e-5."""
    p = Pair((96, 54))
    try:
        steps = 10
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
        va, _ = p.gpu.download_grid()
        vb = p.ora.field("VEL", np.float32, (p.N, 4)).reshape(va.shape)
        live = (np.abs(vb[..., 2]) < 1e29) & (np.abs(va[..., 2]) < 1e29)
        assert rel_l2(va[..., 0][live], vb[..., 0][live]) <= 1e-4
        pra, _ = p.gpu.download_pressure()
        prb = p.ora.field("PRES", np.float32, (p.N, 2))[:, 0].reshape(pra.shape)
        assert rel_l2(pra, prb) <= 1e-4
    finally:
        p.close()


def test_sort_permutation_bitexact():
    p = Pair((64, 36))
    try:
        p.gpu.record(3, 3 / 240)
        p.gpu.wait()
        p.ora.begin_batch(3 / 240, 3)
        for k in range(3):
            p.ora.substep(1 if k > 0 else 0)
        p.sync_to_gpu("PARTICLES")
        p.gpu.sort()
        p.ora.sort()
        a = p.gpu.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        b = p.ora.field("PARTICLES", N.PARTICLE_DTYPE, p.ora.n_slots())
        assert np.array_equal(a["id"], b["id"])
        assert a.tobytes() == b.tobytes()
    finally:
        p.close()


def test_cell_counts_and_stats(small):
    p = small
    p.gpu.reset(p.parts)
    p.ora.reset(p.parts)
    a = p.gpu.cell_counts((3, 3), (40, 30))
    b = p.ora.cell_counts(3, 3, 37, 27)
    assert np.array_equal(a, b)
    s = p.gpu.stats()
    live, count, conc = p.ora.stats()
    assert s.live == live == len(p.parts)
    assert list(s.count) == count.tolist()
