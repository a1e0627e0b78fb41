"""ctypes binding of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY: used
by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FIELD = dict(PARTICLES=0, ACC=1, VEL=2, CELL=3, COMP=4, COEF=5, SURF=6, PJUMP=7, MGA=8, MGD=9,
             PRHS=10, PRES=11, STEP=12, SOLVE=13, SPROP=14, AUX=15, GRHO=16, VEL0=17, MGZ=18,
             MGR=19, PTILE=20, REGION_AIR=21, SURF2=22, VW=23, VS=24, GBASE=25, VR=26, VP=27,
             VP2=28, VQ=29, VD=30, VR2=31, VQ2=32, VAZ=33, VTILES=34, VTILE=35, VMA=36, VMB=37,
             VMD=38, VMR=39, VMX=40, VMZ=41, P2KEEP=42, THERM=43, DHEAT=44, HCELL=45, SOLID_T=46,
             LEDGER=47, RL=48, RBODY=49, RH=50, ESTR=51, PSTRESS=52, FAMOCC=53, DSETTLE=54, MAGC=55, RMASK=56, MGX=57)
STAGE = dict(STEP=0, CLEAR_ACC=1, PARTICLES=2, GRID_PREP=3, COEF=4, SURF=5, VISC=6, RHS=7,
             PRESSURE=8, PROJECT=9, SUBSTEP=10, MG_BUILD=11, VCYCLE=12, VISC_SPLIT=13,
             VISC_SOLVE=14, RHS_VISC=15, PRESSURE2=16, PROJECT_VISC=17, VISC_VCYCLE=18, HEAT=19,
             RIGID=20, RHEO=21, ELASTIC=22, SETTLE=23, MAGNET=24)

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.or_create.argtypes = [C.c_int32, C.c_int32, C.c_uint32]
        L.or_create.restype = P
        L.or_destroy.argtypes = [P]
        L.or_set_gparams.argtypes = [P, P, P]
        L.or_set_solids.argtypes = [P, P, P, P, P]
        L.or_reset.argtypes = [P, P, C.c_uint32]
        L.or_begin_batch.argtypes = [P, C.c_float, C.c_uint32]
        L.or_substep.argtypes = [P, C.c_int]
        L.or_stage.argtypes = [P, C.c_int, C.c_int]
        L.or_get.argtypes = [P, C.c_int, P, C.c_size_t]
        L.or_get.restype = C.c_long
        L.or_put.argtypes = [P, C.c_int, P, C.c_size_t]
        L.or_put.restype = C.c_long
        L.or_n_slots.argtypes = [P]
        L.or_n_slots.restype = C.c_uint32
        L.or_sort.argtypes = [P]
        L.or_spawn.argtypes = [P, P, C.c_uint32]
        L.or_cell_counts.argtypes = [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P]
        L.or_stats.argtypes = [P, P, P, P]
        L.or_last_iters.argtypes = [P]
        L.or_last_iters.restype = C.c_uint32
        L.or_air_velocity.argtypes = [P, C.c_float, C.c_float, C.c_float, P, P]
        _lib = L
    return _lib


class OracleSim:
    """The CPU restatement of one tank (oracle.c), driven like the device handle."""

    def __init__(self, grid, capacity):
        self.grid = (int(grid[0]), int(grid[1]))
        self.capacity = int(capacity)
        self._lib = lib()
        self._h = self._lib.or_create(self.grid[0], self.grid[1], self.capacity)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.or_destroy(self._h)
            self._h = None

    def set_gparams(self, g, mats):
        self._lib.or_set_gparams(self._h, C.addressof(g), C.addressof(mats))

    def set_solids(self, field):
        sdf = np.ascontiguousarray(field.sdf, np.float32)
        solid = np.ascontiguousarray(field.solid, np.float32)
        region = np.ascontiguousarray(field.region, np.uint32)
        temp = getattr(field, "temp", None)
        temp = None if temp is None or len(temp) == 0 else np.ascontiguousarray(temp, np.float32)
        self._lib.or_set_solids(self._h, sdf.ctypes.data, solid.ctypes.data, region.ctypes.data,
                                None if temp is None else temp.ctypes.data)

    def reset(self, particles):
        p = np.ascontiguousarray(particles)
        self._lib.or_reset(self._h, p.ctypes.data if len(p) else None, len(p))

    def begin_batch(self, target, substeps):
        self._lib.or_begin_batch(self._h, float(target), int(substeps))

    def substep(self, gather):
        self._lib.or_substep(self._h, int(gather))

    def stage(self, name, arg=0):
        self._lib.or_stage(self._h, STAGE[name], int(arg))

    def field(self, name, dtype, shape):
        out = np.zeros(shape, dtype)
        n = self._lib.or_get(self._h, FIELD[name], out.ctypes.data, out.nbytes)
        if n < 0:
            raise ValueError(f"oracle field {name}: buffer too small")
        return out

    def put_field(self, name, arr):
        arr = np.ascontiguousarray(arr)
        n = self._lib.or_put(self._h, FIELD[name], arr.ctypes.data, arr.nbytes)
        if n < 0:
            raise ValueError(f"oracle put {name}: size mismatch")

    def n_slots(self):
        return self._lib.or_n_slots(self._h)

    def air_velocity(self, x, y, lean):
        """The air's velocity (cells / s) and the liquid occupancy at (x, y) (air.glsl)."""
        v = (C.c_float * 2)()
        occ = C.c_float()
        self._lib.or_air_velocity(self._h, float(x), float(y), float(lean), C.addressof(v), C.addressof(occ))
        return float(v[0]), float(v[1]), float(occ.value)

    def sort(self):
        self._lib.or_sort(self._h)

    def spawn(self, particles):
        """The record's spawn step: fresh slots first, then dead slots in slot order."""
        p = np.ascontiguousarray(particles)
        self._lib.or_spawn(self._h, p.ctypes.data if len(p) else None, len(p))

    def cell_counts(self, x0, y0, w, h):
        out = np.zeros(w * h, np.uint32)
        self._lib.or_cell_counts(self._h, x0, y0, w, h, out.ctypes.data)
        return out

    def stats(self):
        live = np.zeros(1, np.uint32)
        count = np.zeros(96, np.uint32)
        conc = np.zeros(96, np.uint32)
        self._lib.or_stats(self._h, live.ctypes.data, count.ctypes.data, conc.ctypes.data)
        return int(live[0]), count, conc

    def last_iters(self):
        return self._lib.or_last_iters(self._h)
