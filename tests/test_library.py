"""The C-ABI library loads and exports every entry point include/sprout_hip.h declares (no
compute: these run without a GPU)."""
import ctypes as C
import os
import re

import numpy as np

from sprout_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|void|const char\*)\s+(sp_\w+)\s*\(", text, re.M))
    return names


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(N.LIB_PATH)
    names = declared()
    assert len(names) >= 25
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert declared() == set(N.SIGNATURES)


def test_struct_layouts():
    assert N.PARTICLE_DTYPE.itemsize == 48  # params.rs:1319
    assert C.sizeof(N.GMat) == 208
    assert N.STEP_DTYPE.itemsize == 48      # 12 words (mod.rs:123 STEP_WORDS)


def test_python_structs_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets the C compiler gives the header's
    struct (a field added on one side only shifts everything after it)."""
    import subprocess
    pairs = {"sp_material": N.Material, "sp_sim_params": N.SimParams, "sp_gmat": N.GMat,
             "sp_gparams": N.GParams, "sp_solver_stats": N.SolverStats, "sp_levels_info": N.LevelsInfo,
             "sp_pool_stats": N.PoolStats, "sp_magnet": N.Magnet, "sp_device_view": N.DeviceView,
             "sp_pass_time": N.PassTime}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sprout_hip.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append('printf("sp_gpumat %zu\\n", sizeof(sp_gpumat));')
    for fname in N.GPUMAT_DTYPE.names:
        lines.append(f'printf("sp_gpumat.{fname} %zu\\n", offsetof(sp_gpumat, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for cname, py in pairs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(got[f"{cname}.{fname}"]) == getattr(py, fname).offset, f"{cname}.{fname}"
    assert int(got["sp_gpumat"]) == N.GPUMAT_DTYPE.itemsize == 208
    for fname in N.GPUMAT_DTYPE.names:
        assert int(got[f"sp_gpumat.{fname}"]) == N.GPUMAT_DTYPE.fields[fname][1], f"sp_gpumat.{fname}"


def test_errors_cross_the_abi_as_codes():
    L = N.lib()
    h = C.c_void_p()
    cfg = N.Config(4, 4, 16, 0)
    assert L.sp_create(C.byref(cfg), C.byref(h)) == N.E_INVALID
    assert b"grid" in L.sp_create_error()
    assert L.sp_create(None, C.byref(h)) == N.E_INVALID
    p = N.default_params()
    g = N.GParams()
    mats = (N.GMat * N.MAX_MATERIALS)()
    assert L.sp_derive_grid_params(C.byref(p), 64, 64, None, -1, 0, None, 0, None, 0, C.byref(g), mats) == N.E_INVALID
    assert L.sp_derive_grid_params(C.byref(p), 64, 64, None, 0, 0, None, 0, None, 2, C.byref(g), mats) == N.E_INVALID
    # thirteen materials do not fit a tank (MAX_PRESENT 12, mod.rs:2051-2081)
    rows = (C.c_int32 * 13)(*range(13))
    assert L.sp_derive_grid_params(C.byref(p), 64, 64, rows, 13, 0, None, 0, None, 0, C.byref(g), mats) == N.E_CAPACITY


def test_viscosity_modes_follow_the_diffusion_number():
    """viscosity_mode (mod.rs:2768-2792): none for an inviscid tank, explicit for water,
    implicit for honey."""
    p = N.default_params()
    p.mat[0].viscosity = 0.0
    g, _ = N.derive_grid_params(p, 64, 64, [0])
    assert g.visc_mode == 0
    p = N.default_params()
    g, _ = N.derive_grid_params(p, 64, 64, [0])
    assert g.visc_mode == 1
    g, _ = N.derive_grid_params(p, 64, 64, [1])
    assert g.visc_mode == 2


def test_tank_bytes_scale_with_the_tank():
    """sp_tank_bytes needs no device: evictable bytes grow with the grid and the capacity,
    and a 1024 x 576 tank at 4 per cell fits hundreds of times in 288 GB."""
    small, fs = N.tank_bytes(128, 72, 128 * 72 * 4)
    mid, fm = N.tank_bytes(512, 288, 512 * 288 * 4)
    big, fb = N.tank_bytes(1024, 576, 1024 * 576 * 4)
    assert 0 < small < mid < big
    assert 0 < fs < fm < fb
    # the arena alone: at least the particles (48 B) and the accumulator (128 B per cell)
    assert big > 48 * 1024 * 576 * 4 + 128 * 1024 * 576
    assert 288e9 / (big + fb) > 300
    L = N.lib()
    ev, fx = C.c_uint64(), C.c_uint64()
    assert L.sp_tank_bytes(4, 4, 16, C.byref(ev), C.byref(fx)) == N.E_INVALID


def test_pool_and_batch_refuse_bad_arguments():
    L = N.lib()
    h = C.c_void_p()
    assert L.sp_batch_create(None, 0, C.byref(h)) == N.E_INVALID
    r = C.c_int32()
    assert L.sp_is_resident(None, C.byref(r)) == N.E_INVALID
    # the pool checks its arguments before touching a device
    assert L.sp_pool_create(0, 1 << 30, 0, C.byref(h)) == N.E_INVALID
    assert L.sp_pool_create(0, 1 << 30, 17, C.byref(h)) == N.E_INVALID
    assert L.sp_pool_create(0, 0, 2, C.byref(h)) == N.E_INVALID
    assert L.sp_pool_submit(None, 0, 1, 0.1, None) == N.E_INVALID
    assert L.sp_pool_step(None, 0.0, None) == N.E_INVALID
    assert L.sp_pool_drain(None) == N.E_INVALID
    assert L.sp_pool_wait(None, 1, None) == N.E_INVALID
    assert L.sp_pool_last_error(None) == b"null pool"


def test_autotune_modes_and_shapes_without_device():
    shapes = N.autotune_shapes()
    assert len(shapes) >= 2 and shapes[0] == (256, 12288, 1)
    assert len(set(shapes)) == len(shapes)
    assert all(t % 64 == 0 and 0 < b * 4 <= 64 * 1024 and m >= 1 for t, b, m in shapes)
    L = N.lib()
    assert L.sp_autotune(len(shapes), 0) == N.E_INVALID
    assert L.sp_autotune(-3, 0) == N.E_INVALID
    N.autotune(1)
    N.autotune(-1)
    assert isinstance(N.autotune_report(), list)
