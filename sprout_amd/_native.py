"""ctypes view of the C ABI in include/sprout_hip.h (libsprout_hip.so).

The library is built in-tree (sprout_amd/libsprout_hip.so) by sprout_amd.build; loading it
fails loudly when it is missing: there is no CPU fallback for the fluid step.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# SPROUT_HIP_LIB: another build of the library (a variant or an installed copy)
LIB_PATH = os.environ.get("SPROUT_HIP_LIB") or os.path.join(HERE, "libsprout_hip.so")

GROUPS = 4
MAX_PRESENT = 12
MAX_MATERIALS = 96
MAX_LEVELS = 16
PPC = 4
SPAWN_MAX = 32768
BIN_NOALLOC_MAX = 1048576  # SP_BIN_NOALLOC_MAX

KIND_DEAD = 0
KIND_FRESH = 0x10000
KIND_REAL_T = 0x20000
SPAWN_TEMP = -1.0e4
SPAWN_S = -1.0

ROW_WATER, ROW_HONEY, ROW_OIL, ROW_MERCURY = 0, 1, 2, 8
ROW_BRINE, ROW_SALT, ROW_STEAM = 14, 56, 57  # material.rs: Brine, Salt, Steam
KIND_PUFF = 0x40000  # params.rs:28
KIND_SRC = 0xFF00    # params.rs:31: of steam, the row + 1 of the liquid it boiled from

# Status codes
OK, E_INVALID, E_DEVICE, E_NOMEM, E_UNSUPPORTED, E_CAPACITY, E_BUSY = 0, -1, -2, -3, -4, -5, -6
RIGID_BODIES, BODY_WORDS, RIGID_FIX = 4096, 32, 65536.0
RIGID_WORDS = 16 + BODY_WORDS * RIGID_BODIES
CAP_FAMILIES, CAP_THERMAL, CAP_RIGID, CAP_RHEO, CAP_MAGNET, CAP_ELASTIC, CAP_SETTLE = 1, 2, 4, 8, 16, 32, 64
MAX_MAGNETS = 8
MAX_CURRENTS = 16  # params.rs:1069
CURRENT_SPEED, CURRENT_RATE, DRAIN_RATE = 150.0, 20.0, 40.0  # params.rs:1065, 1099, 1101
CAP_CURRENTS = 0x40000
CAP_VOLATILE = 0x10000
CAP_SELF_HEAT = 0x20000
CAP_REACT = 256
CAP_STEAM = 128
CAP_BOILS = 0x80000
CAP_PUFFS = 0x100000
LEDGER_WORDS = 16 + 3 * 96
STRESS_FIX = 65536.0

# Field ids (SP_FIELD_*)
FIELD = dict(PARTICLES=0, ACC=1, VEL=2, CELL=3, COMP=4, COEF=5, SURF=6, PJUMP=7, MGA=8, MGD=9,
             PRHS=10, PRES=11, STEP=12, SOLVE=13, SPROP=14, AUX=15, GRHO=16, VEL0=17, MGZ=18,
             MGR=19, PTILE=20, REGION_AIR=21, SURF2=22, VW=23, VS=24, GBASE=25, VR=26, VP=27,
             VP2=28, VQ=29, VD=30, VR2=31, VQ2=32, VAZ=33, VTILES=34, VTILE=35, VMA=36, VMB=37,
             VMD=38, VMR=39, VMX=40, VMZ=41, P2KEEP=42, THERM=43, DHEAT=44, HCELL=45, SOLID_T=46,
             LEDGER=47, RL=48, RBODY=49, RH=50, ESTR=51, PSTRESS=52, FAMOCC=53, DSETTLE=54, MAGC=55, RMASK=56, MGX=57)
# Stage ids (SP_STAGE_*)
STAGE = dict(STEP=0, CLEAR_ACC=1, PARTICLES=2, GRID_PREP=3, COEF=4, SURF=5, VISC=6, RHS=7,
             PRESSURE=8, PROJECT=9, SUBSTEP=10, MG_BUILD=11, VCYCLE=12, VISC_SPLIT=13,
             VISC_SOLVE=14, RHS_VISC=15, PRESSURE2=16, PROJECT_VISC=17,
             VISC_VCYCLE=18, HEAT=19, RIGID=20, RHEO=21, ELASTIC=22, SETTLE=23, MAGNET=24)

# `struct Particle` (common.glsl:11-19): 48 bytes.
PARTICLE_DTYPE = np.dtype([("pos", "<f4", (2,)), ("vel", "<f4", (2,)), ("c", "<f4", (4,)),
                           ("s", "<f4"), ("kind", "<u4"), ("id", "<u4"), ("t", "<f4")])
assert PARTICLE_DTYPE.itemsize == 48


class Material(C.Structure):
    _fields_ = [("group", C.c_int32), ("diluent", C.c_int32), ("density", C.c_float),
                ("density_solid", C.c_float), ("viscosity", C.c_float), ("temp_model", C.c_int32),
                ("temp_b", C.c_float), ("temp_c", C.c_float), ("temp_ref", C.c_float),
                ("tension", C.c_float), ("contact_angle", C.c_float), ("mix_rate", C.c_float),
                ("expansion", C.c_float), ("melt_lo", C.c_float), ("melt_hi", C.c_float),
                ("visc_exponent", C.c_float), ("load", C.c_float), ("spawn_temp", C.c_float),
                ("latent", C.c_float), ("heat_capacity", C.c_float), ("conductivity", C.c_float),
                ("emissivity", C.c_float), ("boil", C.c_float), ("latent_vap", C.c_float),
                ("freeze_shift", C.c_float), ("rheology", C.c_uint32), ("yield_stress", C.c_float),
                ("flow_index", C.c_float), ("shear_time", C.c_float), ("visc_ratio", C.c_float),
                ("skin", C.c_float), ("friction", C.c_float), ("shear_modulus", C.c_float),
                ("relax_time", C.c_float), ("grain", C.c_float), ("susceptibility", C.c_float),
                ("saturation", C.c_float), ("volatility", C.c_float), ("fume", C.c_float),
                ("react_with", C.c_int32), ("react_into", C.c_int32), ("react_rate", C.c_float),
                ("react_heat", C.c_float), ("self_heat", C.c_float), ("residue", C.c_int32),
                ("residue_share", C.c_float)]


class SimParams(C.Structure):
    _fields_ = [("cell_size", C.c_float), ("gravity", C.c_float), ("ambient_temp", C.c_float),
                ("rigid_viscosity", C.c_float), ("pair_tension", (C.c_float * GROUPS) * GROUPS),
                ("droplet_cells", C.c_float), ("cfl", C.c_float), ("advection_order", C.c_uint32),
                ("dt_max", C.c_float), ("fixed_hz", C.c_float), ("pressure_tol", C.c_float),
                ("pressure_max_iters", C.c_uint32), ("correction_tol", C.c_float),
                ("viscosity_tol", C.c_float), ("viscosity_max_iters", C.c_uint32),
                ("implicit_viscosity_min", C.c_float), ("spray_cfl", C.c_float),
                ("speed_cap", C.c_float), ("correction_time", C.c_float),
                ("correction_max", C.c_float), ("correction_deadband", C.c_float),
                ("correction_step_max", C.c_float), ("crowd_jitter", C.c_float),
                ("tool_crowd_max", C.c_float), ("apic_clamp", C.c_float),
                ("curvature_max", C.c_float), ("slip_cap", C.c_uint32), ("sealed_vent", C.c_uint32),
                ("min_substep_share", C.c_float), ("viscous_projection", C.c_uint32),
                ("viscous_warm", C.c_uint32), ("viscosity_warm", C.c_uint32),
                ("viscosity_precond", C.c_uint32), ("viscosity_mg_levels", C.c_uint32),
                ("viscosity_mg_sweeps", C.c_uint32), ("iters_margin", C.c_float),
                ("mg_precision", C.c_uint32),
                ("heat_rate", C.c_float), ("air_cooling", C.c_float), ("contact_heat", C.c_float),
                ("eddy_heat", C.c_float), ("evaporation", C.c_uint32),
                ("rigid_contact", C.c_uint32), ("rigid_friction", C.c_float),
                ("elastic_stress_max", C.c_float), ("elastic_step_max", C.c_float),
                ("magnet_demag", C.c_float),
                ("steam", C.c_uint32), ("steam_boil", C.c_float), ("steam_condense", C.c_float),
                ("steam_mixing", C.c_float), ("steam_rise", C.c_float), ("steam_life", C.c_float),
                ("steam_swirl", C.c_float), ("steam_eddy", C.c_float), ("heat_power", C.c_float),
                ("mat_count", C.c_int32),
                ("mat", Material * MAX_MATERIALS)]


class Magnet(C.Structure):
    """`Magnet` (params.rs:1082-1094): position and radius in cells, strength in tesla on the
    axis at the rim, north pole direction."""
    _fields_ = [("pos", C.c_float * 2), ("radius", C.c_float), ("strength", C.c_float), ("dir", C.c_float * 2)]

    @staticmethod
    def new(pos, radius, degrees, strength):
        """Magnet::new: north pole `degrees` clockwise from straight up."""
        a = np.radians(np.float32(degrees))
        return Magnet((C.c_float * 2)(*pos), radius, strength, (C.c_float * 2)(float(np.sin(a)), float(np.cos(a))))


class Current(C.Structure):
    """`Current` (params.rs:1113-1121): position and radius in cells, the velocity the liquid is
    driven towards (cells/s), rate (1/s; negative: a drain)."""
    _fields_ = [("pos", C.c_float * 2), ("radius", C.c_float), ("vel", C.c_float * 2), ("rate", C.c_float)]

    @staticmethod
    def new(pos, radius, degrees, strength):
        """Current::new (params.rs:1126-1130): flowing `degrees` clockwise from straight down at
        `strength` x CURRENT_SPEED, at CURRENT_RATE."""
        a = np.radians(np.float32(degrees))
        sp = np.float32(CURRENT_SPEED * strength)
        return Current((C.c_float * 2)(*pos), radius, (C.c_float * 2)(float(np.sin(a) * sp), float(-np.cos(a) * sp)),
                       CURRENT_RATE)

    @staticmethod
    def drain(pos, radius, strength):
        """Current::drain (params.rs:1134-1136)."""
        return Current((C.c_float * 2)(*pos), radius, (C.c_float * 2)(0.0, 0.0), -DRAIN_RATE * strength)


class ToolState(C.Structure):
    _fields_ = [("pos", C.c_float * 2), ("vel", C.c_float * 2), ("radius", C.c_float),
                ("mode", C.c_int32), ("strength", C.c_float)]


class SolverStats(C.Structure):
    _fields_ = [("substeps", C.c_uint64), ("pressure_unconverged", C.c_uint64),
                ("viscosity_unconverged", C.c_uint64), ("pressure_iters_max", C.c_uint32),
                ("viscosity_iters_max", C.c_uint32), ("pressure_residual_max", C.c_float),
                ("viscosity_residual_max", C.c_float), ("dt", C.c_float), ("dt_cfl", C.c_float),
                ("umax", C.c_float), ("umax_grid", C.c_float), ("umax_part", C.c_float),
                ("pressure_iters_last", C.c_uint32)]


class RequestInfo(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("scene", C.c_uint32), ("substeps", C.c_uint32),
                ("status", C.c_int32), ("tanks", C.c_uint32), ("submit_ms", C.c_double),
                ("launch_ms", C.c_double), ("done_ms", C.c_double), ("latency_ms", C.c_double)]


class PoolStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("particle_steps", C.c_uint64), ("launches", C.c_uint64), ("evictions", C.c_uint64),
                ("restores", C.c_uint64), ("resident_bytes", C.c_uint64), ("scenes", C.c_uint32),
                ("resident", C.c_uint32), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("span_s", C.c_double),
                ("tanks_per_launch", C.c_double), ("rounds", C.c_uint64)]


class TuneEntry(C.Structure):
    _fields_ = [("grid_w", C.c_int32), ("grid_h", C.c_int32), ("tanks", C.c_int32), ("stride", C.c_int32),
                ("best", C.c_int32), ("best_ms", C.c_float), ("default_ms", C.c_float), ("samples", C.c_uint32)]


class StatsOut(C.Structure):
    _fields_ = [("live", C.c_uint32), ("count", C.c_uint32 * MAX_MATERIALS),
                ("conc", C.c_float * MAX_MATERIALS)]


class GMat(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("group", "diluent", "chan_w", "chan_s", "fam", "rheo",
                                          "temp_model", "residue")] + \
               [(n, C.c_float) for n in ("rho", "ln_nu", "sigma", "mix_rate", "cos_c", "sin_c",
                                          "grain", "rho_solid", "visc_exp", "expansion", "load",
                                          "freeze_shift")] + \
               [("th", C.c_float * 4), ("th2", C.c_float * 4), ("temp", C.c_float * 4),
                ("rheo_p", C.c_float * 4), ("gr", C.c_float * 4), ("rx", C.c_float * 4),
                ("ic", C.c_int32 * 4), ("residue_share", C.c_float), ("fume", C.c_float),
                ("pad_m", C.c_float * 2)]


class GParams(C.Structure):
    _fields_ = [("W", C.c_int32), ("H", C.c_int32), ("n_slots", C.c_uint32),
                ("spawn_count", C.c_uint32), ("dt_fixed", C.c_float), ("gravity", C.c_float),
                ("inv_ppc", C.c_float), ("apic_clamp", C.c_float),
                ("pair", (C.c_float * GROUPS) * GROUPS), ("kappa_max", C.c_float),
                ("sealed_vent", C.c_float), ("cell_size", C.c_float), ("rigid_friction", C.c_float),
                ("n_present", C.c_int32), ("acc_stride", C.c_int32), ("n_fam", C.c_int32),
                ("group_mask", C.c_uint32), ("caps", C.c_uint32), ("primary", C.c_int32),
                ("group_heat", C.c_uint32), ("visc_mode", C.c_int32),
                ("present", C.c_int32 * MAX_PRESENT), ("fams", (C.c_int32 * 2) * MAX_PRESENT),
                ("sep_tau", C.c_float), ("cfl", C.c_float), ("tool_crowd_max", C.c_float),
                ("density_k", C.c_float), ("dt_max", C.c_float), ("crowd_jitter", C.c_float),
                ("ext_speed", C.c_float), ("wall_margin", C.c_float), ("tol", C.c_float * 4),
                ("ext", C.c_float * 4), ("corr", C.c_float * 4), ("p_max_iters", C.c_int32),
                ("v_max_iters", C.c_int32), ("heat_x", C.c_float), ("ln_rigid", C.c_float),
                ("v_twice", C.c_int32), ("v_p2warm", C.c_int32), ("v_warm", C.c_int32),
                ("v_precond", C.c_int32), ("v_mg_levels", C.c_int32), ("v_mg_sweeps", C.c_int32),
                ("mg_half", C.c_int32), ("heat_rate", C.c_float), ("film", C.c_float),
                ("radiation", C.c_float), ("contact", C.c_float), ("eddy", C.c_float), ("evap", C.c_float),
                ("chan_t", C.c_int32), ("chan_e", C.c_int32), ("el_g", C.c_float),
                ("el_stress_max", C.c_float), ("el_step_max", C.c_float), ("mag", C.c_float * 4),
                ("magnet", (C.c_float * 4) * MAX_MAGNETS), ("magnet_dir", (C.c_float * 2) * MAX_MAGNETS),
                ("demag", C.c_float), ("current_zone", (C.c_float * 4) * MAX_CURRENTS),
                ("current_vel", (C.c_float * 4) * MAX_CURRENTS), ("steam", C.c_float * 4),
                ("steam2", C.c_float * 4), ("tool_pos", C.c_float * 2), ("tool_vel", C.c_float * 2),
                ("tool_radius", C.c_float), ("tool_mode", C.c_int32), ("tool_strength", C.c_float),
                ("heat_power", C.c_float)]


class DeviceView(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("bytes", C.c_uint64), ("count", C.c_uint32), ("stride", C.c_uint32),
                ("device", C.c_int32), ("pad", C.c_uint32), ("generation", C.c_uint64)]


RENDER = dict(PARTICLES=0, SDF=1, CELL=2, COMPOSITION=3, SURFACE=4, MATERIALS=5, VELOCITY=6, GPUMAT=7)


class LevelsInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("tail", C.c_int32), ("total", C.c_int32), ("tiles", C.c_int32),
                ("dims", (C.c_int32 * 2) * MAX_LEVELS), ("offs", C.c_int32 * MAX_LEVELS),
                ("tile_offs", C.c_int32 * MAX_LEVELS)]


class Config(C.Structure):
    _fields_ = [("grid_w", C.c_int32), ("grid_h", C.c_int32), ("capacity", C.c_uint32),
                ("device", C.c_int32)]


STEP_DTYPE = np.dtype([("frame_left", "<f4"), ("advanced", "<f4"), ("batch_steps", "<u4"),
                       ("batch_done", "<u4"), ("dt", "<f4"), ("dt_prev", "<f4"),
                       ("umax_grid", "<u4"), ("umax_part", "<u4"), ("dt_cfl", "<f4"),
                       ("substeps", "<u4"), ("umax_seen", "<f4"), ("pad0", "<u4")])
# sp_cg_state: the pressure CG's scalars entering an iteration (two, by parity).
CG_STATE_DTYPE = np.dtype([("rz", "<f8", (2,)), ("rr", "<f8", (2,)), ("bb", "<f8", (2,)), ("beta", "<f8", (2,)),
                           ("alpha", "<f8", (2,)), ("done", "<u4", (2,)), ("iters", "<u4", (2,)), ("sweeps", "<u4"),
                           ("halt", "<u4"), ("cap", "<u4"), ("pad", "<u4")])
assert CG_STATE_DTYPE.itemsize == 112
SOLVE_DTYPE = np.dtype([("rz", "<f8", (2,)), ("alpha", "<f8", (2,)), ("beta", "<f8", (2,)),
                        ("bb", "<f8", (2,)), ("rr", "<f8", (2,)), ("done", "<u4", (2,)),
                        ("iters", "<u4", (2,)), ("it_max", "<u4"), ("unconv", "<u4"),
                        ("solves", "<u4"), ("arrive", "<u4"), ("res_max", "<f4"),
                        ("iters_last", "<u4"), ("p2_it_max", "<u4"), ("p2_unconv", "<u4"),
                        ("v_rz", "<f8"), ("v_alpha", "<f8"), ("v_beta", "<f8"), ("v_r0", "<f8"),
                        ("v_rr", "<f8"), ("v_pq", "<f8"), ("v_done", "<u4"), ("v_iters", "<u4"),
                        ("v_it_max", "<u4"), ("v_unconv", "<u4"), ("v_res_max", "<f4"),
                        ("v_iters_last", "<u4"), ("p2_iters_last", "<u4"), ("iters_sum", "<u4"),
                        ("cap_p", "<u4"), ("cap_p2", "<u4"), ("cap_v", "<u4"), ("cap_now", "<u4"),
                        ("sweeps", "<u4"), ("halt", "<u4"), ("v_sweeps", "<u4"), ("v_halt", "<u4"),
                        ("st", CG_STATE_DTYPE, (2,))])
assert SOLVE_DTYPE.itemsize == 464

# sp_gpumat: a row of the reference's GpuMat table (material.rs:508-549), 208 bytes.
GPUMAT_DTYPE = np.dtype([("ia", "<i4", (4,)), ("ib", "<i4", (4,)), ("a", "<f4", (4,)), ("b", "<f4", (4,)),
                         ("rheo", "<f4", (4,)), ("temp", "<f4", (4,)), ("th", "<f4", (4,)), ("th2", "<f4", (4,)),
                         ("el", "<f4", (4,)), ("misc", "<f4", (4,)), ("ic", "<i4", (4,)), ("rx", "<f4", (4,)),
                         ("gr", "<f4", (4,))])
assert GPUMAT_DTYPE.itemsize == 208


class PassTime(C.Structure):
    _fields_ = [("name", C.c_char * 24), ("ms", C.c_double)]


# Exported symbols of include/sprout_hip.h, with their argument types.
_P = C.c_void_p
SIGNATURES = {
    "sp_create": ([C.POINTER(Config), C.POINTER(C.c_void_p)], C.c_int),
    "sp_destroy": ([_P], None),
    "sp_last_error": ([_P], C.c_char_p),
    "sp_create_error": ([], C.c_char_p),
    "sp_set_params": ([_P, C.POINTER(SimParams)], C.c_int),
    "sp_set_solids": ([_P, _P, _P, _P, _P], C.c_int),
    "sp_reset": ([_P, _P, C.c_uint32], C.c_int),
    "sp_spawn": ([_P, _P, C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "sp_record": ([_P, C.c_uint32, C.c_float, C.POINTER(ToolState)], C.c_int),
    "sp_wait": ([_P], C.c_int),
    "sp_poll": ([_P, C.POINTER(SolverStats)], C.c_int),
    "sp_advance_blocking": ([_P, C.c_double, C.POINTER(ToolState)], C.c_int),
    "sp_step_blocking": ([_P, C.c_uint32, C.POINTER(ToolState)], C.c_int),
    "sp_sim_time": ([_P, C.POINTER(C.c_double)], C.c_int),
    "sp_frame_blocking": ([_P, C.c_uint32, C.c_float, C.POINTER(ToolState), C.POINTER(C.c_double)], C.c_int),
    "sp_slots": ([_P, C.POINTER(C.c_uint32)], C.c_int),
    "sp_dt_estimate": ([_P, C.POINTER(C.c_float)], C.c_int),
    "sp_recorded_iters": ([_P, _P], C.c_int),
    "sp_substep_cost": ([_P, C.POINTER(C.c_float)], C.c_int),
    "sp_profile_passes": ([_P, C.POINTER(PassTime), C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "sp_material_table": ([_P, _P, C.c_uint32], C.c_int),
    "sp_download_labels": ([_P, _P], C.c_int),
    "sp_download_bodies": ([_P, _P, C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "sp_download_particles": ([_P, _P, C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "sp_download_grid": ([_P, _P, _P], C.c_int),
    "sp_download_pressure": ([_P, _P, C.POINTER(C.c_float)], C.c_int),
    "sp_cell_counts": ([_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P], C.c_int),
    "sp_stats": ([_P, C.POINTER(StatsOut)], C.c_int),
    "sp_sort": ([_P], C.c_int),
    "sp_bin_particles": ([_P, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P], C.c_int),
    "sp_bin_particles_noalloc": ([_P, C.c_uint32, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P], C.c_int),
    "sp_error_text": ([], C.c_char_p),
    "sp_render_view": ([_P, C.c_int32, _P], C.c_int),
    "sp_render_download": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "sp_stream": ([_P, C.POINTER(C.c_void_p)], C.c_int),
    "sp_resets": ([_P, C.POINTER(C.c_uint64)], C.c_int),
    "sp_shares_a_group": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "sp_default_params": ([C.POINTER(SimParams)], None),
    "sp_derive_grid_params": ([C.POINTER(SimParams), C.c_int32, C.c_int32, _P, C.c_int32, C.c_int32,
                               _P, C.c_uint32, _P, C.c_uint32, C.POINTER(GParams), _P], C.c_int),
    "sp_set_magnets": ([_P, _P, C.c_uint32], C.c_int),
    "sp_set_currents": ([_P, _P, C.c_uint32], C.c_int),
    "sp_download_ledger_counts": ([_P, _P, _P, _P], C.c_int),
    "sp_levels": ([C.c_int32, C.c_int32, C.POINTER(LevelsInfo)], C.c_int),
    "sp_debug_download": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "sp_debug_upload": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "sp_debug_stage": ([_P, C.c_int32, C.c_int32], C.c_int),
    "sp_debug_gparams": ([_P, C.POINTER(GParams), _P], C.c_int),
    "sp_timings": ([_P, _P], C.c_int),
    "sp_kernel_time": ([_P, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)], C.c_int),
    "sp_batch_create": ([_P, C.c_uint32, C.POINTER(C.c_void_p)], C.c_int),
    "sp_batch_destroy": ([_P], None),
    "sp_batch_last_error": ([_P], C.c_char_p),
    "sp_batch_record": ([_P, C.c_uint32, C.c_float], C.c_int),
    "sp_batch_wait": ([_P], C.c_int),
    "sp_batch_groups": ([_P, C.POINTER(C.c_uint32), _P, C.c_uint32], C.c_int),
    "sp_batch_timings": ([_P, C.c_uint32, _P], C.c_int),
    "sp_download_ledger": ([_P, _P], C.c_int),
    "sp_evict": ([_P], C.c_int),
    "sp_is_resident": ([_P, C.POINTER(C.c_int32)], C.c_int),
    "sp_tank_bytes": ([C.c_int32, C.c_int32, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)], C.c_int),
    "sp_autotune": ([C.c_int32, C.c_uint32], C.c_int),
    "sp_autotune_shapes": ([_P, _P, _P, C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "sp_autotune_report": ([C.POINTER(TuneEntry), C.c_uint32, C.POINTER(C.c_uint32)], C.c_int),
    "sp_pool_create": ([C.c_int32, C.c_uint64, C.c_uint32, C.POINTER(C.c_void_p)], C.c_int),
    "sp_pool_destroy": ([_P], None),
    "sp_pool_last_error": ([_P], C.c_char_p),
    "sp_pool_add": ([_P, _P, C.POINTER(C.c_uint32)], C.c_int),
    "sp_pool_submit": ([_P, C.c_uint32, C.c_uint32, C.c_float, C.POINTER(C.c_uint64)], C.c_int),
    "sp_pool_step": ([_P, C.c_double, C.POINTER(C.c_uint32)], C.c_int),
    "sp_pool_poll": ([_P, C.c_uint64, C.POINTER(RequestInfo)], C.c_int),
    "sp_pool_wait": ([_P, C.c_uint64, C.POINTER(RequestInfo)], C.c_int),
    "sp_pool_drain": ([_P], C.c_int),
    "sp_pool_stats_get": ([_P, C.POINTER(PoolStats)], C.c_int),
    "sp_pool_reset_stats": ([_P], C.c_int),
}

_lib = None


def lib():
    """The loaded libsprout_hip.so. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run sprout_amd.build.build() (hipcc, gfx950) first; "
                               "the fluid step has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


_hip = None


def hip():
    """The HIP runtime libsprout_hip.so links (the same loaded library, by its soname), for
    device buffers handed to the handle-free entry points. Another copy of the runtime in
    the process (a framework's bundled one) would see no devices once this one holds them."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        H.hipDeviceSynchronize.argtypes = []
        for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipMemset, H.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = H
    return _hip


class DeviceBuffer:
    """`nbytes` of device memory on the current device (hipMalloc), with host copies in and
    out; freed on close()."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), max(self.nbytes, 1))
        if rc != 0:
            raise SproutError(E_NOMEM, f"hipMalloc({self.nbytes}) failed with HIP error {rc}")
        self.ptr = p.value
        hip().hipMemset(self.ptr, 0, max(self.nbytes, 1))

    @classmethod
    def of(cls, arr):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        if arr.nbytes and hip().hipMemcpy(self.ptr, arr.ctypes.data, arr.nbytes, 1) != 0:
            raise SproutError(E_DEVICE, "hipMemcpy to the device failed")

    def download(self, dtype, count):
        out = np.zeros(count, dtype)
        assert out.nbytes <= self.nbytes
        if out.nbytes and hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:
            raise SproutError(E_DEVICE, "hipMemcpy from the device failed")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bin_particles(dev_in, n, W, H, tiled, dev_out, dev_perm=None, stream=None):
    """sp_bin_particles over device buffers: the stable counting sort of sort.comp."""
    L = lib()
    rc = L.sp_bin_particles(dev_in, n, W, H, tiled, dev_out, dev_perm, stream)
    if rc != OK:
        raise SproutError(rc, f"bin_particles: {L.sp_error_text().decode()}")


def tank_bytes(W, H, capacity):
    """sp_tank_bytes: (evictable, fixed) device bytes of a tank of this grid and capacity."""
    ev, fx = C.c_uint64(), C.c_uint64()
    rc = lib().sp_tank_bytes(W, H, capacity, C.byref(ev), C.byref(fx))
    if rc != OK:
        raise SproutError(rc, "tank_bytes")
    return ev.value, fx.value


def autotune(mode=-1, samples=0):
    """sp_autotune: -1 tune the particle pass's launch shape on the launches (default), -2
    forget and tune again, k >= 0 use shape k everywhere; samples > 0: timings per shape."""
    rc = lib().sp_autotune(int(mode), int(samples))
    if rc != OK:
        raise SproutError(rc, "autotune")


def autotune_shapes():
    """The candidate shapes: [(threads per group, LDS scatter-box ints, groups per compute
    unit the registers are limited for), ...]."""
    n = C.c_uint32()
    t = np.zeros(16, np.int32)
    b = np.zeros(16, np.int32)
    m = np.zeros(16, np.int32)
    rc = lib().sp_autotune_shapes(t.ctypes.data, b.ctypes.data, m.ctypes.data, 16, C.byref(n))
    if rc != OK:
        raise SproutError(rc, "autotune_shapes")
    return [(int(t[k]), int(b[k]), int(m[k])) for k in range(min(n.value, 16))]


def autotune_report():
    """The tuned keys: dicts of grid, tanks, stride, best shape (-1 while tuning), the
    median particle-pass ms of the best shape and of shape 0, and the timings taken."""
    n = C.c_uint32()
    rc = lib().sp_autotune_report(None, 0, C.byref(n))
    if rc != OK:
        raise SproutError(rc, "autotune_report")
    arr = (TuneEntry * max(n.value, 1))()
    rc = lib().sp_autotune_report(arr, n.value, C.byref(n))
    if rc != OK:
        raise SproutError(rc, "autotune_report")
    return [{name: getattr(arr[k], name) for name, _ in TuneEntry._fields_} for k in range(n.value)]


class SproutError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sprout_hip error {code}: {msg}")
        self.code = code


def default_params():
    p = SimParams()
    lib().sp_default_params(C.byref(p))
    return p


def derive_grid_params(params, W, H, rows, thermal=False, magnets=(), currents=()):
    """sp_derive_grid_params: the grid-unit block and material table for a tank of `rows`
    with `magnets` (Magnet records) placed and `currents` (Current records)."""
    g = GParams()
    mats = (GMat * MAX_MATERIALS)()
    arr = (C.c_int32 * max(len(rows), 1))(*rows)
    mg = (Magnet * max(len(magnets), 1))(*magnets)
    cu = (Current * max(len(currents), 1))(*currents)
    rc = lib().sp_derive_grid_params(C.byref(params), W, H, arr, len(rows), int(thermal), mg, len(magnets),
                                     cu, len(currents), C.byref(g), mats)
    if rc != OK:
        raise SproutError(rc, "derive_grid_params")
    return g, mats


def levels(W, H):
    L = LevelsInfo()
    rc = lib().sp_levels(W, H, C.byref(L))
    if rc != OK:
        raise SproutError(rc, "levels")
    return L
