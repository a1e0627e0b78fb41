/*
 * sprout_types.h — plain-data types of the Sprout fluid step shared by the C ABI, the HIP
 * kernels and the CPU oracle.
 *
 * Units follow the reference (shaders/common.glsl:3-9): lengths in grid cells, time in
 * seconds, densities relative to water. Grid layout is the MAC layout of the reference:
 * u on the left face (i, j+0.5), v on the bottom face (i+0.5, j), scalars at the centre,
 * row-major cell index c = y*W + x.
 */
#ifndef SPROUT_TYPES_H
#define SPROUT_TYPES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_GROUPS 4          /* material.rs: GROUPS */
#define SP_MAX_PRESENT 12    /* material.rs: MAX_PRESENT */
#define SP_MAX_MATERIALS 96  /* material.rs: MAX_MATERIALS */
#define SP_PPC 4             /* params.rs:11 PPC */
#define SP_ACC_MIN 8         /* mod.rs:68 ACC_MIN */
#define SP_ACC_MAX 32        /* mod.rs:69 ACC_MAX */
#define SP_MAX_MAGNETS 8     /* params.rs:1071 MAX_MAGNETS */
#define SP_ACC_FIRST 5       /* mod.rs:71 ACC_FIRST */
#define SP_REGION_MASK 31u   /* common.glsl:220 */
#define SP_REGION_VENT 256u  /* common.glsl:221 */
#define SP_SPAWN_MAX 32768   /* mod.rs:63 SPAWN_MAX */

/* Particle kinds (params.rs:21-31). */
#define SP_KIND_DEAD 0u
#define SP_KIND_FRESH 0x10000u
#define SP_KIND_REAL_T 0x20000u
#define SP_KIND_PUFF 0x40000u
#define SP_KIND_SRC 0xff00u

/* Material rows carried by this build, at the reference's table rows
 * (src/sim/material.rs:35-96). */
#define SP_ROW_WATER 0
#define SP_ROW_HONEY 1
#define SP_ROW_OIL 2
#define SP_ROW_MERCURY 8
#define SP_ROW_BRINE 14   /* material.rs: Material::Brine */
#define SP_ROW_SALT 56    /* material.rs: Material::Salt, what boiling brine leaves */
#define SP_ROW_STEAM 57   /* material.rs: Material::Steam */
#define SP_BUILTIN_COUNT 59  /* presets.rs:1130-1135: REAL_COUNT + SIMPLE_COUNT + 3 */

/* `struct Particle` (common.glsl:11-19, params.rs:37-54): 48 bytes, std430. */
typedef struct sp_particle {
    float pos[2];   /* cells */
    float vel[2];   /* cells / s */
    float c[4];     /* APIC affine matrix du/dx, du/dy, dv/dx, dv/dy (1/s) */
    float s;        /* concentration */
    uint32_t kind;  /* 0 dead, else table row + 1 in the low byte, flags above */
    uint32_t id;    /* permanent identity */
    float t;        /* enthalpy temperature, deg C */
} sp_particle;

/* A material row in SI units: the fields of `MatProps` (src/sim/material.rs:159-340) that
 * the isothermal step reads. */
typedef struct sp_material {
    int32_t group;         /* phase group 0 aqueous, 1 oily, 2 metal, 3 silicate */
    int32_t diluent;       /* row of the diluent, -1 none */
    float density;         /* kg/m^3 */
    float density_solid;   /* kg/m^3 when frozen, 0 = as liquid */
    float viscosity;       /* kinematic, m^2/s, at temp_ref */
    int32_t temp_model;    /* 0 none, 1 Arrhenius, 2 VFT (material.rs TEMP_*) */
    float temp_b, temp_c;  /* temperature law (K) */
    float temp_ref;        /* deg C */
    float tension;         /* against air, N/m */
    float contact_angle;   /* degrees */
    float mix_rate;        /* 1/s */
    float expansion;       /* 1/K */
    float melt_lo, melt_hi;/* deg C (-1e9: never freezes) */
    float visc_exponent;   /* viscosity blend exponent against the diluent (visc_exponent) */
    float load;            /* emitted concentration of a concentrate */
    float spawn_temp;      /* deg C */
    float latent;          /* J/kg */
    float heat_capacity;   /* J/(kg K) */
    float conductivity;    /* W/(m K) */
    float emissivity;      /* grey body, at the free surface */
    float boil;            /* deg C (1e9: never boils) */
    float latent_vap;      /* J/kg */
    float freeze_shift;    /* K per unit concentration (a concentrate freezes like its diluent, shifted) */
    /* Rheology (material.rs:146-182, 239-240, 311-314): model (0 Newtonian, 1 Carreau,
     * 2 Herschel-Bulkley, 3 shear thickening), yield stress (Pa), flow index n, shear time
     * lambda (s), viscosity ratio (Carreau: infinite-shear / zero-shear; thickening: the
     * cap); a surface skin (N/m) and the internal friction of a granular liquid. */
    uint32_t rheology;
    float yield_stress, flow_index, shear_time, visc_ratio;
    float skin, friction;
    /* Viscoelasticity (material.rs:233-235): the shear modulus G (Pa; 0 = not elastic) and
     * the relaxation time lambda (s) of an upper-convected Maxwell liquid. */
    float shear_modulus, relax_time;
    /* Suspensions (material.rs:198-200): the diameter of the grains of a concentrate (m), which
     * settle (or cream) through the diluent at their terminal velocity; 0 = dissolved. */
    float grain;
    /* Magnetism (material.rs:236-238): the volume susceptibility (SI; 0 = not magnetic) and
     * the saturation magnetisation as mu0 Ms (T). */
    float susceptibility, saturation;
    /* Evaporation below the boiling point (material.rs:307-310, 326-327): the share of the
     * liquid at an open surface that evaporates per second at 20 deg C (doubling every 10 K;
     * 0 = none), and how much visible vapour it gives off where it evaporates or boils away (a
 * puff of mist, kind SP_KIND_PUFF; 0 = none). */
    float volatility, fume;
    /* Reactions and heat made inside (material.rs:305-320): in a cell that also holds row
     * `react_with` (-1: none) the liquid turns into row `react_into` (-1: it is used up) at
     * `react_rate` per second, giving off `react_heat` J/kg; `self_heat` W/kg of its own. */
    int32_t react_with, react_into;
    float react_rate, react_heat, self_heat;
    /* Boiling (material.rs:221-225): the table row a liquid leaves behind when it boils away
     * (-1: none) and the volume share of its concentrate that becomes that residue. */
    int32_t residue;
    float residue_share;
} sp_material;

/* `SimParams` (src/sim/params.rs:149-376), the fields the isothermal step reads, SI units. */
typedef struct sp_sim_params {
    float cell_size;            /* m */
    float gravity;              /* m/s^2 */
    float ambient_temp;         /* deg C */
    float rigid_viscosity;      /* m^2/s */
    float pair_tension[SP_GROUPS][SP_GROUPS];  /* N/m */
    float droplet_cells;
    float cfl;
    uint32_t advection_order;
    float dt_max;
    float fixed_hz;
    float pressure_tol;
    uint32_t pressure_max_iters;
    float correction_tol;
    float viscosity_tol;
    uint32_t viscosity_max_iters;
    float implicit_viscosity_min;
    float spray_cfl;
    float speed_cap;
    float correction_time;
    float correction_max;
    float correction_deadband;
    float correction_step_max;
    float crowd_jitter;
    float tool_crowd_max;
    float apic_clamp;
    float curvature_max;
    uint32_t slip_cap;
    uint32_t sealed_vent;
    float min_substep_share;
    uint32_t viscous_projection;   /* 1: a second pressure solve after the implicit viscosity */
    uint32_t viscous_warm;         /* 1: each pressure solve starts from its own part of the last */
    uint32_t viscosity_warm;       /* 1: the viscosity solve starts from the last substep's change */
    uint32_t viscosity_precond;    /* 1: multigrid V-cycle, 0: Jacobi */
    uint32_t viscosity_mg_levels;  /* levels below the grid in the viscosity V-cycle */
    uint32_t viscosity_mg_sweeps;  /* Gauss-Seidel sweeps per visit of a level (1 or 2) */
    /* Pressure solves record this share more iterations than the last batch needed
     * (params.rs:284, range [0, 2], default 0.25; mod.rs:3098-3114). */
    float iters_margin;
    /* Storage of the pressure multigrid (no counterpart in the reference, which keeps FP32):
     * 0 FP32; 1 mixed: the V-cycle's iterates and coarse right-hand sides in BF16, the
     * coarse levels' matrices in FP16, smoothing arithmetic in FP32; the level-0 matrix
     * (also the CG operator) and the CG vectors stay FP32, the reductions FP64. */
    uint32_t mg_precision;
    /* Heat (params.rs:168-193): conduction speed-up over real time (0: no heat pass), the
     * air's film coefficient (W/(m^2 K)), the heat transfer coefficient between different
     * liquids (W/(m^2 K)), the Smagorinsky constant of the sub-grid mixing (0 off), and
     * evaporative cooling of aqueous surfaces (1 on). */
    float heat_rate;
    float air_cooling;
    float contact_heat;
    float eddy_heat;
    uint32_t evaporation;
    /* Rigid bodies of frozen or too viscous liquid (params.rs:198-208): 1 = a body resting on
     * solids is supported by its contacts, with this Coulomb friction. */
    uint32_t rigid_contact;
    float rigid_friction;
    /* Viscoelastic stabilisation (params.rs:366-371): the largest stress, times G (0: no
     * cap), and the largest deformation per substep that enters the stress (0: no limit). */
    float elastic_stress_max;
    float elastic_step_max;
    /* Magnetism (params.rs:218): how much of a local shape demagnetisation is blended in
     * (0 = a sphere's, N = 1/3, everywhere). */
    float magnet_demag;
    /* Steam (params.rs:219-249): 1 = boiling aqueous liquid turns into bubbles of steam that
     * rise, condense in colder liquid and escape at the surface as puffs (0: boiled liquid
     * leaves at once); the vapour quality at which boiling water becomes steam and below which
     * steam condenses; the exchange rate among the particles of a cell holding steam (1/s);
     * the escaped puffs' rise speed (m/s), lifetime (s), and the speed (m/s) and size (m) of
     * the air's eddies they ride. */
    uint32_t steam;
    float steam_boil, steam_condense, steam_mixing, steam_rise, steam_life, steam_swirl, steam_eddy;
    /* The heat and cool tools' power (params.rs:210): kelvin per second for water at the
     * brush's centre. */
    float heat_power;
    int32_t mat_count;
    sp_material mat[SP_MAX_MATERIALS];
} sp_sim_params;

/* `Magnet` (params.rs:1082-1094): a permanent magnet, the field of a point dipole at `pos`
 * (cells) in the plane of the tank, `strength` tesla on its axis at distance `radius` (cells),
 * north pole towards `dir` (unit vector; Magnet::new(pos, radius, degrees, strength) gives
 * (sin, cos) of the angle clockwise from straight up). */
typedef struct sp_magnet {
    float pos[2];
    float radius;
    float strength;
    float dir[2];
} sp_magnet;

/* `Current` (src/sim/params.rs:1113-1121): a zone of the tank, position and radius in cells,
 * that drives the liquid in it towards `vel` (cells/s) at `rate` (1/s); a negative rate makes
 * it a drain, which removes the liquid inside it at minus that rate per second. */
#define SP_MAX_CURRENTS 16  /* params.rs:1069 MAX_CURRENTS */
typedef struct sp_current {
    float pos[2];
    float radius;
    float vel[2];
    float rate;
} sp_current;

/* `ToolState` (params.rs:1141-1150). */
typedef struct sp_tool_state {
    float pos[2];
    float vel[2];
    float radius;
    int32_t mode;
    float strength;
} sp_tool_state;

/* `SolverStats` (src/sim/mod.rs:334-353). */
typedef struct sp_solver_stats {
    uint64_t substeps;
    uint64_t pressure_unconverged;
    uint64_t viscosity_unconverged;
    uint32_t pressure_iters_max;
    uint32_t viscosity_iters_max;
    float pressure_residual_max;
    float viscosity_residual_max;
    float dt, dt_cfl, umax, umax_grid, umax_part;
    uint32_t pressure_iters_last;   /* iterations of the latest solve (this build) */
} sp_solver_stats;

/* `Stats` (src/sim/mod.rs:128-143). */
typedef struct sp_stats_out {
    uint32_t live;
    uint32_t count[SP_MAX_MATERIALS];
    float conc[SP_MAX_MATERIALS];
} sp_stats_out;

/* One row of the reference's GPU material table in its own 208-byte layout (`GpuMat`,
 * src/sim/material.rs:508-549, filled by gpu_materials, src/sim/mod.rs:2179-2236), for a
 * renderer that reads `material_buffer`. The kernels of this library read `sp_gmat`; this
 * table is the same numbers in the reference's order. The one field this build does not
 * carry (fizz, the renderer's bubbles) is 0. */
typedef struct sp_gpumat {
    int32_t ia[4];   /* phase group, diluent row, weight channel, concentration channel */
    int32_t ib[4];   /* family, rheology model, temperature model, residue row */
    float a[4];      /* relative density, ln nu (cells^2/s), tension / rho_w, mixing rate */
    float b[4];      /* cos, sin of the contact angle, grain diameter (cells), relative density frozen */
    float rheo[4];   /* yield stress / rho_w (cells^2/s^2), flow index, shear time, viscosity ratio */
    float temp[4];   /* temperature law B, C (K), reference temperature (K), skin / rho_w */
    float th[4];     /* heat capacity vs water, conductivity, solidus, liquidus */
    float th2[4];    /* latent heat as degrees, emissivity, boiling point, vaporisation as degrees */
    float el[4];     /* shear modulus / rho_w, relaxation time, susceptibility, residue share */
    float misc[4];   /* freeze shift, load, viscosity blending exponent, expansion */
    int32_t ic[4];   /* reaction partner row, product row, partner's index among the rows, 0 */
    float rx[4];     /* reaction rate, its heat as degrees, self-heating (K/s), volatility */
    float gr[4];     /* friction, fume, fizz, 0 */
} sp_gpumat;

/* One entry of sp_profile_passes: a phase of the substep and its device time. */
typedef struct sp_pass_time {
    char name[24];
    double ms;
} sp_pass_time;

/* One material row in grid units (`GpuMat`, material.rs:508-547, filled at mod.rs:
 * 2179-2236), the fields the isothermal step reads. */
typedef struct sp_gmat {
    int32_t group, diluent, chan_w, chan_s;   /* ia */
    int32_t fam, rheo, temp_model, residue;   /* ib */
    float rho, ln_nu, sigma, mix_rate;        /* a: relative density, ln nu (cells^2/s), tension/rho_w, 1/s */
    float cos_c, sin_c, grain, rho_solid;     /* b: grain = its diameter in cells */
    float visc_exp, expansion, load, freeze_shift;  /* misc */
    float th[4];     /* volumetric heat capacity / water's; conductivity (cells^2/s as water's); solidus, liquidus (deg C) */
    float th2[4];    /* latent heat as degrees (L / c); emissivity; boiling point (deg C); latent heat of vaporisation as degrees */
    float temp[4];   /* temperature law: B, C (K); reference temperature (K); skin (cells^3/s^2) */
    float rheo_p[4]; /* yield stress (cells^2/s^2), flow index, shear time (s), viscosity ratio */
    float gr[4];     /* internal friction; shear modulus (cells^2/s^2, GpuMat el.x); relaxation time (s, el.y); susceptibility (el.z) */
    float rx[4];     /* GpuMat rx: reaction rate (1/s), heat of reaction (K of the liquid), self-heating (K/s), volatility (1/s) */
    int32_t ic[4];   /* GpuMat ic: the partner row while it is in the tank (-1: no reaction), the product row, the partner's index among the tank's rows, 0 */
    float residue_share;  /* GpuMat el.w: volume share of the concentrate left as residue */
    float fume;           /* GpuMat gr.y: vapour of a puff it leaves (0: none) */
    float pad_m[2];
} sp_gmat;

/* The grid-unit parameter block (`GpuParams`, params.rs:1215-1302, filled at mod.rs:
 * 2238-2330): the fields the isothermal step reads, plus the solve set-up. */
typedef struct sp_gparams {
    int32_t W, H;
    uint32_t n_slots, spawn_count;
    float dt_fixed, gravity, inv_ppc, apic_clamp;
    float pair[SP_GROUPS][SP_GROUPS];
    float kappa_max, sealed_vent, cell_size, rigid_friction;   /* misc */
    int32_t n_present, acc_stride, n_fam;
    uint32_t group_mask;
    uint32_t caps;
    int32_t primary;
    uint32_t group_heat;
    int32_t visc_mode;             /* 0 none, 1 explicit, 2 implicit (viscosity_mode) */
    int32_t present[SP_MAX_PRESENT];
    int32_t fams[SP_MAX_PRESENT][2];
    float sep_tau, cfl, tool_crowd_max, density_k;
    float dt_max, crowd_jitter, ext_speed, wall_margin;
    float tol[4];    /* pressure, correction, viscosity, correction time */
    float ext[4];    /* speed cap (cells/s), advection order, spray cfl, currents */
    float corr[4];   /* correction step max, dead band, slip cap, implicit viscosity min */
    int32_t p_max_iters, v_max_iters;
    float heat_x;    /* ambient temperature (P.heat.x) */
    float ln_rigid;  /* ln rigid viscosity, cells^2/s (P.heat2.x) */
    /* The implicit viscosity solve (visc_mode 2), from SimParams: */
    int32_t v_twice;      /* viscous_projection: pressure, viscosity, pressure again */
    int32_t v_p2warm;     /* viscous_warm: P1 and P2 warm-start their own solves */
    int32_t v_warm;       /* viscosity_warm */
    int32_t v_precond;    /* viscosity_precond: 1 multigrid, 0 Jacobi */
    int32_t v_mg_levels;  /* levels below the grid in its V-cycle (1 .. tail) */
    int32_t v_mg_sweeps;  /* 1 or 2 */
    int32_t mg_half;      /* mg_precision: 1 mixed-precision pressure multigrid */
    /* Heat (GpuParams heat, heat2, heat3; mod.rs:2299-2321), used where caps has CAP_THERMAL: */
    float heat_rate;      /* heat.y: conduction speed-up (0: no heat pass) */
    float film;           /* heat.z: air film coefficient, cells / s */
    float radiation;      /* heat.w: Stefan-Boltzmann constant over water's heat per cell, cells / s / K^3 */
    float contact;        /* heat2.w: heat transfer coefficient between different liquids, cells / s */
    float eddy;           /* heat3.x: Smagorinsky constant of the sub-grid mixing (0 off) */
    float evap;           /* heat3.y: L_vap / c_p,air (K; 0: no evaporation) */
    int32_t chan_t;       /* heat2.z: accumulator channel of the temperature sum (-1: isothermal) */
    /* Viscoelasticity (GpuParams elastic, mod.rs:2316-2321), used where caps has CAP_ELASTIC: */
    int32_t chan_e;       /* elastic.x: first of the three accumulator channels of the stress (-1: none) */
    float el_g;           /* elastic.y: the largest shear modulus in the tank (cells^2/s^2, at least 1e-3 Pa) */
    float el_stress_max;  /* elastic.z: most stress, times G (0: no cap) */
    float el_step_max;    /* elastic.w: most deformation per substep (0: no limit) */
    /* Magnetism (GpuParams mag, magnet, magnet_dir, heat3.w; mod.rs:2258-2264, 2313), used
     * where caps has CAP_MAGNET: */
    float mag[4];         /* magnets; largest susceptibility in the tank (at least 1e-6); its saturation mu0 Ms (T); 1 / (rho_w dx^2) */
    float magnet[SP_MAX_MAGNETS][4];      /* position (cells), radius (cells, at least 0.5), strength (T at the rim) */
    float magnet_dir[SP_MAX_MAGNETS][2];  /* north pole direction */
    float demag;          /* heat3.w: magnet_demag in [0, 1] */
    /* currents and drains (mod.rs:2265-2270): position (cells), radius (at least 0.01 cells),
     * rate (1/s; negative: a drain); the velocity driven towards (cells/s). ext[3] counts them. */
    float current_zone[SP_MAX_CURRENTS][4];
    float current_vel[SP_MAX_CURRENTS][4];
    /* Steam (GpuParams steam, steam2; mod.rs:2322-2323): the quality below which steam
     * condenses, the puffs' rise (cells/s), their lifetime (s), the quality at which boiling
     * water becomes steam; the exchange rate in a cell holding steam (1/s), 1 if boiling makes
     * steam, the air's eddy speed (cells/s) and size (cells). */
    float steam[4];
    float steam2[4];
    /* The mouse tool of the record (GpuParams tool_*, mod.rs:2300-2305): position (cells),
     * velocity (cells/s), radius (cells, at least 0.01), mode (0 idle, 1 grab / stir, 2 erase,
     * 3 push / pull, 4 stream nozzle, 5 heat, 6 cool, 7 whirl, 8 hold), strength (1/s); the
     * heat and cool tools' power (P.heat2.y, K/s for water). */
    float tool_pos[2], tool_vel[2];
    float tool_radius;
    int32_t tool_mode;
    float tool_strength, heat_power;
} sp_gparams;

/* Step state (`StepBuf`, common.glsl:122-135): 12 words. */
typedef struct sp_stepbuf {
    float frame_left;
    float advanced;
    uint32_t batch_steps;
    uint32_t batch_done;
    float dt;
    float dt_prev;
    uint32_t umax_grid;  /* float bits */
    uint32_t umax_part;  /* float bits */
    float dt_cfl;
    uint32_t substeps;
    float umax_seen;
    uint32_t pad0;       /* float bits: Kahan compensation of `advanced` */
} sp_stepbuf;

/* The scalar state of a pressure CG in one iteration (reduce_tail.glsl:59-116's rz, rr, bb,
 * alpha, beta, done and counts), FP64 scalars. The solve keeps two, by iteration parity:
 * iteration k's state is worked out from iteration k - 1's into the other slot, so that
 * nothing reads a slot while it is written. */
typedef struct sp_cg_state {
    double rz[2], rr[2], bb[2], beta[2], alpha[2];
    uint32_t done[2], iters[2];
    uint32_t sweeps, halt, cap, pad;
} sp_cg_state;

/* Krylov state of the pressure solve (`SolveBuf`, common.glsl:138-149), with the scalar
 * reductions in FP64. */
typedef struct sp_solve {
    double rz[2], alpha[2], beta[2], bb[2], rr[2];
    uint32_t done[2], iters[2];
    uint32_t it_max, unconv, solves, arrive;
    float res_max;
    uint32_t iters_last;
    uint32_t p2_it_max, p2_unconv;   /* the second pressure solve (after the viscosity) */
    /* the viscosity solve: one right-hand side (u and v faces together) */
    double v_rz, v_alpha, v_beta, v_r0, v_rr, v_pq;
    uint32_t v_done, v_iters, v_it_max, v_unconv;
    float v_res_max;
    uint32_t v_iters_last, p2_iters_last;
    uint32_t iters_sum;  /* iterations of every first pressure solve so far, wrapping at 2^32 */
    /* This tank's recorded iterations of the pressure, second pressure and viscosity solves
     * (written by the host before each batch; 0 = no cap), the iterations the running solves
     * have done, and whether they stopped at the cap. A tank whose launches are shared with
     * tanks that record more iterations stops where it would alone. */
    uint32_t cap_p, cap_p2, cap_v, cap_now;
    uint32_t sweeps, halt, v_sweeps, v_halt;
    sp_cg_state st[2];   /* the pressure CG's state entering iterations of each parity */
} sp_solve;

#define SP_MAX_LEVELS 16
/* The multigrid hierarchy (`Levels`, src/sim/mod.rs:715-784). */
typedef struct sp_levels_info {
    int32_t n;                   /* levels */
    int32_t tail;                /* first level of the single-dispatch tail */
    int32_t total;               /* cells of all levels */
    int32_t tiles;               /* tile flags of levels 0..tail */
    int32_t dims[SP_MAX_LEVELS][2];
    int32_t offs[SP_MAX_LEVELS];
    int32_t tile_offs[SP_MAX_LEVELS];
} sp_levels_info;

/* Constants of the solver (common.glsl:455-466, mod.rs:100-119). */
#define SP_MASS_SCALE 1048576.0f
#define SP_MOM_SCALE 4194304.0f
#define SP_MIN_FACE_MASS 64
#define SP_THETA_MIN 0.1f
#define SP_REST_PHI 1.0f
#define SP_LIQUID_THR 0.5f
#define SP_FACE_EMPTY 1e30f
#define SP_FACE_CLOSED (-1e30f)
#define SP_MG_OMEGA 1.8f
#define SP_COARSEST_CELLS 4
#define SP_TAIL_CELLS 1500
#define SP_TAIL_SHARED 3328
#define SP_TAIL_SHARED_COARSE 1024
#define SP_TAIL_HALF_SWEEPS 9
#define SP_TAIL_WARP_CELLS 128
#define SP_SORT_INTERVAL 128
#define SP_SORT_MIN_PARTICLES 16384
/* Capabilities (common.glsl:341-352) */
#define SP_CAP_FAMILIES 1u
#define SP_CAP_THERMAL 2u
#define SP_CAP_RIGID 4u
#define SP_CAP_RHEO 8u
#define SP_CAP_ELASTIC 32u
#define SP_CAP_SETTLE 64u
#define SP_CAP_MAGNET 16u
/* Not a reference capability: set where a liquid in the tank evaporates (volatility > 0), so
 * that the particle kernel with the evaporation code runs only there. */
#define SP_CAP_VOLATILE 0x10000u
/* Not a reference capability either: a thermal tank with a liquid that heats itself. */
#define SP_CAP_SELF_HEAT 0x20000u
/* Nor this: the tank has currents or drains (the particle kernel with their code runs). */
#define SP_CAP_CURRENTS 0x40000u
#define SP_CAP_REACT 256u
#define SP_CAP_STEAM 128u  /* a thermal tank holding the steam row (mod.rs:2752-2756) */
/* Not reference capabilities: a thermal tank with a liquid that can boil (the particle kernel
 * with the boiling code runs), and a tank that may hold puffs of vapour (steam escaped at the
 * surface, or a liquid that fumes as it evaporates or boils away). */
#define SP_CAP_BOILS 0x80000u
#define SP_CAP_PUFFS 0x100000u
#define SP_REACT_MIN 0.02f  /* common.glsl:259 */
#define SP_MU0 1.25663706e-6f  /* magnet.comp MU0 */
/* Viscoelastic stress (common.glsl:376-382): fixed-point scale of stress / el_g in the
 * accumulators */
#define SP_STRESS_FIX 65536.0f
/* Heat (common.glsl:431-436) */
#define SP_NO_TEMP (-1e9f)
#define SP_LEDGER_SCALE 65536.0f
#define SP_T_SCALE 4096.0f
/* 8 energy entries as 64-bit fixed point, then per table row the particles boiled or
 * evaporated away, the residues left, the steam made (mod.rs:82 LEDGER_WORDS) */
#define SP_LEDGER_WORDS (16 + 3 * SP_MAX_MATERIALS)
/* Rigid bodies (rigid.comp, common.glsl:263-275; mod.rs:77-78) */
#define SP_RIGID_BODIES 4096
#define SP_BODY_WORDS 32
#define SP_RIGID_WORDS (16 + SP_BODY_WORDS * SP_RIGID_BODIES)
#define SP_RIGID_FIX 65536.0f
#define SP_RIGID_SHARE 0.5f

/* Internal fields (sp_debug_download / sp_debug_upload). Sizes: cells N = W*H. */
enum {
    SP_FIELD_PARTICLES = 0,  /* sp_particle[n_slots] */
    SP_FIELD_ACC = 1,        /* int32[N*stride] */
    SP_FIELD_VEL = 2,        /* float4[N] */
    SP_FIELD_CELL = 3,       /* float4[N] */
    SP_FIELD_COMP = 4,       /* float4[N] */
    SP_FIELD_COEF = 5,       /* float4[N] */
    SP_FIELD_SURF = 6,       /* float4[N] */
    SP_FIELD_PJUMP = 7,      /* float2[N] */
    SP_FIELD_MGA = 8,        /* float4[total] */
    SP_FIELD_MGD = 9,        /* float[total] */
    SP_FIELD_PRHS = 10,      /* float4[N] */
    SP_FIELD_PRES = 11,      /* float2[N] */
    SP_FIELD_STEP = 12,      /* sp_stepbuf */
    SP_FIELD_SOLVE = 13,     /* sp_solve */
    SP_FIELD_SPROP = 14,     /* float4[N] */
    SP_FIELD_AUX = 15,       /* float4[N] */
    SP_FIELD_GRHO = 16,      /* float4[N] */
    SP_FIELD_VEL0 = 17,      /* float2[N] */
    SP_FIELD_MGZ = 18,       /* float2[total] */
    SP_FIELD_MGR = 19,       /* float2[total] */
    SP_FIELD_PTILE = 20,     /* uint32[tiles] */
    SP_FIELD_REGION_AIR = 21,/* uint32 */
    SP_FIELD_SURF2 = 22,     /* float4[N] */
    SP_FIELD_VW = 23,        /* float4[N] */
    SP_FIELD_VS = 24,        /* float4[N] */
    SP_FIELD_GBASE = 25,     /* float4[N] */
    SP_FIELD_VR = 26,        /* float2[N]  viscosity residual */
    SP_FIELD_VP = 27,        /* float2[N]  search direction */
    SP_FIELD_VP2 = 28,       /* float2[N]  the other direction buffer (Jacobi) */
    SP_FIELD_VQ = 29,        /* float2[N]  A p */
    SP_FIELD_VD = 30,        /* float2[N]  the change x - u* of the last solve */
    SP_FIELD_VR2 = 31,       /* float2[N]  second residual buffer (multigrid) */
    SP_FIELD_VQ2 = 32,       /* float2[N]  second A p buffer (multigrid) */
    SP_FIELD_VAZ = 33,       /* float2[N]  A z */
    SP_FIELD_VTILES = 34,    /* uint32[tiles of level 0] active viscosity tiles */
    SP_FIELD_VTILE = 35,     /* uint32[tiles] per-level tile flags of the viscosity multigrid */
    SP_FIELD_VMA = 36,       /* float4[total] */
    SP_FIELD_VMB = 37,       /* float4[total] */
    SP_FIELD_VMD = 38,       /* float2[total] */
    SP_FIELD_VMR = 39,       /* float2[total] */
    SP_FIELD_VMX = 40,       /* float2[total] */
    SP_FIELD_VMZ = 41,       /* float2[total] */
    SP_FIELD_P2KEEP = 42,    /* float[N] */
    SP_FIELD_THERM = 43,     /* float4[N]  temperature, heat capacity, conductivity, frozen share */
    SP_FIELD_DHEAT = 44,     /* float4[N]  change of enthalpy temperature, exchange share, mean, 0 */
    SP_FIELD_HCELL = 45,     /* int2[N]    heat capacity and heat sums of the particles in the cell */
    SP_FIELD_SOLID_T = 46,   /* float[N]   temperature a solid is held at (-1e9: insulating) */
    SP_FIELD_LEDGER = 47,    /* uint32[16] energy entries 0-7, 64-bit fixed point */
    SP_FIELD_RL = 48,        /* int2[N]    rigid body: root cell (-1: not rigid), at a root its slot */
    SP_FIELD_RBODY = 49,     /* uint32[16 + 32*4096] bodies claimed, then per slot 64-bit sums */
    SP_FIELD_RH = 50,        /* float4[N]  non-Newtonian share, its row, its concentration; 1 where stiff */
    SP_FIELD_ESTR = 51,      /* float4[N]  viscoelastic stress per cell: xx, yy, xy, 0 (cells^2/s^2) */
    SP_FIELD_PSTRESS = 52,   /* float4[n_slots] viscoelastic stress per particle slot: xx, yy, xy, 0 */
    SP_FIELD_FAMOCC = 53,    /* float[N*n_fam] occupancy of each concentration family per cell */
    SP_FIELD_DSETTLE = 54,   /* float2[N*n_fam] settling's gain and loss factors per cell and family */
    SP_FIELD_MAGC = 55,      /* float2[N] susceptibility of the cell's liquid, magnetic energy density (grid units) */
    SP_FIELD_RMASK = 56,     /* uint32[N] bit j: the tank's j-th row makes up more than REACT_MIN of the cell */
    SP_FIELD_MGX = 57,       /* float2[total]: the V-cycle's pre-smoothed iterate per level */
    SP_FIELD_COUNT
};

/* Stages of one substep (record_substep, src/sim/mod.rs:2615-2708), for sp_debug_stage. */
enum {
    SP_STAGE_STEP = 0,       /* step.comp */
    SP_STAGE_CLEAR_ACC = 1,  /* fill acc */
    SP_STAGE_PARTICLES = 2,  /* particles.comp; arg = gather flag */
    SP_STAGE_GRID_PREP = 3,
    SP_STAGE_COEF = 4,
    SP_STAGE_SURF = 5,       /* surf_smooth + surf_curv + surf_apply */
    SP_STAGE_VISC = 6,       /* viscosity before the first solve: explicit, or the implicit split (+ solve) */
    SP_STAGE_RHS = 7,        /* fill ptile + rhs.comp */
    SP_STAGE_PRESSURE = 8,   /* record_pressure (first solve) */
    SP_STAGE_PROJECT = 9,
    SP_STAGE_SUBSTEP = 10,   /* the whole substep; arg = gather flag */
    SP_STAGE_MG_BUILD = 11,  /* the Galerkin hierarchy only */
    SP_STAGE_VCYCLE = 12,    /* one V-cycle: mgr level 0 -> mgz level 0 */
    SP_STAGE_VISC_SPLIT = 13,   /* record_viscosity_split: weights, tiles, explicit step, matrices */
    SP_STAGE_VISC_SOLVE = 14,   /* record_viscosity_solve; arg = restrain */
    SP_STAGE_RHS_VISC = 15,     /* rhs_visc.comp */
    SP_STAGE_PRESSURE2 = 16,    /* the second pressure solve */
    SP_STAGE_PROJECT_VISC = 17, /* project_visc.comp */
    SP_STAGE_VISC_VCYCLE = 18,  /* one viscosity V-cycle: vr -> vmz level 0 (and vaz) */
    SP_STAGE_HEAT = 19,         /* heat.comp */
    SP_STAGE_RIGID = 20,        /* rigid.comp; arg = mode 0..4 */
    SP_STAGE_RHEO = 21,         /* rheo.comp */
    SP_STAGE_ELASTIC = 22,      /* elastic.comp */
    SP_STAGE_SETTLE = 23,       /* settle.comp */
    SP_STAGE_MAGNET = 24,       /* magnet.comp; arg = mode 0 (energy) or 1 (force) */
    SP_STAGE_COUNT
};

#ifdef __cplusplus
}
#endif

/* Byte layouts shared with the reference's std430 buffers and with the device code. */
#ifdef __cplusplus
#define SP_STATIC_ASSERT static_assert
#else
#define SP_STATIC_ASSERT _Static_assert
#endif
SP_STATIC_ASSERT(sizeof(sp_particle) == 48, "Particle is 48 bytes (common.glsl:11-19)");
SP_STATIC_ASSERT(sizeof(sp_stepbuf) == 48, "StepBuf is 12 words (common.glsl:122-135)");
SP_STATIC_ASSERT(sizeof(sp_gmat) == 208, "one grid-unit material row is 52 words");
SP_STATIC_ASSERT(sizeof(sp_gpumat) == 208, "GpuMat is 208 bytes (material.rs:549)");
SP_STATIC_ASSERT(sizeof(sp_solve) % 8 == 0, "FP64 Krylov scalars stay 8-byte aligned");
SP_STATIC_ASSERT(sizeof(sp_cg_state) == 112, "CG state: 8 doubles, 12 words, no padding");
SP_STATIC_ASSERT(sizeof(sp_solve) == 240 + 2 * 112, "solve record layout (mirrored by the Python binding)");

#endif
