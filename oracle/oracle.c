/*
 * oracle.c — CPU restatement of the Sprout fluid step (the reference's GLSL compute
 * shaders, run through Vulkano). TEST INFRASTRUCTURE ONLY: it checks the HIP path and
 * times the CPU baseline; nothing in the product calls it.
 *
 * Scope: the substep with heat, rigid bodies, rheology, viscoelasticity, settling and
 * magnetism, evaporation, reactions and heat made inside, boiling, steam, residues and the
 * puffs of vapour riding the air (air.glsl), currents and drains; without the mouse tools,
 * the smoke field puffs feed (a renderer input) or film spatter:
 *   step.comp, particles.comp (gather, drift, family exchange, advect, collide, scatter),
 *   grid_prep.comp, grid_coef.comp, surf_smooth/curv/apply.comp, visc_setup.comp +
 *   visc_apply.comp mode 2 (explicit viscosity), the implicit viscosity solve
 *   (visc_setup/split/diag/strain/apply/iter/update.comp, vmg_build/down/up.comp,
 *   visc_mgu/mgz.comp), rhs.comp, mg_build/down/up/tail.comp, cg_scale/init/apply/update.comp
 *   with the reduce_tail.glsl scalar steps, project.comp, rhs_visc.comp + the second
 *   pressure solve + project_visc.comp, heat.comp, rheo.comp, rigid.comp, elastic.comp,
 *   settle.comp, magnet.comp, sort.comp, count.comp, stats.comp.
 *
 * Every branch above has a check of its own on the CPU (closed forms, invariants, brute
 * force): tests/test_oracle*.py. The last ones pinned: the surface-tension jump by Laplace's
 * law (test_oracle_surface.py), the drift flux between and within phase groups by the
 * terminal velocity of the drag balance and the cancelling volume flux, the family exchange
 * by its relaxation rate and conservation, the crowd jitter by its bounds and uniformity
 * (test_oracle_drift.py), self-heating and the heat of reaction by the energy they add
 * (test_oracle_react.py), the magnetic resting rule by the direction of the pull
 * (test_oracle_magnet.py), currents and drains by their relaxation and removal rates
 * (test_oracle_currents.py), and boiling, steam, residues and puffs
 * (test_oracle_steam.py): the exact latent-heat threshold and the heat booked for what
 * left, the residue share and its solid state at the boiling point, steam forming at its
 * vapour quality, escaping at the surface with probability x and condensing into the
 * liquid it boiled from with its heat content kept, the bubble's drift at the drag
 * balance's terminal velocity, the exchange rate of a cell holding steam, a puff's
 * exponential take-up of the air's rise and its lifetime, and divergence-free eddies.
 *
 * Every float operation follows the shader's order (compile with -ffp-contract=off), so
 * that the integer scatter sums and cell indices can be compared bit for bit. The CG
 * scalar reductions are FP64 (the build's precision plan); the reference sums FP32 in a
 * fixed tree (reduce_tail.glsl:28-41).
 *
 * Built by oracle/Makefile into oracle/liboracle.so (gcc, no GPU).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NO_REDUCE 0xffffffffu

typedef struct { float x, y; } v2;
typedef struct { float x, y, z, w; } v4;

struct or_sim {
    int W, H, N;
    uint32_t cap, n_slots;
    sp_gparams P;
    sp_gmat mats[SP_MAX_MATERIALS];
    sp_particle* part;
    int32_t* acc;
    v4 *vel, *cell, *comp, *coef, *surf, *surf2, *sprop, *aux, *grho, *gbase, *solid, *prhs;
    v4 *vw, *vs;
    v2 *vel0, *pjump, *pres;
    float *sdf, *mixb, *p2keep;
    uint32_t* region;
    uint32_t region_air;
    sp_levels_info lv;
    v4* mga;
    float* mgd;
    v2 *mgx, *mgz, *mgr, *cgp, *cgq;
    uint32_t* ptile;
    /* implicit viscosity */
    v2 *vr, *vp, *vp2, *vq, *vd, *vr2, *vq2, *vaz;
    uint32_t* vtiles;  /* active viscosity tiles (16 x 16) of the grid */
    uint32_t* vtile;   /* per-level tile flags of the viscosity multigrid */
    v4 *vma, *vmb;
    v2 *vmd, *vmr, *vmx, *vmz;
    sp_stepbuf S;
    sp_solve K;
    /* heat */
    v4 *therm, *dheat;
    int32_t* hcell;   /* 2 per cell */
    float* solid_t;
    uint32_t ledger[SP_LEDGER_WORDS];
    /* rigid bodies */
    int32_t* rl;      /* 2 per cell: root cell (-1: not rigid); at a root, its slot */
    uint32_t* rbody;  /* SP_RIGID_WORDS: bodies claimed, then 64-bit sums per slot */
    v4* rh;           /* non-Newtonian share, its row, its concentration; w = 1 where stiff */
    /* viscoelasticity */
    v4* estr;         /* per cell: stress xx, yy, xy (cells^2/s^2) */
    v4* pstress;      /* per particle slot: its stress xx, yy, xy */
    /* settling, per cell and family (cell * n_fam + f) */
    float* famocc;    /* the family's occupancy of the cell */
    v2* dsettle;      /* gain and loss factors of the family's concentration this substep */
    /* magnetism: per cell the susceptibility of its liquid and the magnetic energy density */
    v2* magc;
    /* reactions: per cell, bit j where the tank's j-th row makes up more than REACT_MIN */
    uint32_t* rmask;
};

/* ---- small helpers with the shader's float semantics ---------------------------------- */

static float fmaxf_(float a, float b) { return a > b ? a : b; }
static float fminf_(float a, float b) { return a < b ? a : b; }
static float clampf(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }
static int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
/* GLSL mix(x, y, a) = x (1 - a) + y a. */
static float mixf(float x, float y, float a) { return x * (1.0f - a) + y * a; }
static float len2(float x, float y) { return sqrtf(x * x + y * y); }
static float signf(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
static float smoothstepf(float e0, float e1, float x) {
    float t = clampf((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static int face_live(float s) { return fabsf(s) < 1e29f; }
static int face_closed(float s) { return s < -1e29f; }
static float v4c(v4 a, int k) { return k == 0 ? a.x : k == 1 ? a.y : k == 2 ? a.z : a.w; }

/* hash (air.glsl:4-11) */
static uint32_t hash_u(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

/* bspline (common.glsl:638-643) */
static void bspline(float f, float w[3]) {
    w[0] = 0.5f * (1.5f - f) * (1.5f - f);
    w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
    w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}

static int cidx(const or_sim* s, int x, int y) { return y * s->W + x; }

/* row_density (common.glsl:475-479) */
static float row_density(const or_sim* s, int row, float sc) {
    int d = s->mats[row].diluent;
    float rho = s->mats[row].rho;
    return d >= 0 ? mixf(s->mats[d].rho, rho, sc) : rho;
}

/* terminal (common.glsl:370-374) */
/* ---- heat helpers (common.glsl:408-453, 485-530, 579-606; particles.comp:67-80) ------- */

static int thermal(const or_sim* s) { return (s->P.caps & SP_CAP_THERMAL) != 0u; }

/* temp_law (common.glsl:410-417): ln of the viscosity factor of row r at T (deg C). */
static float temp_law(const or_sim* s, int r, float T) {
    const sp_gmat* m = &s->mats[r];
    if (m->temp_model == 0) return 0.0f;
    float Tk = fmaxf_(T + 273.15f, 1.0f);
    float x = m->temp_model == 1 ? m->temp[0] * (1.0f / Tk - 1.0f / m->temp[2])
                                 : m->temp[0] / fmaxf_(Tk - m->temp[1], 5.0f) - m->temp[0] / fmaxf_(m->temp[2] - m->temp[1], 5.0f);
    return clampf(x, -30.0f, 30.0f);
}

/* expansion (common.glsl:485-494): density factor of row r at concentration sc and T. */
static float expansion(const or_sim* s, int row, float sc, float T) {
    const sp_gmat* m = &s->mats[row];
    float b = m->expansion, tr = m->temp[2] - 273.15f, tb = m->th2[2];
    if (m->diluent >= 0) {
        const sp_gmat* d = &s->mats[m->diluent];
        b = mixf(d->expansion, b, sc);
        tr = mixf(d->temp[2] - 273.15f, tr, sc);
        tb = mixf(d->th2[2], tb, sc);
    }
    return clampf(1.0f - b * (fminf_(T, tb) - tr), 0.7f, 1.3f);
}

/* thermal_of (common.glsl:500-509): solidus, liquidus, latent heat (K), heat capacity. */
static v4 thermal_of(const or_sim* s, int row, float sc) {
    const sp_gmat* m = &s->mats[row];
    float ld = m->th2[0];
    if (m->diluent < 0) return (v4){m->th[2], m->th[3], ld, m->th[0]};
    const sp_gmat* d = &s->mats[m->diluent];
    float shift = m->freeze_shift * sc;
    float C = mixf(d->th[0], m->th[0], sc);
    return (v4){d->th[2] + shift, d->th[3] + shift, mixf(d->th[0] * d->th2[0], m->th[0] * ld, sc) / C, C};
}

/* phase_of (common.glsl:515-521): temperature and solid share of enthalpy temperature te. */
static v2 phase_of(float te, v4 m) {
    if (te >= m.y) return (v2){te, 0.0f};
    float lo = m.x - m.z;
    if (te <= lo) return (v2){te + m.z, 1.0f};
    float fr = (te - lo) / (m.y - lo);
    return (v2){m.x + fr * (m.y - m.x), 1.0f - fr};
}

/* heat_quanta (particles.comp:70-73) */
static void heat_quanta(float C, float T, int32_t q[2]) {
    int ci = (int)floorf(C * SP_MASS_SCALE + 0.5f);
    q[0] = ci;
    q[1] = (int)floorf((float)ci * (T + 273.15f) / SP_T_SCALE + 0.5f);
}

/* keep_heat (particles.comp:78-80) */
static float keep_heat(float t, float c0, float c1) { return c0 == c1 ? t : (t + 273.15f) * c0 / c1 - 273.15f; }

/* The boiling point of row at concentration sc (particles.comp:368, 957). */
static float boil_of(const or_sim* s, int row, float sc) {
    const sp_gmat* m = &s->mats[row];
    return m->diluent >= 0 ? mixf(s->mats[m->diluent].th2[2], m->th2[2], sc) : m->th2[2];
}

/* group_heat_chan (common.glsl:582-584) */
static int group_heat_chan(const or_sim* s, int g) { return (int)((s->P.group_heat >> (8u * (uint32_t)g)) & 255u) - 1; }

/* ledger_add (common.glsl:446-451): a 64-bit fixed-point sum in two words. */
static void ledger_add(or_sim* s, int e, float v) {
    uint32_t x = (uint32_t)(v * SP_LEDGER_SCALE + 0.5f);
    if (x == 0u) return;
    uint32_t old = s->ledger[2 * e];
    s->ledger[2 * e] = old + x;
    if (old + x < old) s->ledger[2 * e + 1] += 1u;
}

static float terminal(const or_sim* s, float drho, float rho_c, float nu_c, float d) {
    float qa = 0.44f * rho_c, qb = 24.0f * nu_c * rho_c / d, qc = 4.0f / 3.0f * fabsf(drho) * s->P.gravity * d;
    return qc > 0.0f ? 2.0f * qc / (qb + sqrtf(qb * qb + 4.0f * qa * qc)) : 0.0f;
}

/* group_sorts (common.glsl:590-602) */
static int group_sorts(const or_sim* s, int g) {
    if (g <= 0) return 0;
    int n = 0;
    for (int j = 0; j < s->P.n_present; j++) {
        int r = s->P.present[j];
        const sp_gmat* m = &s->mats[r];
        if (m->diluent >= 0 && s->mats[m->diluent].group == g) return 0;
        if (m->group != g) continue;
        if (m->diluent >= 0) return 0;
        n++;
    }
    return n >= 2;
}

/* sample_sdf (air.glsl:13-22) */
static float sample_sdf(const or_sim* s, float px, float py, float* gx, float* gy) {
    float x = clampf(px, 0.0f, (float)s->W - 0.001f), y = clampf(py, 0.0f, (float)s->H - 0.001f);
    int ix = (int)x, iy = (int)y;
    float fx = x - (float)ix, fy = y - (float)iy;
    int w1 = s->W + 1;
    int k = iy * w1 + ix;
    float d00 = s->sdf[k], d10 = s->sdf[k + 1], d01 = s->sdf[k + w1], d11 = s->sdf[k + w1 + 1];
    *gx = mixf(d10 - d00, d11 - d01, fy);
    *gy = mixf(d01 - d00, d11 - d10, fx);
    return mixf(mixf(d00, d10, fx), mixf(d01, d11, fx), fy);
}

/* ---- rigid bodies (common.glsl:263-340) ----------------------------------------------- */

/* row_lnu (common.glsl:631-636) */
static float row_lnu(const or_sim* s, int row, float sc, float T, int th) {
    const sp_gmat* m = &s->mats[row];
    float l = m->ln_nu + (th ? temp_law(s, row, T) : 0.0f);
    int d = m->diluent;
    if (d >= 0) {
        float bs = m->visc_exp == 1.0f ? sc : powf(sc, m->visc_exp);
        l = mixf(s->mats[d].ln_nu + (th ? temp_law(s, d, T) : 0.0f), l, bs);
    }
    return l;
}

/* get64 (common.glsl:277-279) */
static float get64(const or_sim* s, uint32_t w) {
    return ((float)(int32_t)s->rbody[w + 1] * 4294967296.0f + (float)s->rbody[w]) / SP_RIGID_FIX;
}

/* body_contact (common.glsl:290-324) */
static void body_contact(const or_sim* s, uint32_t o, v2 cm, v2* U, float* w) {
    const float mu = s->P.rigid_friction;
    const float nc = get64(s, o + 14u);
    if (mu < 0.0f || nc < 0.5f) return;
    const float nsx = get64(s, o + 16u), nsy = get64(s, o + 18u);
    const float nsl = len2(nsx, nsy);
    if (nsl < 0.25f * nc) return;
    const float nx = nsx / nsl, ny = nsy / nsl;
    const float tx = ny, ty = -nx;
    const float ccx = get64(s, o + 20u) / nc, ccy = get64(s, o + 22u) / nc;
    const float sqx = get64(s, o + 24u) / nc - ccx * ccx;
    const float sqy = get64(s, o + 26u) / nc - ccy * ccy;
    const float sqz = get64(s, o + 28u) / nc - ccx * ccy;
    const float ext = sqrtf(3.0f * fmaxf_(tx * tx * sqx + ty * ty * sqy + 2.0f * tx * ty * sqz, 0.0f)) + 0.5f;
    const float rcx = ccx - cm.x, rcy = ccy - cm.y;
    float ucx = U->x + *w * -rcy, ucy = U->y + *w * rcx;
    const float vn = ucx * nx + ucy * ny;
    const float sp = (cm.x - ccx) * tx + (cm.y - ccy) * ty;
    float vn1 = vn, w1 = *w;
    if (fabsf(sp) <= ext) {
        vn1 = fmaxf_(vn, 0.0f);
        w1 = clampf(*w, -vn1 / ext, vn1 / ext);
    } else {
        const float edge = signf(sp) * ext;
        vn1 = fmaxf_(vn + *w * edge, 0.0f) - *w * edge;
    }
    if (vn1 == vn && w1 == *w) return;
    const float vt = ucx * tx + ucy * ty;
    const float vt1 = signf(vt) * fmaxf_(fabsf(vt) - mu * (vn1 - vn), 0.0f);
    ucx += (vn1 - vn) * nx + (vt1 - vt) * tx;
    ucy += (vn1 - vn) * ny + (vt1 - vt) * ty;
    *w = w1;
    U->x = ucx - *w * -rcy;
    U->y = ucy - *w * rcx;
}

/* body_motion (common.glsl:326-340): (u, v, angular velocity) of slot b (root r) at (px, py) */
static int body_motion(const or_sim* s, uint32_t b, int r, float px, float py, float u[3]) {
    const uint32_t o = 16u + (uint32_t)SP_BODY_WORDS * b;
    const float M = get64(s, o);
    if (M <= 1e-3f) return 0;
    const v2 cm = {get64(s, o + 2u) / M, get64(s, o + 4u) / M};
    v2 U = {get64(s, o + 6u) / M, get64(s, o + 8u) / M};
    const float L = get64(s, o + 10u) - M * (cm.x * U.y - cm.y * U.x);
    const float I = get64(s, o + 12u) - M * (cm.x * cm.x + cm.y * cm.y);
    float w = I > 1e-3f ? L / I : 0.0f;
    body_contact(s, o, cm, &U, &w);
    const float dx = px - ((float)(r % s->W) + 0.5f) - cm.x;
    const float dy = py - ((float)(r / s->W) + 0.5f) - cm.y;
    u[0] = U.x + w * -dy;
    u[1] = U.y + w * dx;
    u[2] = w;
    return 1;
}

/* sample_vel (air.glsl:41-78) */
static v2 sample_vel(const or_sim* s, float px, float py, v2 fb) {
    v2 o = fb;
    {
        float gx = px, gy = py - 0.5f;
        int bx = (int)floorf(gx), by = (int)floorf(gy);
        float fx = gx - (float)bx, fy = gy - (float)by;
        float sw = 0.0f, su = 0.0f;
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                v4 q = s->vel[cidx(s, bx + a, by + b)];
                if (face_live(q.z)) {
                    float wt = (a == 0 ? 1.0f - fx : fx) * (b == 0 ? 1.0f - fy : fy);
                    sw += wt;
                    su += wt * q.x;
                }
            }
        if (sw > 1e-3f) o.x = su / sw;
    }
    {
        float gx = px - 0.5f, gy = py;
        int bx = (int)floorf(gx), by = (int)floorf(gy);
        float fx = gx - (float)bx, fy = gy - (float)by;
        float sw = 0.0f, su = 0.0f;
        for (int b = 0; b < 2; b++)
            for (int a = 0; a < 2; a++) {
                v4 q = s->vel[cidx(s, bx + a, by + b)];
                if (face_live(q.w)) {
                    float wt = (a == 0 ? 1.0f - fx : fx) * (b == 0 ? 1.0f - fy : fy);
                    sw += wt;
                    su += wt * q.y;
                }
            }
        if (sw > 1e-3f) o.y = su / sw;
    }
    return o;
}

/* collide (particles.comp:51-65) */
static int collide(const or_sim* s, float* x, float* y, float* vx, float* vy, int beyond) {
    float nx, ny;
    float d = sample_sdf(s, *x, *y, &nx, &ny);
    if (d < s->P.wall_margin) {
        float l = len2(nx, ny);
        if (l > 1e-5f) {
            nx /= l;
            ny /= l;
            float k = s->P.wall_margin - d + (beyond ? fmaxf_(-d, 0.0f) : 0.0f);
            *x += nx * k;
            *y += ny * k;
            float vn = *vx * nx + *vy * ny;
            if (vn < 0.0f) {
                *vx -= vn * nx;
                *vy -= vn * ny;
            }
            return 1;
        }
    }
    return 0;
}

/* ---- the mouse tool (air.glsl:24-37) ----------------------------------------------------- */

/* tool_target: the velocity the tool drives a particle towards, (dx, dy) to the cursor. */
static v2 tool_target(const sp_gparams* P, float dx, float dy) {
    float l = len2(dx, dy);
    if (P->tool_mode == 3) {
        float m = fmaxf_(l, 0.15f * P->tool_radius);
        return (v2){-dx / m * P->tool_vel[0], -dy / m * P->tool_vel[0]};
    }
    if (P->tool_mode == 4) return (v2){P->tool_vel[0], P->tool_vel[1]};
    if (P->tool_mode == 7) {
        float m = fmaxf_(l, 0.4f * P->tool_radius);
        return (v2){dy / m * P->tool_vel[0] + dx * 2.0f, -dx / m * P->tool_vel[0] + dy * 2.0f};
    }
    if (P->tool_mode == 8) return (v2){0.0f, 0.0f};
    return (v2){P->tool_vel[0] + dx * 4.0f, P->tool_vel[1] + dy * 4.0f};
}

/* ---- the air above the liquid (air.glsl:91-175): a procedural flow escaped vapour rides --- */

#define PUFF_FOLLOW 0.08f /* air.glsl:102 */

/* vnoise3 (air.glsl:108-128): value noise in [-0.5, 0.5) and its gradient. */
static v4 vnoise3(float x, float y, float z) {
    float ix = floorf(x), iy = floorf(y), iz = floorf(z);
    float fx = x - ix, fy = y - iy, fz = z - iz;
    float ux = fx * fx * fx * (fx * (fx * 6.0f - 15.0f) + 10.0f);
    float uy = fy * fy * fy * (fy * (fy * 6.0f - 15.0f) + 10.0f);
    float uz = fz * fz * fz * (fz * (fz * 6.0f - 15.0f) + 10.0f);
    float dux = 30.0f * fx * fx * (fx * (fx - 2.0f) + 1.0f);
    float duy = 30.0f * fy * fy * (fy * (fy - 2.0f) + 1.0f);
    float duz = 30.0f * fz * fz * (fz * (fz - 2.0f) + 1.0f);
    uint32_t cx = (uint32_t)((int)ix + 4096), cy = (uint32_t)((int)iy + 4096), cz = (uint32_t)((int)iz + 4096);
    float r[8];
    for (int k = 0; k < 8; k++) {
        uint32_t qx = cx + (uint32_t)(k & 1), qy = cy + (uint32_t)((k >> 1) & 1), qz = cz + (uint32_t)((k >> 2) & 1);
        r[k] = (float)(hash_u(qx * 73856093u ^ qy * 19349663u ^ qz * 83492791u) >> 8) / 16777216.0f;
    }
    float k0 = r[0], k1 = r[1] - r[0], k2 = r[2] - r[0], k3 = r[4] - r[0];
    float k4 = r[0] - r[1] - r[2] + r[3], k5 = r[0] - r[2] - r[4] + r[6], k6 = r[0] - r[1] - r[4] + r[5];
    float k7 = -r[0] + r[1] + r[2] - r[3] + r[4] - r[5] - r[6] + r[7];
    float n = k0 + k1 * ux + k2 * uy + k3 * uz + k4 * ux * uy + k5 * uy * uz + k6 * uz * ux + k7 * ux * uy * uz;
    return (v4){n - 0.5f, dux * (k1 + k4 * uy + k6 * uz + k7 * uy * uz), duy * (k2 + k4 * ux + k5 * uz + k7 * ux * uz),
                duz * (k3 + k5 * uy + k6 * ux + k7 * ux * uy)};
}

/* air_psi (air.glsl:130-141): the eddies' stream function and its gradient. */
static v4 air_psi(const sp_gparams* P, float x, float y, float t) {
    float L = fmaxf_(P->steam2[3], 1.0f);
    float turn = 0.7f * t * P->steam2[2] / L;
    float ax = x / L, ay = (y - 0.5f * P->steam[1] * t) / L;
    v4 n1 = vnoise3(ax, ay, turn);
    v4 n2 = vnoise3(ax * 2.17f + 17.3f, ay * 2.17f + 17.3f, 2.2f * turn);
    float amp = 2.2f * P->steam2[2] * L;
    return (v4){amp * (n1.x + 0.45f * n2.x), amp * ((n1.y + 0.45f * 2.17f * n2.y) / L), amp * ((n1.z + 0.45f * 2.17f * n2.z) / L), 0.0f};
}

/* puff_lean (air.glsl:143-146) */
static float puff_lean(uint32_t pid) { return ((float)(hash_u(pid ^ 0x5bd1e995u) >> 8) / 16777216.0f - 0.5f) * 0.3f; }

/* sample_occ (air.glsl:81-88): liquid occupancy, bilinear over the cell centres. */
static float sample_occ(const or_sim* s, float x, float y) {
    float gx = clampf(x - 0.5f, 0.0f, (float)s->W - 1.001f), gy = clampf(y - 0.5f, 0.0f, (float)s->H - 1.001f);
    int x0 = (int)gx, y0 = (int)gy;
    float fx = gx - (float)x0, fy = gy - (float)y0;
    int k = y0 * s->W + x0;
    return mixf(mixf(s->cell[k].x, s->cell[k + 1].x, fx), mixf(s->cell[k + s->W].x, s->cell[k + s->W + 1].x, fx), fy);
}

/* air_velocity (air.glsl:148-175). */
static v2 air_velocity(const or_sim* s, float x, float y, float lean, float* occ) {
    const sp_gparams* P = &s->P;
    float t = s->S.advanced;
    float nx, ny;
    float d = sample_sdf(s, x, y, &nx, &ny);
    float nl = len2(nx, ny);
    if (nl > 1e-5f) {
        nx = nx / nl;
        ny = ny / nl;
    } else {
        nx = 0.0f;
        ny = 1.0f;
    }
    float Lb = fmaxf_(P->steam2[3], 1.0f);
    float r = clampf(d / Lb, 0.0f, 1.0f);
    float ramp = r * (1.875f - r * r * (1.25f - 0.375f * r * r));
    float dramp = d < Lb && d > 0.0f ? (1.875f - r * r * (3.75f - 1.875f * r * r)) / Lb : 0.0f;
    float vx = 0.0f, vy = 0.0f;
    if (P->steam2[2] > 0.0f) {
        v4 psi = air_psi(P, x, y, t);
        vx = ramp * psi.z + psi.x * dramp * ny;
        vy = ramp * -psi.y + psi.x * dramp * -nx;
    }
    float rx = lean * P->steam[1], ry = 1.0f * P->steam[1];
    float into = rx * nx + ry * ny;
    if (into < 0.0f) {
        float tx = ny, ty = -nx;
        float side = (rx * tx + ry * ty) + lean * 1e-3f >= 0.0f ? 1.0f : -1.0f;
        rx -= (nx + side * tx * 0.7f) * into * (1.0f - ramp);
        ry -= (ny + side * ty * 0.7f) * into * (1.0f - ramp);
    }
    vx += rx;
    vy += ry;
    *occ = sample_occ(s, x, y);
    v2 vl = sample_vel(s, x, y, (v2){vx, vy - P->steam[1]});
    float k = smoothstepf(0.0f, 0.4f, *occ);
    vx = mixf(vx, vl.x, k);
    vy = mixf(vy, vl.y + P->steam[1], k);
    int tm = P->tool_mode;
    if (tm == 1 || tm == 3 || tm == 4 || tm == 7) {
        float tdx = P->tool_pos[0] - x, tdy = P->tool_pos[1] - y;
        float r2 = (tdx * tdx + tdy * tdy) / (P->tool_radius * P->tool_radius);
        if (r2 < 1.0f) {
            v2 tt = tool_target(P, tdx, tdy);
            vx = mixf(vx, tt.x, 1.0f - r2);
            vy = mixf(vy, tt.y, 1.0f - r2);
        }
    }
    return (v2){vx, vy};
}

/* A puff of vapour (particles.comp:103-134): a tracer in the air, gone after steam.z seconds,
 * at the grid's edge, or swallowed by the liquid. (The smoke field it feeds is the
 * renderer's and is not restated.) */
static void advance_puff(or_sim* s, uint32_t id, sp_particle p) {
    const sp_gparams* P = &s->P;
    float dt = s->S.dt;
    p.s += dt;
    if (p.s >= P->steam[2] || p.pos[0] < 2.01f || p.pos[1] < 2.01f || p.pos[0] > (float)s->W - 2.01f || p.pos[1] > (float)s->H - 2.01f) {
        s->part[id].kind = SP_KIND_DEAD;
        return;
    }
    float occ;
    v2 target = air_velocity(s, p.pos[0], p.pos[1], puff_lean(p.id), &occ);
    if (occ > 0.6f && p.s > 0.1f) {
        s->part[id].kind = SP_KIND_DEAD;
        return;
    }
    if (P->tool_mode == 2 && len2(P->tool_pos[0] - p.pos[0], P->tool_pos[1] - p.pos[1]) < P->tool_radius) {
        s->part[id].kind = SP_KIND_DEAD;
        return;
    }
    float k = 1.0f - expf(-dt / PUFF_FOLLOW);
    float vx = mixf(p.vel[0], target.x, k), vy = mixf(p.vel[1], target.y, k);
    float x = p.pos[0] + vx * dt, y = p.pos[1] + vy * dt;
    collide(s, &x, &y, &vx, &vy, 0);
    p.pos[0] = x;
    p.pos[1] = y;
    p.vel[0] = vx;
    p.vel[1] = vy;
    s->part[id] = p;
}

/* The particle leaves as a puff of vapour `fume` (particles.comp:704-712, 796-800). */
static void become_puff(or_sim* s, uint32_t id, sp_particle p, float vx, float vy, float fume) {
    p.kind = SP_KIND_PUFF;
    p.s = 0.0f;
    p.vel[0] = vx;
    p.vel[1] = vy;
    p.c[0] = fume;
    p.c[1] = 0.0f;
    p.c[2] = 0.0f;
    p.c[3] = 0.0f;
    s->part[id] = p;
}

/* ---- step.comp (step.comp:22-50) ------------------------------------------------------ */

static void stage_step(or_sim* s) {
    sp_stepbuf* S = &s->S;
    const sp_gparams* P = &s->P;
    float u = fmaxf_(fminf_(u2f(S->umax_grid), P->ext[0]), u2f(S->umax_part));
    u = fmaxf_(u, P->ext_speed);
    float dt_cfl = P->cfl / (u + sqrtf(5.0f * fmaxf_(P->gravity, 0.0f)) + 1e-6f);
    float dt;
    if (P->dt_fixed > 0.0f) {
        dt = P->dt_fixed;
    } else {
        uint32_t left = S->batch_steps > S->batch_done ? S->batch_steps - S->batch_done : 1u;
        dt = fminf_(fminf_(dt_cfl, P->dt_max), fmaxf_(S->frame_left, 0.0f) / (float)left);
    }
    dt = fmaxf_(dt, 1e-6f);
    S->dt_prev = S->dt > 0.0f ? S->dt : dt;
    S->dt = dt;
    S->dt_cfl = fminf_(dt_cfl, P->dt_max);
    S->umax_seen = u;
    S->frame_left -= dt;
    float lo = u2f(S->pad0);
    float y = dt - lo;
    float t = S->advanced + y;
    S->pad0 = f2u((t - S->advanced) - y);
    S->advanced = t;
    S->batch_done += 1u;
    S->substeps += 1u;
    S->umax_grid = 0u;
    S->umax_part = 0u;
    s->region_air = 0u;
}

/* ---- particles.comp ------------------------------------------------------------------- */

typedef struct {
    int live;
    float px, py, vsx, vsy, Csx, Csy, Csz, Csw, rho, sc, CT;
    int cw, cs, ct, cg;
    int ce;               /* first stress channel, -1: no stress to scatter */
    float te[3];          /* the stress over el_g, in STRESS_FIX units */
} scatter_in;

static void acc_add(or_sim* s, int x, int y, int ch, int v) {
    s->acc[(size_t)cidx(s, x, y) * s->P.acc_stride + ch] += v;
}

/* deposit (particles.comp:1006-1062) */
static void deposit(or_sim* s, const scatter_in* g) {
    float w[2][3];
    {
        float gx = g->px, gy = g->py - 0.5f;
        int bx = (int)floorf(gx - 0.5f), by = (int)floorf(gy - 0.5f);
        float fx = gx - (float)bx, fy = gy - (float)by;
        bspline(fx, w[0]);
        bspline(fy, w[1]);
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                float m = w[0][a] * w[1][b] * g->rho;
                float dx = (float)a - fx, dy = (float)b - fy;
                acc_add(s, bx + a, by + b, 0, (int)floorf(m * SP_MASS_SCALE + 0.5f));
                acc_add(s, bx + a, by + b, 1, (int)floorf(m * (g->vsx + (g->Csx * dx + g->Csy * dy)) * SP_MOM_SCALE + 0.5f));
            }
    }
    {
        float gx = g->px - 0.5f, gy = g->py;
        int bx = (int)floorf(gx - 0.5f), by = (int)floorf(gy - 0.5f);
        float fx = gx - (float)bx, fy = gy - (float)by;
        bspline(fx, w[0]);
        bspline(fy, w[1]);
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                float m = w[0][a] * w[1][b] * g->rho;
                float dx = (float)a - fx, dy = (float)b - fy;
                acc_add(s, bx + a, by + b, 2, (int)floorf(m * SP_MASS_SCALE + 0.5f));
                acc_add(s, bx + a, by + b, 3, (int)floorf(m * (g->vsy + (g->Csz * dx + g->Csw * dy)) * SP_MOM_SCALE + 0.5f));
            }
    }
    {
        float gx = g->px - 0.5f, gy = g->py - 0.5f;
        int bx = (int)floorf(gx - 0.5f), by = (int)floorf(gy - 0.5f);
        float fx = gx - (float)bx, fy = gy - (float)by;
        bspline(fx, w[0]);
        bspline(fy, w[1]);
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                int wi = (int)floorf(w[0][a] * w[1][b] * SP_MASS_SCALE + 0.5f);
                acc_add(s, bx + a, by + b, 4, wi);
                if (g->cw >= 0) acc_add(s, bx + a, by + b, g->cw, wi);
                if (g->cs >= 0 && g->sc > 0.0f)
                    acc_add(s, bx + a, by + b, g->cs, (int)floorf(w[0][a] * w[1][b] * g->sc * SP_MASS_SCALE + 0.5f));
                /* heat: C (T + 273.15) / T_SCALE into the sum and the group's own (particles.comp:1047-1051) */
                if (g->ct >= 0) {
                    int h = (int)floorf(w[0][a] * w[1][b] * g->CT * SP_MASS_SCALE + 0.5f);
                    acc_add(s, bx + a, by + b, g->ct, h);
                    if (g->cg >= 0) acc_add(s, bx + a, by + b, g->cg, h);
                }
                /* viscoelastic stress, fixed point (particles.comp:1053-1058) */
                if (g->ce >= 0)
                    for (int q = 0; q < 3; q++) acc_add(s, bx + a, by + b, g->ce + q, (int)floorf(w[0][a] * w[1][b] * g->te[q] + 0.5f));
            }
    }
}

/* The upper-convected Maxwell stress of one particle over a substep (particles.comp:249-279),
 * in the shader's matrix form: L the velocity gradient [[du/dx, du/dy], [dv/dx, dv/dy]] (the
 * APIC matrix) scaled to at most el_step_max deformation per substep, T the stress,
 * D = L T + T L^T + G (L + L^T), T' = (T + dt D) / (1 + dt / lambda); the norm capped at
 * el_stress_max G; no stress without a cell's worth of liquid in the 3 x 3 cells around. */
static v4 stress_update(const or_sim* s, float px, float py, const float C[4], v4 tau, float G, float lam, float dt) {
    const sp_gparams* P = &s->P;
    float L[2][2] = {{C[0], C[1]}, {C[2], C[3]}};
    float lnorm = sqrtf(((C[0] * C[0] + C[1] * C[1]) + C[2] * C[2]) + C[3] * C[3]) * dt;
    if (P->el_step_max > 0.0f && lnorm > P->el_step_max) {
        float k = P->el_step_max / lnorm;
        for (int r = 0; r < 2; r++)
            for (int c = 0; c < 2; c++) L[r][c] *= k;
    }
    float T[2][2] = {{tau.x, tau.z}, {tau.z, tau.y}};
    float D[2][2], Tn[2][2];
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) {
            float lt = L[r][0] * T[0][c] + L[r][1] * T[1][c];
            float tl = T[r][0] * L[c][0] + T[r][1] * L[c][1];
            D[r][c] = (lt + tl) + G * (L[r][c] + L[c][r]);
        }
    float k = 1.0f + dt / fmaxf_(lam, 1e-4f);
    for (int r = 0; r < 2; r++)
        for (int c = 0; c < 2; c++) Tn[r][c] = (T[r][c] + dt * D[r][c]) / k;
    v4 t = {Tn[0][0], Tn[1][1], 0.5f * (Tn[0][1] + Tn[1][0]), 0.0f};
    float tn = sqrtf((t.x * t.x + t.y * t.y) + t.z * t.z);
    if (P->el_stress_max > 0.0f && tn > P->el_stress_max * G) {
        float f = P->el_stress_max * G / tn;
        t.x *= f;
        t.y *= f;
        t.z *= f;
        t.w *= f;
    }
    int x1 = clampi((int)px, 1, s->W - 2), y1 = clampi((int)py, 1, s->H - 2);
    float around = 0.0f;
    for (int b = -1; b <= 1; b++)
        for (int a = -1; a <= 1; a++) around += fminf_(s->cell[cidx(s, x1 + a, y1 + b)].x, 1.0f);
    if (around < 1.0f) t = (v4){0.0f, 0.0f, 0.0f, 0.0f};
    return t;
}

/* G2P of one face component (particles.comp:165-222). comp = 0: u faces, 1: v faces. */
static int gather_face(const or_sim* s, float px, float py, int comp, float* u, float* cr, float* Ca, float* Cb) {
    float gx = comp == 0 ? px : px - 0.5f, gy = comp == 0 ? py - 0.5f : py;
    int bx = (int)floorf(gx - 0.5f), by = (int)floorf(gy - 0.5f);
    float fx = gx - (float)bx, fy = gy - (float)by;
    float wx[3], wy[3];
    bspline(fx, wx);
    bspline(fy, wy);
    float sw = 0.0f, su = 0.0f, sc = 0.0f, sdx = 0.0f, sdy = 0.0f, sudx = 0.0f, sudy = 0.0f;
    for (int b = 0; b < 3; b++)
        for (int a = 0; a < 3; a++) {
            v4 q = s->vel[cidx(s, bx + a, by + b)];
            float st = comp == 0 ? q.z : q.w, val = comp == 0 ? q.x : q.y;
            if (face_live(st)) {
                float wt = wx[a] * wy[b];
                float dx = (float)a - fx, dy = (float)b - fy;
                sw += wt;
                su += wt * val;
                sc += wt * st;
                sdx += wt * dx;
                sdy += wt * dy;
                float wu = wt * val;
                sudx += wu * dx;
                sudy += wu * dy;
            }
        }
    if (sw > 1e-4f) {
        float uu = su / sw;
        *u = uu;
        *cr = sc / sw;
        *Ca = 4.0f * (sudx - uu * sdx) / sw;
        *Cb = 4.0f * (sudy - uu * sdy) / sw;
        return 1;
    }
    return 0;
}

/* advance (particles.comp:98-985), the in-scope branches. Returns the scatter inputs. */
static void advance(or_sim* s, uint32_t id, int gather, scatter_in* g, float* umax_part) {
    const sp_gparams* P = &s->P;
    g->live = 0;
    if (id >= s->n_slots) return;
    sp_particle p = s->part[id];
    if (p.kind == SP_KIND_DEAD) return;
    if ((p.kind & 255u) == 0u) {
        advance_puff(s, id, p);
        return;
    }
    float dt = s->S.dt;
    int row = (int)(p.kind & 255u) - 1;
    const sp_gmat* M = &s->mats[row];
    int Mgroup = M->group, Mdil = M->diluent, Mcw = M->chan_w, Mcs = M->chan_s;
    int fam = M->fam;
    const int th_on = thermal(s);
    const int rig = (P->caps & SP_CAP_RIGID) != 0u;
    /* frozen, or too viscous to flow: part of a rigid body (particles.comp:145-151) */
    int frozen = 0;
    if (rig) {
        v2 ph = {P->heat_x, 0.0f};
        if (th_on) ph = phase_of(p.t, thermal_of(s, row, p.s));
        frozen = ph.y > 0.5f || row_lnu(s, row, p.s, ph.x, th_on) >= P->ln_rigid;
    }
    /* viscoelastic stress: a particle that has not run a substep starts unstressed
     * (particles.comp:152-156) */
    const int el_on = (P->caps & SP_CAP_ELASTIC) != 0u;
    v4 tau = {0.0f, 0.0f, 0.0f, 0.0f};
    if (el_on && (p.kind & SP_KIND_FRESH) == 0u && s->mats[row].gr[1] > 0.0f) tau = s->pstress[id];
    if (gather) {
        float vx = p.vel[0], vy = p.vel[1];
        float C[4] = {p.c[0], p.c[1], p.c[2], p.c[3]};
        float corrx = 0.0f, corry = 0.0f, slip = 0.0f, crowd = 0.0f;
        gather_face(s, p.pos[0], p.pos[1], 0, &vx, &corrx, &C[0], &C[1]);
        gather_face(s, p.pos[0], p.pos[1], 1, &vy, &corry, &C[2], &C[3]);
        /* held by the rigid body around it (particles.comp:223-248) */
        int held = 0;
        if (rig) {
            const int px1 = clampi((int)floorf(p.pos[0]), 1, s->W - 2), py1 = clampi((int)floorf(p.pos[1]), 1, s->H - 2);
            const int in_body = row != SP_ROW_STEAM && s->rl[2 * cidx(s, px1, py1)] >= 0;
            const int px0 = clampi((int)floorf(p.pos[0]), 0, s->W - 1), py0 = clampi((int)floorf(p.pos[1]), 0, s->H - 1);
            const int plug = (s->mats[row].rheo_p[0] > 0.0f || s->mats[row].temp[3] > 0.0f) && s->cell[cidx(s, px0, py0)].w <= 0.0f;
            if (frozen || plug || in_body) {
                static const int ox[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1}, oy[9] = {0, 0, 0, -1, 1, -1, -1, 1, 1};
                for (int k = 0; k < 9 && !held; k++) {
                    const int root = s->rl[2 * cidx(s, px1 + ox[k], py1 + oy[k])];
                    if (root < 0 || s->rl[2 * root + 1] < 0) continue;
                    float u[3];
                    if (body_motion(s, (uint32_t)s->rl[2 * root + 1], root, p.pos[0], p.pos[1], u)) {
                        vx = u[0];
                        vy = u[1];
                        C[0] = 0.0f;
                        C[1] = -u[2];
                        C[2] = u[2];
                        C[3] = 0.0f;
                        held = 1;
                    }
                }
            }
        }
        /* the stress over the substep, from the velocity gradient before the clamp
         * (particles.comp:249-279) */
        if (el_on) {
            float G = s->mats[row].gr[1];
            if (G > 0.0f) tau = stress_update(s, p.pos[0], p.pos[1], C, tau, G, s->mats[row].gr[2], dt);
            s->pstress[id] = tau;
        }
        /* Stabilisation (particles.comp:281-285). */
        if (P->apic_clamp > 0.0f) {
            float lim = P->apic_clamp / dt;
            for (int k = 0; k < 4; k++) C[k] = clampf(C[k], -lim, lim);
        }
        if (P->corr[0] > 0.0f) {
            float cl = len2(corrx, corry), cmax = P->corr[0] / dt;
            if (cl > cmax) {
                float k = cmax / cl;
                corrx *= k;
                corry *= k;
            }
        }
        /* Composition gather (particles.comp:288-327). */
        {
            float gx = p.pos[0] - 0.5f, gy = p.pos[1] - 0.5f;
            int bx = (int)floorf(gx - 0.5f), by = (int)floorf(gy - 0.5f);
            float fx = gx - (float)bx, fy = gy - (float)by;
            float wx[3], wy[3];
            bspline(fx, wx);
            bspline(fy, wy);
            float sw = 0.0f, sfw = 0.0f, sfs = 0.0f, sgx = 0.0f, sgy = 0.0f, sgz = 0.0f;
            int more = (P->group_mask & 12u) != 0u;
            int groups = (P->group_mask & (P->group_mask - 1u)) != 0u;
            int sorts = group_sorts(s, Mgroup);
            float sbg = 0.0f;
            /* the share of steam around it (particles.comp:303-305) */
            int steamy = (P->caps & SP_CAP_STEAM) && Mgroup == 0;
            float sfv = 0.0f;
            int nmix = P->n_fam - 1;
            for (int b = 0; b < 3; b++)
                for (int a = 0; a < 3; a++) {
                    int ci = cidx(s, bx + a, by + b);
                    v4 q = s->comp[ci];
                    float wt = wx[a] * wy[b];
                    crowd += wt * q.w;
                    if (q.x > 0.0f) {
                        sw += wt;
                        sgx += wt * q.y;
                        if (more) {
                            sgy += wt * s->aux[ci].z;
                            sgz += wt * s->aux[ci].w;
                        }
                        if (sorts) sbg += wt * v4c(s->gbase[ci], Mgroup - 1);
                        if (steamy) sfv += wt * s->gbase[ci].w;
                        float sb = fam > 0 ? s->mixb[(size_t)ci * nmix + fam - 1] : q.z;
                        if (sb >= 0.0f) {
                            sfw += wt;
                            sfs += wt * sb;
                        }
                    }
                }
            /* Settling (particles.comp:332-351): the cells' gain and loss factors of the
             * particle's family with its scatter weights; water that is the primary family's
             * diluent takes part at concentration 0 and becomes the concentrate if it gains. */
            if (P->caps & SP_CAP_SETTLE) {
                int dil0 = fam < 0 && P->primary >= 0 && P->fams[0][1] == row;
                int sf = dil0 ? 0 : fam;
                if (sf >= 0 && (p.kind & SP_KIND_FRESH) == 0u && s->mats[dil0 ? P->primary : row].grain > 0.0f) {
                    float dsx = 0.0f, dsy = 0.0f;
                    for (int b = 0; b < 3; b++)
                        for (int a = 0; a < 3; a++) {
                            float wt = wx[a] * wy[b];
                            v2 d = s->dsettle[(size_t)cidx(s, bx + a, by + b) * P->n_fam + sf];
                            dsx += wt * d.x;
                            dsy += wt * d.y;
                        }
                    float s0 = dil0 ? 0.0f : p.s;
                    float s1 = clampf(s0 + (1.0f - s0) * dsx - s0 * dsy, 0.0f, 1.0f);
                    float c0 = th_on ? thermal_of(s, row, p.s).w : 0.0f;
                    if (dil0 && s1 > 0.0f) {
                        row = P->primary;
                        M = &s->mats[row];
                        Mgroup = M->group;
                        Mdil = M->diluent;
                        Mcw = M->chan_w;
                        Mcs = M->chan_s;
                        fam = 0;
                        p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                    }
                    if (!dil0 || s1 > 0.0f) p.s = s1;
                    if (th_on) p.t = keep_heat(p.t, c0, thermal_of(s, row, p.s).w);
                }
            }
            /* Heat (particles.comp:353-373): the change of enthalpy temperature the heat pass
             * worked out for the cells this particle scattered to, with the same weights, and
             * the exchange among the particles of its own cell against its own temperature
             * as it entered the cell's sums. */
            if (th_on && (p.kind & SP_KIND_FRESH) == 0u) {
                float dT = 0.0f;
                for (int b = 0; b < 3; b++)
                    for (int a = 0; a < 3; a++) dT += wx[a] * wy[b] * s->dheat[cidx(s, bx + a, by + b)].x;
                int ox = clampi((int)floorf(p.pos[0]), 0, s->W - 1), oy = clampi((int)floorf(p.pos[1]), 0, s->H - 1);
                v4 own = s->dheat[cidx(s, ox, oy)];
                if (own.y > 0.0f) {
                    v4 th = thermal_of(s, row, p.s);
                    int32_t q[2];
                    heat_quanta(th.w, fminf_(phase_of(p.t, th).x, boil_of(s, row, p.s)), q);
                    dT += own.y * (own.z - (float)q[1] / (float)q[0]) * SP_T_SCALE;
                }
                p.t += dT;
            }
            if (sw > 1e-4f) {
                /* Drift between phase groups (particles.comp:374-477). */
                sgx /= sw;
                sgy /= sw;
                sgz /= sw;
                float sgv[3] = {sgx, sgy, sgz};
                int gg = Mgroup;
                float fg = gg == 0 ? 1.0f - sgx - sgy - sgz : sgv[gg - 1];
                float others = 1.0f - fg;
                if (others > 1e-4f && groups && !held) {
                    float sgr[4] = {0, 0, 0, 0}, snu[4] = {0, 0, 0, 0}, snw[4] = {0, 0, 0, 0};
                    for (int b = 0; b < 3; b++)
                        for (int a = 0; a < 3; a++) {
                            int ci = cidx(s, bx + a, by + b);
                            v4 q = s->comp[ci];
                            if (q.x <= 0.0f) continue;
                            float wt = wx[a] * wy[b];
                            v4 gr = s->grho[ci];
                            sgr[0] += wt * gr.x;
                            sgr[1] += wt * gr.y;
                            sgr[2] += wt * gr.z;
                            sgr[3] += wt * gr.w;
                            v4 ax = s->aux[ci];
                            float fc[4] = {1.0f - q.y - ax.z - ax.w, q.y, ax.z, ax.w};
                            float top = fmaxf_(fmaxf_(fc[0], fc[1]), fmaxf_(fc[2], fc[3]));
                            float oh[4];
                            for (int k = 0; k < 4; k++) oh[k] = fc[k] == top ? 1.0f : 0.0f;
                            float ohs = fmaxf_(oh[0] + oh[1] + oh[2] + oh[3], 1.0f);
                            for (int k = 0; k < 4; k++) oh[k] /= ohs;
                            float wa = wt * ax.x;
                            for (int k = 0; k < 4; k++) {
                                snu[k] += wa * oh[k];
                                snw[k] += wt * oh[k];
                            }
                        }
                    for (int k = 0; k < 4; k++) sgr[k] /= sw;
                    float rho_g = fg > 1e-4f ? sgr[gg] / fg : row_density(s, row, p.s);
                    float rho_o = (sgr[0] + sgr[1] + sgr[2] + sgr[3] - sgr[gg]) / others;
                    float drho = rho_g - rho_o;
                    int pcx = clampi((int)p.pos[0], 1, s->W - 2), pcy = clampi((int)p.pos[1], 1, s->H - 2);
                    int k = cidx(s, pcx, pcy);
                    float fsv[4] = {1.0f - sgx - sgy - sgz, sgx, sgy, sgz};
                    int G = fsv[0] >= fmaxf_(fsv[1], fmaxf_(fsv[2], fsv[3])) ? 0 : fsv[1] >= fmaxf_(fsv[2], fsv[3]) ? 1 : fsv[2] >= fsv[3] ? 2 : 3;
                    float rho_c = sgr[G] / fmaxf_(fsv[G], 1e-4f);
                    float nu_c = snw[G] > 1e-6f ? snu[G] / snw[G] : s->aux[k].x;
                    float vt = terminal(s, drho, rho_c, nu_c, P->sep_tau);
                    float sep_max = P->corr[2] > 0.0f ? P->cfl / dt : 1e30f;
                    slip = -fminf_(vt, sep_max) * signf(drho) * others;
                    float fs[4] = {1.0f - sgx - sgy - sgz, sgx, sgy, sgz};
                    fs[gg] = -1.0f;
                    int h = fs[0] >= fmaxf_(fs[1], fmaxf_(fs[2], fs[3])) ? 0 : fs[1] >= fmaxf_(fs[2], fs[3]) ? 1 : fs[2] >= fs[3] ? 2 : 3;
                    float bond = fabsf(drho) * P->gravity / fmaxf_(P->pair[gg][h], 1e-9f);
                    float hold = clampf(1.0f - bond, 0.0f, 1.0f);
                    if (slip != 0.0f && hold > 0.0f) {
                        int hi = gg > h ? gg : h;
                        int W = s->W;
                        float grx, gry;
                        if (hi == 1) {
                            grx = 0.5f * (s->surf[k + 1].y - s->surf[k - 1].y);
                            gry = 0.5f * (s->surf[k + W].y - s->surf[k - W].y);
                        } else if (hi == 2) {
                            grx = 0.5f * (s->surf2[k + 1].x - s->surf2[k - 1].x);
                            gry = 0.5f * (s->surf2[k + W].x - s->surf2[k - W].x);
                        } else {
                            grx = 0.5f * (s->surf2[k + 1].y - s->surf2[k - 1].y);
                            gry = 0.5f * (s->surf2[k + W].y - s->surf2[k - W].y);
                        }
                        float resolved = 1.0f - clampf((0.12f - len2(grx, gry)) / (0.12f - 0.04f), 0.0f, 1.0f);
                        slip *= 1.0f - hold * resolved;
                    }
                }
                /* Drift within a phase group (particles.comp:490-502). */
                if (sorts && !held) {
                    float fgi = groups ? sgv[Mgroup - 1] : 1.0f;
                    if (fgi > 1e-4f) {
                        float rmean = sbg / sw / fgi;
                        float drho = M->rho - rmean;
                        int pcx = clampi((int)p.pos[0], 0, s->W - 1), pcy = clampi((int)p.pos[1], 0, s->H - 1);
                        v4 cq = s->cell[cidx(s, pcx, pcy)];
                        float nu_c = cq.y > 0.0f ? cq.z / cq.y : s->aux[cidx(s, pcx, pcy)].x;
                        float vt = terminal(s, drho, rmean, nu_c, P->sep_tau);
                        float sep_max = P->corr[2] > 0.0f ? P->cfl / dt : 1e30f;
                        slip -= fminf_(vt, sep_max) * signf(drho);
                    }
                }
                /* Steam rises through the water around it (particles.comp:505-517): a bubble
                 * of the unresolved size at its terminal velocity in water, by the share of
                 * water; the water drifts down by the share of steam. */
                if (steamy && !held) {
                    float fv = clampf(sfv / sw, 0.0f, 1.0f);
                    if (fv > 1e-4f) {
                        float rw = s->mats[SP_ROW_WATER].rho;
                        float vt = terminal(s, s->mats[SP_ROW_STEAM].rho - rw, rw, expf(s->mats[SP_ROW_WATER].ln_nu), P->sep_tau);
                        float sep_max = P->corr[2] > 0.0f ? P->cfl / dt : 1e30f;
                        vt = fminf_(vt, sep_max);
                        slip += row == SP_ROW_STEAM ? vt * (1.0f - fv) : -vt * fv;
                    }
                }
                /* Miscible exchange within a family (particles.comp:527-543). */
                int diluent0 = fam < 0 && P->primary >= 0 && P->fams[0][1] == row;
                if (fam >= 0 || diluent0) {
                    if (sfw > 1e-4f) {
                        float s0 = diluent0 ? 0.0f : p.s;
                        int rr = diluent0 ? P->primary : row;
                        float s1 = clampf(s0 + (sfs / sfw - s0) * (1.0f - expf(-s->mats[rr].mix_rate * dt)), 0.0f, 1.0f);
                        float c0 = th_on ? thermal_of(s, row, p.s).w : 0.0f;
                        if (!diluent0) {
                            p.s = s1;
                        } else if (s1 > 0.0f) {
                            row = P->primary;
                            M = &s->mats[row];
                            Mgroup = M->group;
                            Mdil = M->diluent;
                            Mcw = M->chan_w;
                            Mcs = M->chan_s;
                            p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                            p.s = s1;
                        }
                        /* the same heat content at the new heat capacity (particles.comp:541) */
                        if (th_on) p.t = keep_heat(p.t, c0, thermal_of(s, row, p.s).w);
                    }
                }
                /* A pure diluent joins a family at concentration 0 (particles.comp:544-569). */
                if (fam < 0 && (P->caps & 1u) && (int)(p.kind & 255u) - 1 == row && s->mats[row].fam < 0) {
                    for (int j = P->primary >= 0 ? 1 : 0; j < P->n_fam; j++) {
                        if (P->fams[j][1] != row) continue;
                        float found = 0.0f;
                        if (j == 0) {
                            found = sfw;
                        } else {
                            for (int b = 0; b < 3; b++)
                                for (int a = 0; a < 3; a++)
                                    if (s->mixb[(size_t)cidx(s, bx + a, by + b) * nmix + j - 1] >= 0.0f) found += wx[a] * wy[b];
                        }
                        if (found > 1e-4f) {
                            row = P->fams[j][0];
                            M = &s->mats[row];
                            Mgroup = M->group;
                            Mdil = M->diluent;
                            Mcw = M->chan_w;
                            Mcs = M->chan_s;
                            p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                            p.s = 0.0f;
                            break;
                        }
                    }
                }
            }
        }
        (void)Mdil;
        /* held: the density correction only where the body holds a clump (particles.comp:573-576) */
        if (held && crowd <= SP_REST_PHI + fmaxf_(P->corr[1], 0.05f)) {
            corrx = 0.0f;
            corry = 0.0f;
        }
        /* Drains (particles.comp:578-592): each particle in a zone of negative rate leaves with
         * probability -rate dt (1 - 0.75 r^2). */
        if (P->ext[3] > 0.0f) {
            for (int i = 0; i < (int)P->ext[3]; i++) {
                const float* z = P->current_zone[i];
                if (z[3] >= 0.0f) continue;
                float dzx = p.pos[0] - z[0], dzy = p.pos[1] - z[1];
                float r2 = (dzx * dzx + dzy * dzy) / (z[2] * z[2]);
                if (r2 < 1.0f) {
                    float u = (float)(hash_u(p.id * 2246822519u + s->S.substeps * 3266489917u + (uint32_t)i) >> 8) / 16777216.0f;
                    if (u < -z[3] * (1.0f - 0.75f * r2) * dt) {
                        s->part[id].kind = SP_KIND_DEAD;
                        return;
                    }
                }
            }
        }
        /* The mouse tool (particles.comp:594-631) inside its radius: Erase removes the
         * particle, Heat and Cool add or take heat_power (1 - r^2) dt kelvin for water (not
         * below -150 deg C), booked in the ledger; the others drive the velocity towards the
         * tool's target by strength (1 - r^2) dt, less in a crowd (but the nozzle). */
        int pushed = 0;
        if (P->tool_mode != 0) {
            float tdx = P->tool_pos[0] - p.pos[0], tdy = P->tool_pos[1] - p.pos[1];
            float r2 = (tdx * tdx + tdy * tdy) / (P->tool_radius * P->tool_radius);
            if (r2 < 1.0f) {
                if (P->tool_mode == 2) {
                    s->part[id].kind = SP_KIND_DEAD;
                    return;
                }
                float fall = 1.0f - r2;
                if (P->tool_mode == 5 || P->tool_mode == 6) {
                    if (th_on) {
                        float e = P->heat_power * fall * dt;
                        v4 tth = thermal_of(s, row, p.s);
                        float C = tth.w;
                        if (P->tool_mode == 6) e = fminf_(e, fmaxf_(phase_of(p.t, tth).x + 150.0f, 0.0f) * C);
                        p.t += (P->tool_mode == 5 ? e : -e) / C;
                        ledger_add(s, P->tool_mode == 5 ? 4 : 5, e * P->inv_ppc);
                    }
                } else {
                    float k = clampf(P->tool_strength * fall * dt, 0.0f, 1.0f);
                    if (P->tool_crowd_max > 0.0f && P->tool_mode != 4)
                        k *= clampf((P->tool_crowd_max - crowd) / (0.25f * P->tool_crowd_max), 0.0f, 1.0f);
                    v2 tt = tool_target(P, tdx, tdy);
                    vx = mixf(vx, tt.x, k);
                    vy = mixf(vy, tt.y, k);
                    pushed = 1;
                }
            }
        }
        /* Boiling (particles.comp:633-731): past its boiling point liquid takes up latent
         * heat, its share the vapour quality x. Aqueous liquid in a tank that makes steam turns
         * into steam at x = steam.w (keeping mass, heat and the row it boiled from in KIND_SRC);
         * steam condenses below steam.x, and at the free surface escapes as a puff with
         * probability x, else gives up its latent heat there and stays as its liquid. Any other
         * liquid leaves at x = 1, leaving its residue, one particle per residue share. */
        if (th_on && (P->caps & SP_CAP_BOILS)) {
            v4 th = thermal_of(s, row, p.s);
            int dil = s->mats[row].diluent;
            float bo_t = s->mats[row].th2[2], bo_l = s->mats[row].th2[3];
            if (dil >= 0) {
                bo_t = mixf(s->mats[dil].th2[2], bo_t, p.s);
                bo_l = mixf(s->mats[dil].th2[3], bo_l, p.s);
            }
            int vapour = (P->caps & SP_CAP_STEAM) && Mgroup == 0 && (P->steam2[1] > 0.5f || row == SP_ROW_STEAM);
            int leave = 0;
            if (vapour && row == SP_ROW_STEAM) {
                float x = (p.t - bo_t) / fmaxf_(bo_l, 1e-3f);
                int pcx = clampi((int)p.pos[0], 0, s->W - 1), pcy = clampi((int)p.pos[1], 0, s->H - 1);
                int surface = s->cell[cidx(s, pcx, pcy)].w <= 0.0f || s->cell[cidx(s, pcx, clampi(pcy + 1, 0, s->H - 1))].w <= 0.0f;
                float u = (float)(hash_u(p.id ^ hash_u(s->S.substeps ^ 0x2545f491u)) >> 8) / 16777216.0f;
                if (surface && u < x) {
                    leave = 1;
                } else if (x < P->steam[0] || surface) {
                    if (surface) {
                        ledger_add(s, 6, fmaxf_(th.w * (p.t - bo_t), 0.0f) * P->inv_ppc);
                        p.t = bo_t;
                    }
                    uint32_t src = (p.kind & SP_KIND_SRC) >> 8;
                    row = src > 0u ? (int)src - 1 : SP_ROW_WATER;
                    M = &s->mats[row];
                    Mgroup = M->group;
                    Mdil = M->diluent;
                    Mcw = M->chan_w;
                    Mcs = M->chan_s;
                    fam = M->fam;
                    p.kind = (p.kind & ~(255u | SP_KIND_SRC)) | (uint32_t)(row + 1);
                    p.t = keep_heat(p.t, th.w, thermal_of(s, row, p.s).w);
                }
            } else if (p.t >= bo_t + bo_l * (vapour ? P->steam[3] : 1.0f)) {
                if (vapour) {
                    float c1 = thermal_of(s, SP_ROW_STEAM, 1.0f).w;
                    s->ledger[16 + 2 * SP_MAX_MATERIALS + row] += 1u;
                    p.kind = (p.kind & ~(255u | SP_KIND_SRC)) | ((uint32_t)(row + 1) << 8) | (uint32_t)(SP_ROW_STEAM + 1);
                    row = SP_ROW_STEAM;
                    M = &s->mats[row];
                    Mgroup = M->group;
                    Mdil = M->diluent;
                    Mcw = M->chan_w;
                    Mcs = M->chan_s;
                    fam = -1;
                    p.t = keep_heat(p.t, th.w, c1);
                } else {
                    leave = 1;
                }
            }
            if (leave) {
                uint32_t src = row == SP_ROW_STEAM ? (p.kind & SP_KIND_SRC) >> 8 : (uint32_t)(row + 1);
                int from = src > 0u ? (int)src - 1 : SP_ROW_WATER;
                int res = s->mats[from].residue;
                float share = s->mats[from].residue_share * (s->mats[from].diluent >= 0 ? p.s : 1.0f);
                float e = th.w * (p.t + 273.15f);
                s->ledger[16 + row] += 1u;
                if (res >= 0 && (float)(hash_u(p.id ^ 0x9e3779b9u) >> 8) / 16777216.0f < share) {
                    s->ledger[16 + SP_MAX_MATERIALS + res] += 1u;
                    row = res;
                    M = &s->mats[row];
                    Mgroup = M->group;
                    Mdil = M->diluent;
                    Mcw = M->chan_w;
                    Mcs = M->chan_s;
                    fam = M->fam;
                    p.kind = (uint32_t)(row + 1);
                    p.s = 1.0f;
                    p.t = bo_t - thermal_of(s, row, 1.0f).z; /* the residue is a solid */
                    e -= thermal_of(s, row, 1.0f).w * (p.t + 273.15f);
                    ledger_add(s, 6, fmaxf_(e, 0.0f) * P->inv_ppc);
                } else {
                    ledger_add(s, 6, fmaxf_(e, 0.0f) * P->inv_ppc);
                    float fume = row == SP_ROW_STEAM ? 1.0f : s->mats[row].fume;
                    if (fume > 0.0f) become_puff(s, id, p, vx, vy, fume);
                    else s->part[id].kind = SP_KIND_DEAD;
                    return;
                }
            }
        }
        /* Heat made inside (particles.comp:733-739), counted in the ledger. */
        if (th_on && s->mats[row].rx[2] > 0.0f) {
            float e = s->mats[row].rx[2] * s->mats[row].th[0] * (s->mats[row].diluent >= 0 ? p.s : 1.0f) * dt * P->heat_rate;
            p.t += e / thermal_of(s, row, p.s).w;
            ledger_add(s, 7, e * P->inv_ppc);
        }
        /* A reaction (particles.comp:740-770): where the particle's cell also holds its
         * partner it turns into the product with probability rate x dt, warmed by its heat
         * of reaction; with no product it is used up and counted. */
        if ((P->caps & SP_CAP_REACT) && s->mats[row].ic[0] >= 0) {
            int qx = clampi((int)p.pos[0], 0, s->W - 1), qy = clampi((int)p.pos[1], 0, s->H - 1);
            if ((s->rmask[cidx(s, qx, qy)] & (1u << (uint32_t)s->mats[row].ic[2])) != 0u) {
                float own = s->mats[row].diluent >= 0 ? p.s : 1.0f;
                float u = (float)(hash_u(p.id * 0x9e3779b1u ^ hash_u(s->S.substeps + 0x51ed270bu)) >> 8) / 16777216.0f;
                if (u < s->mats[row].rx[0] * dt && own > SP_REACT_MIN) {
                    float c0 = thermal_of(s, row, p.s).w;
                    float e = s->mats[row].rx[1] * s->mats[row].th[0] * own;
                    if (th_on) {
                        p.t += e / c0;
                        ledger_add(s, 7, e * P->inv_ppc);
                    }
                    int into = s->mats[row].ic[1];
                    if (into < 0) {
                        if (th_on) ledger_add(s, 6, fmaxf_(c0 * (p.t + 273.15f), 0.0f) * P->inv_ppc);
                        s->ledger[16 + row] += 1u;
                        s->part[id].kind = SP_KIND_DEAD;
                        return;
                    }
                    row = into;
                    M = &s->mats[row];
                    Mgroup = M->group;
                    Mdil = M->diluent;
                    Mcw = M->chan_w;
                    Mcs = M->chan_s;
                    fam = M->fam;
                    p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                    if (M->diluent < 0) p.s = 1.0f;
                    if (th_on) p.t = keep_heat(p.t, c0, thermal_of(s, row, p.s).w);
                }
            }
        }
        /* Evaporation below the boiling point (particles.comp:771-807): a particle of a
         * volatile liquid in a cell open to the air leaves with probability rate x dt, the
         * rate doubling every 10 K, a concentrate's scaled by its concentration; it takes its
         * heat with it (ledger entry 6) and is counted. */
        if (s->mats[row].rx[3] > 0.0f) {
            int qx = clampi((int)p.pos[0], 1, s->W - 2), qy = clampi((int)p.pos[1], 1, s->H - 2);
            int qi = cidx(s, qx, qy);
            int open_air = 0;
            for (int q = 0; q < 4; q++) {
                int n = qi + (q == 0 ? -1 : q == 1 ? 1 : q == 2 ? -s->W : s->W);
                if (s->cell[n].w <= 0.0f && s->solid[n].w > 0.5f) open_air = 1;
            }
            if (open_air) {
                v4 th = thermal_of(s, row, p.s);
                float Tp = th_on ? phase_of(p.t, th).x : P->heat_x;
                float rate = s->mats[row].rx[3] * (s->mats[row].diluent >= 0 ? p.s : 1.0f) *
                             exp2f(clampf((Tp - 20.0f) / 10.0f, -6.0f, 5.0f));
                float u = (float)(hash_u(p.id * 0x85ebca6bu ^ hash_u(s->S.substeps + 0x27d4eb2fu)) >> 8) / 16777216.0f;
                if (u < rate * dt) {
                    if (th_on) ledger_add(s, 6, fmaxf_(th.w * (p.t + 273.15f), 0.0f) * P->inv_ppc);
                    s->ledger[16 + row] += 1u;
                    /* a liquid that fumes leaves a puff of mist */
                    if (s->mats[row].fume > 0.0f) become_puff(s, id, p, vx, vy, s->mats[row].fume);
                    else s->part[id].kind = SP_KIND_DEAD;
                    return;
                }
            }
        }
        /* Currents (particles.comp:809-829): v gains dir max(speed - v.dir, 0)
         * (1 - exp(-rate (1 - r^2)^2 dt)) drive inside each zone of positive rate. */
        if (P->ext[3] > 0.0f) {
            float drive = P->tool_crowd_max > 0.0f ? clampf((P->tool_crowd_max - crowd) / (0.25f * P->tool_crowd_max), 0.0f, 1.0f) : 1.0f;
            for (int i = 0; i < (int)P->ext[3] && drive > 0.0f; i++) {
                const float* z = P->current_zone[i];
                if (z[3] <= 0.0f) continue;
                float dzx = p.pos[0] - z[0], dzy = p.pos[1] - z[1];
                float r2 = (dzx * dzx + dzy * dzy) / (z[2] * z[2]);
                if (r2 < 1.0f) {
                    float tx = P->current_vel[i][0], ty = P->current_vel[i][1];
                    float speed = len2(tx, ty);
                    float m = fmaxf_(speed, 1e-6f);
                    float dx = tx / m, dy = ty / m;
                    float fall = (1.0f - r2) * (1.0f - r2);
                    float gain = fmaxf_(speed - (vx * dx + vy * dy), 0.0f);
                    float relax = 1.0f - expf(-z[3] * fall * dt);
                    vx += dx * gain * relax * drive;
                    vy += dy * gain * relax * drive;
                }
            }
        }
        /* Speed cap (particles.comp:831-835). */
        float vcap = P->ext[0];
        {
            float sp = len2(vx, vy);
            if (sp > vcap) {
                float k = vcap / sp;
                vx *= k;
                vy *= k;
            }
        }
        /* Midpoint advection (particles.comp:847-865). */
        float vax = vx, vay = vy;
        if (P->ext[1] >= 2.0f && !pushed && !held) {
            float h = 0.5f * dt;
            float ex = (vx + corrx) + 0.0f, ey = (vy + corry) + slip;
            float xmx = clampf(p.pos[0] + h * ex, 2.01f, (float)s->W - 2.01f);
            float xmy = clampf(p.pos[1] + h * ey, 2.01f, (float)s->H - 2.01f);
            v2 vv = {vx, vy};
            v2 here = sample_vel(s, p.pos[0], p.pos[1], vv);
            v2 mid = sample_vel(s, xmx, xmy, here);
            vax = vx + mid.x - here.x;
            vay = vy + mid.y - here.y;
            float sp = len2(vax, vay);
            if (sp > vcap) {
                float k = vcap / sp;
                vax *= k;
                vay *= k;
            }
            if (crowd < SP_LIQUID_THR && (vax * vax + vay * vay) < 0.01f * (vx * vx + vy * vy)) {
                vx = vax;
                vy = vay;
            }
        }
        float dispx = (vax + corrx) + 0.0f, dispy = (vay + corry) + slip;
        {
            float mine = len2(dispx, dispy);
            if (crowd < SP_LIQUID_THR) mine /= P->ext[2];
            if (mine > *umax_part) *umax_part = mine;
        }
        float sdx = dt * dispx, sdy = dt * dispy;
        int pieces = (int)ceilf(len2(sdx, sdy));
        pieces = clampi(pieces, 1, 32);
        sdx /= (float)pieces;
        sdy /= (float)pieces;
        float x = p.pos[0], y = p.pos[1];
        for (int k = 1; k < pieces; k++) {
            x += sdx;
            y += sdy;
            collide(s, &x, &y, &vx, &vy, 1);
        }
        x += sdx;
        y += sdy;
        float excess = clampf(crowd - 1.5f, 0.0f, 1.0f);
        if (excess > 0.0f && P->crowd_jitter > 0.0f) {
            uint32_t hh = hash_u(p.id ^ hash_u(f2u(p.pos[0]) ^ s->S.substeps));
            x += ((float)(hh & 0xffffu) / 65535.0f - 0.5f) * P->crowd_jitter * excess;
            y += ((float)(hh >> 16) / 65535.0f - 0.5f) * P->crowd_jitter * excess;
        }
        collide(s, &x, &y, &vx, &vy, 1);
        int again = collide(s, &x, &y, &vx, &vy, 1);
        if (again) {
            float gx_, gy_;
            if (sample_sdf(s, x, y, &gx_, &gy_) < -1.0f) collide(s, &x, &y, &vx, &vy, 0);
        }
        x = clampf(x, 2.01f, (float)s->W - 2.01f);
        y = clampf(y, 2.01f, (float)s->H - 2.01f);
        p.pos[0] = x;
        p.pos[1] = y;
        p.vel[0] = vx;
        p.vel[1] = vy;
        for (int k = 0; k < 4; k++) p.c[k] = C[k];
        p.kind &= ~SP_KIND_FRESH;
        s->part[id] = p;
    }
    float rho_p = row_density(s, row, p.s);
    /* Thermal state for the scatter (particles.comp:951-969): temperature capped at the
     * boiling point, the expansion and the solid's density where frozen, and the particle's
     * quanta into its own cell's sums for the exchange inside the cell. */
    float CT = 0.0f;
    if (th_on) {
        const sp_gmat* Mr = &s->mats[row];
        v4 th = thermal_of(s, row, p.s);
        v2 ph = phase_of(p.t, th);
        float tb = boil_of(s, row, p.s);
        rho_p *= expansion(s, row, p.s, ph.x);
        if (ph.y > 0.0f) {
            float rs = Mr->diluent >= 0 ? mixf(s->mats[Mr->diluent].rho_solid, Mr->rho_solid, p.s) : Mr->rho_solid;
            rho_p = mixf(rho_p, rs, ph.y);
        }
        CT = th.w * (fminf_(ph.x, tb) + 273.15f) / SP_T_SCALE;
        int hx = clampi((int)floorf(p.pos[0]), 0, s->W - 1), hy = clampi((int)floorf(p.pos[1]), 0, s->H - 1);
        int32_t q[2];
        heat_quanta(th.w, fminf_(ph.x, tb), q);
        size_t hc = (size_t)cidx(s, hx, hy);
        s->hcell[2 * hc] += q[0];
        s->hcell[2 * hc + 1] += q[1];
    }
    g->live = 1;
    g->px = p.pos[0];
    g->py = p.pos[1];
    g->vsx = p.vel[0] * dt;
    g->vsy = p.vel[1] * dt;
    g->Csx = p.c[0] * dt;
    g->Csy = p.c[1] * dt;
    g->Csz = p.c[2] * dt;
    g->Csw = p.c[3] * dt;
    g->rho = rho_p;
    g->sc = p.s;
    g->cw = Mcw;
    g->cs = Mcs;
    g->CT = CT;
    g->ct = th_on ? s->P.chan_t : -1;
    g->cg = th_on ? group_heat_chan(s, s->mats[row].group) : -1;
    /* the stress to scatter (particles.comp:983-984) */
    g->ce = el_on && (tau.x != 0.0f || tau.y != 0.0f || tau.z != 0.0f || tau.w != 0.0f) ? s->P.chan_e : -1;
    g->te[0] = tau.x / s->P.el_g * SP_STRESS_FIX;
    g->te[1] = tau.y / s->P.el_g * SP_STRESS_FIX;
    g->te[2] = tau.z / s->P.el_g * SP_STRESS_FIX;
}

static void stage_particles(or_sim* s, int gather) {
    float umax = u2f(s->S.umax_part);
    for (uint32_t id = 0; id < s->n_slots; id++) {
        scatter_in g;
        advance(s, id, gather, &g, &umax);
        if (g.live) deposit(s, &g);
    }
    /* atomicMax on float bits of non-negative floats (particles.comp:872-877). */
    if (umax > u2f(s->S.umax_part)) s->S.umax_part = f2u(umax);
}

static void stage_clear_acc(or_sim* s) {
    memset(s->acc, 0, sizeof(int32_t) * (size_t)s->N * s->P.acc_stride);
    if (thermal(s)) memset(s->hcell, 0, sizeof(int32_t) * 2 * (size_t)s->N);
}

/* ---- grid_prep.comp (grid_prep.comp:13-263) ------------------------------------------- */

static int packed(const or_sim* s, int i) {
    v4 sol = s->solid[i];
    float phi_f = (float)s->acc[(size_t)i * s->P.acc_stride + 4] / SP_MASS_SCALE * s->P.inv_ppc;
    return sol.w > 0.02f && phi_f > 0.04f && phi_f + sol.z > SP_LIQUID_THR;
}

static void stage_grid_prep(or_sim* s) {
    const sp_gparams* P = &s->P;
    int W = s->W, H = s->H;
    int nmix = P->n_fam - 1;
    int any_sorts = group_sorts(s, 1) || group_sorts(s, 2) || group_sorts(s, 3);
    int multi = (P->group_mask & (P->group_mask - 1u)) != 0u;
    const int th_on = thermal(s);
    float dt = s->S.dt;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            size_t k = (size_t)i * P->acc_stride;
            int um = s->acc[k], umom = s->acc[k + 1], vm = s->acc[k + 2], vmom = s->acc[k + 3];
            int cw = s->acc[k + 4];
            v4 sol = s->solid[i];
            v4 o = {0.0f, 0.0f, SP_FACE_EMPTY, SP_FACE_EMPTY};
            if (sol.x <= 0.0f) {
                o.z = SP_FACE_CLOSED;
            } else if (um >= SP_MIN_FACE_MASS) {
                o.x = ((float)umom / SP_MOM_SCALE) / ((float)um / SP_MASS_SCALE) / dt;
                o.z = 0.0f;
            }
            if (sol.y <= 0.0f) {
                o.w = SP_FACE_CLOSED;
            } else if (vm >= SP_MIN_FACE_MASS) {
                o.y = ((float)vmom / SP_MOM_SCALE) / ((float)vm / SP_MASS_SCALE) / dt - P->gravity * dt;
                o.w = 0.0f;
            }
            s->vel[i] = o;
            s->vel0[i].x = o.x;
            s->vel0[i].y = o.y;
            float phi_f = (float)cw / SP_MASS_SCALE * P->inv_ppc;
            float f[4] = {0, 0, 0, 0}, rg[4] = {0, 0, 0, 0}, lg[4] = {0, 0, 0, 0};
            float bg[3] = {0, 0, 0};
            float sigma = 0.0f, cx_ = 0.0f, cy_ = 0.0f, fam0 = -1.0f;
            float rho = 0.0f, mu = 0.0f, rho_c = 0.0f, nu_c = 0.0f;
            int major = 0, top_row = -1;
            float top_share = 0.0f, fv = 0.0f; /* fv: the share of steam */
            float T = P->heat_x, kappa = 0.0f, emis = 0.0f, frozen = 0.0f, Craw = 0.0f;
            v4 nn = {0.0f, -1.0f, 1.0f, 0.0f};  /* rh: the cell's non-Newtonian material */
            float chi = 0.0f;
            for (int q = 0; q < nmix; q++) s->mixb[(size_t)i * nmix + q] = -1.0f;
            int react = (P->caps & SP_CAP_REACT) != 0u;
            if (cw <= 0 && react) s->rmask[i] = 0u; /* grid_prep.comp:71 */
            /* each family's occupancy of the cell for settling (grid_prep.comp:69-70, 147) */
            int settle = (P->caps & SP_CAP_SETTLE) != 0u;
            if (settle)
                for (int q = 0; q < P->n_fam; q++) s->famocc[(size_t)i * P->n_fam + q] = 0.0f;
            if (cw > 0) {
                int others = 0;
                for (int j = 0; j < P->n_present; j++) {
                    int ch = s->mats[P->present[j]].chan_w;
                    if (ch >= 0) others += s->acc[k + ch];
                }
                float sh[SP_MAX_PRESENT], scn[SP_MAX_PRESENT];
                float Csh = 0.0f, Cg[4] = {0, 0, 0, 0};
                for (int j = 0; j < P->n_present; j++) {
                    sh[j] = 0.0f;
                    scn[j] = 1.0f;
                    int r = P->present[j];
                    const sp_gmat* m = &s->mats[r];
                    if (m->chan_w < 0 && P->primary >= 0 && r != P->primary) continue;
                    int w = m->chan_w >= 0 ? s->acc[k + m->chan_w] : (cw - others > 0 ? cw - others : 0);
                    if (w <= 0) continue;
                    sh[j] = w >= cw ? 1.0f : clampf((float)w / (float)cw, 0.0f, 1.0f);
                    if (m->diluent >= 0 && m->chan_s >= 0) scn[j] = clampf((float)s->acc[k + m->chan_s] / (float)w, 0.0f, 1.0f);
                    /* with heat: the cell's heat capacity, conductivity and emissivity
                     * (grid_prep.comp:100-108) */
                    if (th_on) {
                        float cr = sh[j] * thermal_of(s, r, scn[j]).w;
                        Csh += cr;
                        Cg[m->group] += cr;
                        float kp = m->th[1];
                        kappa += sh[j] * (m->diluent >= 0 ? mixf(s->mats[m->diluent].th[1], kp, scn[j]) : kp);
                        float em = m->th2[1];
                        emis += sh[j] * (m->diluent >= 0 ? mixf(s->mats[m->diluent].th2[1], em, scn[j]) : em);
                    }
                }
                /* which rows the cell holds, for the reactions (grid_prep.comp:110-118) */
                if (react) {
                    uint32_t mask = 0u;
                    for (int j = 0; j < P->n_present; j++)
                        if (sh[j] * (s->mats[P->present[j]].diluent >= 0 ? scn[j] : 1.0f) > SP_REACT_MIN) mask |= 1u << (uint32_t)j;
                    s->rmask[i] = mask;
                }
                /* The temperature of each phase group from its own sum (group_temperatures,
                 * common.glsl:598-622); T leaves as the cell's mean. */
                float Tg[4] = {T, T, T, T};
                if (th_on && Csh > 0.0f) {
                    int Htot = s->acc[k + P->chan_t];
                    T = (float)Htot * SP_T_SCALE / (Csh * (float)cw) - 273.15f;
                    for (int q = 0; q < 4; q++) Tg[q] = T;
                    int rest = Htot, first = -1;
                    for (int q = 0; q < 4; q++) {
                        int ch = group_heat_chan(s, q);
                        if (ch >= 0) {
                            int h = s->acc[k + ch];
                            rest -= h;
                            if (Cg[q] > 0.0f) Tg[q] = (float)h * SP_T_SCALE / (Cg[q] * (float)cw) - 273.15f;
                        } else if (first < 0 && (P->group_mask & (1u << q)) != 0u) {
                            first = q;
                        }
                    }
                    if (first >= 0 && Cg[first] > 0.0f) Tg[first] = (float)rest * SP_T_SCALE / (Cg[first] * (float)cw) - 273.15f;
                }
                for (int j = 0; j < P->n_present; j++) {
                    float share = sh[j];
                    if (share <= 0.0f) continue;
                    int r = P->present[j];
                    const sp_gmat* m = &s->mats[r];
                    float sc = scn[j];
                    float rho_m = m->rho, lnu = m->ln_nu, sig = m->sigma;
                    float Tm = Tg[m->group];
                    if (th_on) lnu += temp_law(s, r, Tm);
                    float xp = th_on ? expansion(s, r, sc, Tm) : 1.0f;
                    float csx = m->cos_c, csy = m->sin_c;
                    if (m->diluent >= 0) {
                        const sp_gmat* d = &s->mats[m->diluent];
                        rho_m = mixf(d->rho, rho_m, sc);
                        float bs = m->visc_exp == 1.0f ? sc : powf(sc, m->visc_exp);
                        lnu = mixf(th_on ? d->ln_nu + temp_law(s, m->diluent, Tm) : d->ln_nu, lnu, bs);
                        sig = mixf(d->sigma, sig, sc);
                        csx = mixf(d->cos_c, csx, sc);
                        csy = mixf(d->sin_c, csy, sc);
                        int fm = m->fam;
                        if (fm == 0) fam0 = sc;
                        else if (fm > 0) s->mixb[(size_t)i * nmix + fm - 1] = sc;
                        if (settle && fm >= 0) s->famocc[(size_t)i * P->n_fam + fm] = share * phi_f;
                    }
                    rho_m *= xp;
                    /* frozen liquid: the frozen viscosity and the solid's density
                     * (grid_prep.comp:150-159) */
                    if (th_on) {
                        v4 th = thermal_of(s, r, sc);
                        float fs = th.y > -1e8f ? clampf((th.y - Tm) / fmaxf_(th.y - th.x, 1e-3f), 0.0f, 1.0f) : 0.0f;
                        if (fs > 0.0f) {
                            lnu = mixf(lnu, fmaxf_(lnu, P->ln_rigid), fs);
                            float rs = m->rho_solid;
                            rho_m = mixf(rho_m, m->diluent >= 0 ? mixf(s->mats[m->diluent].rho_solid, rs, sc) : rs, fs);
                            frozen += share * fs;
                        }
                    }
                    /* the non-Newtonian material with the largest share (grid_prep.comp:162) */
                    if ((m->rheo != 0 || m->temp[3] > 0.0f || m->gr[0] > 0.0f) && share > nn.x)
                        nn = (v4){share, (float)r, sc, 0.0f};
                    /* steam's share, which never makes it the top material (grid_prep.comp:163-167) */
                    if (r == SP_ROW_STEAM) {
                        fv = share;
                    } else if (share > top_share) {
                        top_share = share;
                        top_row = r;
                    }
                    chi += share * m->gr[3]; /* the cell's susceptibility (grid_prep.comp:168) */
                    int g = m->group;
                    f[g] += share;
                    rg[g] += share * rho_m;
                    if (g > 0) bg[g - 1] += share * m->rho;
                    lg[g] += share * lnu;
                    sigma += share * sig;
                    cx_ += share * csx;
                    cy_ += share * csy;
                }
                for (int g = 0; g < 4; g++) {
                    if (f[g] <= 0.0f) continue;
                    rho += rg[g];
                    mu += rg[g] * expf(lg[g] / f[g]);
                    if (f[g] > f[major]) major = g;
                }
                if (f[major] > 0.0f) {
                    rho_c = rg[major] / f[major];
                    nu_c = expf(lg[major] / f[major]);
                }
                Craw = Csh * (float)cw / SP_MASS_SCALE;
            }
            float phi_t = phi_f + sol.z;
            int has = phi_f > 0.04f && sol.w > 0.02f;
            float level = has ? phi_t - SP_LIQUID_THR : -SP_LIQUID_THR;
            if (has && level <= 0.0f && cx > 0 && cy > 0 && cx < W - 1 && cy < H - 1) {
                if ((packed(s, i - 1) && packed(s, i + 1)) || (packed(s, i - W) && packed(s, i + W))) level = 0.05f;
                else if (P->gravity > 0.0f && nn.x <= 0.0f && phi_t > 0.25f && packed(s, i - W)) level = 0.05f;
                /* with a magnet, "lying on" is towards gravity and the last substep's magnetic
                 * pull together (grid_prep.comp:211-221) */
                else if ((P->caps & SP_CAP_MAGNET) && nn.x <= 0.0f && phi_t > 0.25f && chi > 0.0f) {
                    float km = chi / P->mag[1] / fmaxf_(rho, 0.05f);
                    float px = -km * 0.5f * (s->magc[i + 1].y - s->magc[i - 1].y) - 0.0f;
                    float py = -km * 0.5f * (s->magc[i + W].y - s->magc[i - W].y) - P->gravity;
                    int toward = fabsf(px) > fabsf(py) ? (px > 0.0f ? i + 1 : i - 1) : (py > 0.0f ? i + W : i - W);
                    if (packed(s, toward)) level = 0.05f;
                }
            }
            if (sol.w > 0.0f && level <= 0.0f) s->region_air |= 1u << (s->region[i] & SP_REGION_MASK);
            /* stiff where the viscosity would exceed the rigid cap (grid_prep.comp:230-237) */
            if (P->caps & (SP_CAP_RIGID | SP_CAP_RHEO)) {
                const float capv = rho * expf(P->ln_rigid);
                if (mu >= capv && level > 0.0f) nn.w = 1.0f;
                mu = fminf_(mu, capv);
                s->rh[i] = nn;
            }
            /* the cell's viscoelastic stress, by volume (grid_prep.comp:241-244) */
            if (P->caps & SP_CAP_ELASTIC) {
                float sc = P->el_g * P->inv_ppc / SP_STRESS_FIX;
                s->estr[i] = (v4){(float)s->acc[k + P->chan_e] * sc, (float)s->acc[k + P->chan_e + 1] * sc,
                                  (float)s->acc[k + P->chan_e + 2] * sc, 0.0f};
            }
            if (P->caps & SP_CAP_MAGNET) s->magc[i].x = chi; /* grid_prep.comp:240 */
            s->cell[i] = (v4){phi_f, rho, mu, level};
            s->comp[i] = (v4){rho, f[1], fam0, phi_t};
            if (any_sorts || (P->caps & SP_CAP_STEAM)) s->gbase[i] = (v4){bg[0], bg[1], bg[2], fv};
            if (multi) {
                s->aux[i] = (v4){nu_c, rho_c, f[2], f[3]};
                s->grho[i] = (v4){rg[0], rg[1], rg[2], rg[3]};
            }
            s->sprop[i] = (v4){sigma, cx_, cy_, (float)major + (float)(top_row + 1) / 128.0f};
            if (th_on) {
                s->therm[i] = (v4){T, Craw, kappa, frozen};
                s->dheat[i] = (v4){P->heat_rate > 0.0f ? emis : 0.0f, 0.0f, 0.0f, 0.0f};
            }
            if (level <= 0.0f) s->pres[i] = (v2){0.0f, 0.0f};
        }
}

/* ---- heat.comp (heat.comp:57-198) -------------------------------------------------------- */

/* y_sat (heat.comp:66-70): mass fraction of water vapour in saturated air at T. */
static float y_sat(float T) {
    T = clampf(T, -40.0f, 99.0f);
    float p = 0.61078f * expf(17.27f * T / (T + 237.3f));
    return 0.622f * p / (101.325f - 0.378f * p);
}

/* centre_vel (heat.comp:73-79) */
static v2 centre_vel(const or_sim* s, int i) {
    int last = s->W * s->H - 1;
    v4 v = s->vel[i], vr = s->vel[i + 1 < last ? i + 1 : last], vt = s->vel[i + s->W < last ? i + s->W : last];
    float ax = face_live(v.z) ? v.x : 0.0f, ay = face_live(v.w) ? v.y : 0.0f;
    float bx = face_live(vr.z) ? vr.x : 0.0f, by = face_live(vt.w) ? vt.y : 0.0f;
    return (v2){0.5f * (ax + bx), 0.5f * (ay + by)};
}

/* The leading material of cell i, + 1 (0: none), as grid_prep left it in sprop.w. */
static int lead_of(const or_sim* s, int i) {
    float w = s->sprop[i].w;
    return (int)((w - floorf(w)) * 128.0f + 0.5f);
}

static void stage_heat(or_sim* s) {
    const sp_gparams* P = &s->P;
    const float PHI_MIN = 0.04f, PR_T = 0.85f;
    int W = s->W, H = s->H;
    float k = s->S.dt * P->heat_rate;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            v4 me = s->therm[i];
            float emis = s->dheat[i].x;
            float phi = s->cell[i].x;
            /* exchange inside the cell (heat.comp:88-106) */
            int32_t hx = s->hcell[2 * (size_t)i], hy = s->hcell[2 * (size_t)i + 1];
            float own_x = 0.0f, own_y = 0.0f;
            if (hx > 0 && me.y > 0.0f && phi > 0.0f) {
                float rate = 5.78f * me.z * 3.14159265f / (P->inv_ppc * me.y * P->inv_ppc / phi);
                /* faster where the cell holds steam and no other liquid (heat.comp:95-104) */
                if ((P->caps & SP_CAP_STEAM) && s->gbase[i].w > 0.0f) {
                    int others = (P->group_mask & ~1u) != 0u && s->comp[i].y + s->aux[i].z + s->aux[i].w > 1e-3f;
                    if (!others) rate = fmaxf_(rate, P->steam2[0]);
                }
                own_x = 1.0f - expf(-k * rate);
                own_y = (float)hy / (float)hx;
            }
            if (phi <= PHI_MIN || me.y <= 0.0f || cx < 1 || cy < 1 || cx >= W - 1 || cy >= H - 1) {
                s->dheat[i] = (v4){0.0f, own_x, own_y, 0.0f};
                continue;
            }
            phi = fminf_(phi, 1.0f);
            int nb[4] = {i - 1, i + 1, i - W, i + W};
            float open[4] = {s->solid[i].x, s->solid[i + 1].x, s->solid[i].y, s->solid[i + W].y};
            float Tk = me.x + 273.15f, Ta = P->heat_x + 273.15f;
            float cond = 0.0f, air = 0.0f, wall = 0.0f, contact = 0.0f;
            float cpc = me.y * P->inv_ppc;
            float hc = P->contact;
            int lead = lead_of(s, i);
            float eddy_l2 = P->eddy * P->eddy / PR_T;
            int eddy = eddy_l2 > 0.0f && me.w < 0.5f;
            v2 u_me = eddy ? centre_vel(s, i) : (v2){0.0f, 0.0f};
            float eddy_max = 0.2f / fmaxf_(k, 1e-9f);
            float per_rate = 1.0f / fmaxf_(P->heat_rate, 1e-6f);
            float evap = 0.0f;
            for (int q = 0; q < 4; q++) {
                int n = nb[q];
                float a = open[q];
                float pn = s->cell[n].x;
                v4 o = s->therm[n];
                if (a > 0.0f) {
                    float pl = (pn > PHI_MIN && o.y > 0.0f) ? fminf_(pn, 1.0f) : 0.0f;
                    if (pl > 0.0f) {
                        float gc = 2.0f * me.z * o.z / fmaxf_(me.z + o.z, 1e-12f);
                        int ln = lead_of(s, n);
                        if (eddy && o.w < 0.5f && ln == lead) {
                            v2 un = centre_vel(s, n);
                            float strain = 1.4142f * len2(un.x - u_me.x, un.y - u_me.y);
                            float alpha = fminf_(eddy_l2 * strain * per_rate, eddy_max);
                            float c1 = cpc / phi, c2 = o.y * P->inv_ppc / pl;
                            gc += alpha * 2.0f * c1 * c2 / fmaxf_(c1 + c2, 1e-12f);
                        }
                        cond += a * fminf_(phi, pl) * gc * (o.x - me.x);
                        if (hc > 0.0f && ln != lead) {
                            float g1 = fmaxf_(2.0f * me.z, hc), g2 = fmaxf_(2.0f * o.z, hc);
                            float gx = fmaxf_(g1 * g2 / (g1 + g2) - gc, 0.0f) * a * fminf_(phi, pl);
                            float cn = o.y * P->inv_ppc;
                            float x = gx * k * (1.0f / cpc + 1.0f / cn);
                            contact += 0.25f * (1.0f - expf(-4.0f * x)) * cpc * cn / (cpc + cn) * (o.x - me.x);
                        }
                    }
                    float exposed = fmaxf_(phi - pl, 0.0f);
                    if (exposed > 0.0f && s->solid[n].w > 0.02f) {
                        float h = P->film + emis * P->radiation * (Ta * Ta + Tk * Tk) * (Ta + Tk);
                        air += a * exposed * h * (P->heat_x - me.x);
                        if (P->evap > 0.0f && lead > 0 && s->mats[lead - 1].group == 0 && s->mats[lead - 1].th2[2] < 150.0f && me.w < 0.5f)
                            evap += a * exposed * P->film * P->evap * fmaxf_(y_sat(me.x) - y_sat(P->heat_x), 0.0f);
                    }
                }
                if (a < 1.0f) {
                    float ts = fmaxf_(s->solid_t[n], s->solid_t[i]);
                    if (ts > 0.5f * SP_NO_TEMP) wall += (1.0f - a) * phi * 2.0f * me.z * (ts - me.x);
                }
            }
            if (evap > 0.0f) air -= evap;
            air *= k;
            wall *= k;
            float lim = cpc * fabsf(P->heat_x - me.x);
            air = clampf(air, -lim, lim);
            float e = k * cond + contact + air + wall;
            s->dheat[i] = (v4){e / cpc, own_x, own_y, 0.0f};
            if (air > 0.0f) ledger_add(s, 0, air);
            else ledger_add(s, 1, -air);
            if (wall > 0.0f) ledger_add(s, 2, wall);
            else ledger_add(s, 3, -wall);
        }
}

/* ---- grid_coef.comp (grid_coef.comp:12-38) -------------------------------------------- */

static void face_coef(v4 a, v4 b, float open, float* g, float* rho) {
    float wsum = a.x + b.x;
    *rho = wsum > 1e-6f ? (a.x * a.y + b.x * b.y) / wsum : 1.0f;
    *g = 0.0f;
    int fa = a.w > 0.0f, fb = b.w > 0.0f;
    if (open <= 0.0f || !(fa || fb)) return;
    float theta = 1.0f;
    if (fa != fb) {
        float lf = fa ? a.w : b.w;
        float la = fa ? b.w : a.w;
        theta = clampf(lf / (lf - la), SP_THETA_MIN, 1.0f);
    }
    *g = 1.0f / (fmaxf_(*rho, 0.05f) * theta);
}

static void stage_coef(or_sim* s) {
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < s->W; cx++) {
            int i = cidx(s, cx, cy);
            v4 o = {0.0f, 0.0f, 1.0f, 1.0f};
            if (cx > 0 && cy > 0) {
                v4 me = s->cell[i];
                v4 sol = s->solid[i];
                face_coef(me, s->cell[i - 1], sol.x, &o.x, &o.z);
                face_coef(me, s->cell[i - s->W], sol.y, &o.y, &o.w);
            }
            s->coef[i] = o;
        }
}

/* ---- surf_smooth / surf_curv / surf_apply --------------------------------------------- */

static void stage_surf_smooth(or_sim* s) {
    static const float B5[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    int more = (s->P.group_mask & 12u) != 0u;
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < s->W; cx++) {
            float sx = 0, sy = 0, s2x = 0, s2y = 0, wsum = 0;
            for (int b = -2; b <= 2; b++)
                for (int a = -2; a <= 2; a++) {
                    int nx = cx + a, ny = cy + b;
                    if (nx < 0 || ny < 0 || nx >= s->W || ny >= s->H) continue;
                    int j = cidx(s, nx, ny);
                    float open = s->solid[j].w;
                    if (open <= 0.02f) continue;
                    float phi = clampf(s->cell[j].x / (SP_REST_PHI * open), 0.0f, 1.0f);
                    float w = B5[a + 2] * B5[b + 2];
                    sx += w * phi;
                    sy += w * (phi * s->comp[j].y);
                    if (more) {
                        float wp = w * phi;
                        s2x += wp * s->aux[j].z;
                        s2y += wp * s->aux[j].w;
                    }
                    wsum += w;
                }
            int i = cidx(s, cx, cy);
            s->surf[i].x = wsum > 0.0f ? sx / wsum : 0.0f;
            s->surf[i].y = wsum > 0.0f ? sy / wsum : 0.0f;
            if (more) {
                s->surf2[i].x = wsum > 0.0f ? s2x / wsum : 0.0f;
                s->surf2[i].y = wsum > 0.0f ? s2y / wsum : 0.0f;
            }
        }
}

static float sfield(const or_sim* s, int x, int y, int k) {
    x = clampi(x, 0, s->W - 1);
    y = clampi(y, 0, s->H - 1);
    int i = cidx(s, x, y);
    if (k < 2) return k == 0 ? s->surf[i].x : s->surf[i].y;
    return k == 2 ? s->surf2[i].x : s->surf2[i].y;
}

static float node_sdf(const or_sim* s, int x, int y) {
    x = clampi(x, 0, s->W);
    y = clampi(y, 0, s->H);
    return s->sdf[y * (s->W + 1) + x];
}

static v2 contact_at(const or_sim* s, int nx, int ny) {
    v4 sp = s->sprop[cidx(s, clampi(nx, 0, s->W - 1), clampi(ny, 0, s->H - 1))];
    float cx = sp.y, cy = sp.z;
    if (cx * cx + cy * cy > 1e-12f) {
        float l = len2(cx, cy);
        return (v2){cx / l, cy / l};
    }
    for (int b = -3; b <= 2; b++)
        for (int a = -3; a <= 2; a++) {
            int j = cidx(s, clampi(nx + a, 0, s->W - 1), clampi(ny + b, 0, s->H - 1));
            float occ = s->cell[j].x;
            cx += s->sprop[j].y * occ;
            cy += s->sprop[j].z * occ;
        }
    if (cx * cx + cy * cy > 1e-12f) {
        float l = len2(cx, cy);
        return (v2){cx / l, cy / l};
    }
    return (v2){0.0f, 1.0f};
}

static v2 normal_node(const or_sim* s, int nx, int ny, int k, float* mag) {
    const float G_MIN = 0.05f;
    float a = sfield(s, nx, ny, k), b = sfield(s, nx - 1, ny, k), d = sfield(s, nx, ny - 1, k), e = sfield(s, nx - 1, ny - 1, k);
    float gx = 0.5f * (a + d - b - e), gy = 0.5f * (a + b - d - e);
    *mag = len2(gx, gy);
    if (*mag < G_MIN) return (v2){0.0f, 0.0f};
    v2 nn = {-gx / *mag, -gy / *mag};
    if (k == 0) {
        float dd = node_sdf(s, nx, ny);
        if (dd < 0.5f) {
            float gwx = node_sdf(s, nx + 1, ny) - node_sdf(s, nx - 1, ny);
            float gwy = node_sdf(s, nx, ny + 1) - node_sdf(s, nx, ny - 1);
            float lw = len2(gwx, gwy);
            if (lw > 1e-4f) {
                float nwx = gwx / lw, nwy = gwy / lw;
                float gdn = gx * nwx + gy * nwy;
                float gtx = gx - gdn * nwx, gty = gy - gdn * nwy;
                float along = len2(gtx, gty);
                if (along > 0.5f * G_MIN) {
                    float tx = -gtx / along, ty = -gty / along;
                    v2 cs = contact_at(s, nx, ny);
                    float ncx = cs.x * nwx + cs.y * tx, ncy = cs.x * nwy + cs.y * ty;
                    float sm = smoothstepf(0.5f * G_MIN, G_MIN, along);
                    float mx = mixf(nn.x, ncx, sm), my = mixf(nn.y, ncy, sm);
                    float lm = len2(mx, my);
                    if (lm > 1e-4f) {
                        nn.x = mx / lm;
                        nn.y = my / lm;
                    } else {
                        nn.x = ncx;
                        nn.y = ncy;
                    }
                }
            }
        }
    }
    return nn;
}

static float curvature(const or_sim* s, int cx, int cy, int k, float kmax) {
    float m00, m10, m01, m11;
    v2 n00 = normal_node(s, cx, cy, k, &m00);
    v2 n10 = normal_node(s, cx + 1, cy, k, &m10);
    v2 n01 = normal_node(s, cx, cy + 1, k, &m01);
    v2 n11 = normal_node(s, cx + 1, cy + 1, k, &m11);
    if (0.25f * (m00 + m10 + m01 + m11) < 0.05f) return 0.0f;
    float kap = 0.5f * (n10.x + n11.x - n00.x - n01.x) + 0.5f * (n01.y + n11.y - n00.y - n10.y);
    return kmax > 0.0f ? clampf(kap, -kmax, kmax) : kap;
}

static void stage_surf_curv(or_sim* s) {
    float kmax = s->P.kappa_max;
    uint32_t gm = s->P.group_mask;
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < s->W; cx++) {
            int i = cidx(s, cx, cy);
            float k0 = curvature(s, cx, cy, 0, kmax);
            float k1 = curvature(s, cx, cy, 1, kmax);
            s->surf[i].z = k0;
            s->surf[i].w = k1;
            if (gm & 12u) {
                s->surf2[i].z = (gm & 4u) ? curvature(s, cx, cy, 2, kmax) : 0.0f;
                s->surf2[i].w = (gm & 8u) ? curvature(s, cx, cy, 3, kmax) : 0.0f;
            }
        }
}

static float air_jump(const or_sim* s, int a, int b) {
    int la = s->cell[a].w > 0.0f;
    int liq = la ? a : b, air = la ? b : a;
    float lf = s->cell[liq].w, lgv = s->cell[air].w;
    float theta = clampf(lf / (lf - lgv), SP_THETA_MIN, 1.0f);
    float kl = s->surf[liq].z, kg = s->surf[air].z;
    float kap = kg != 0.0f ? mixf(kl, kg, theta) : kl;
    return s->S.dt * s->sprop[liq].x * kap;
}

static float share_of(const or_sim* s, int i, int g) {
    if (g == 1) return s->comp[i].y;
    if (g == 2) return s->aux[i].z;
    if (g == 3) return s->aux[i].w;
    return 1.0f - s->comp[i].y - s->aux[i].z - s->aux[i].w;
}

static float curv_of(const or_sim* s, int i, int g) {
    return g == 1 ? s->surf[i].w : g == 2 ? s->surf2[i].z : s->surf2[i].w;
}

static float pair_jump(const or_sim* s, int c, int n) {
    int gc = (int)s->sprop[c].w, gn = (int)s->sprop[n].w;
    if (gc == gn) return 0.0f;
    int hi = gc > gn ? gc : gn, lo = gc < gn ? gc : gn;
    float dc = share_of(s, c, hi) - share_of(s, c, lo), dn = share_of(s, n, hi) - share_of(s, n, lo);
    if (dc * dn >= 0.0f) return 0.0f;
    float theta = -dn / (dc - dn);
    float kc = curv_of(s, c, hi), kn = curv_of(s, n, hi);
    float kap = kc == 0.0f ? kn : kn == 0.0f ? kc : mixf(kn, kc, theta);
    float j = s->S.dt * s->P.pair[hi][lo] * kap;
    return gc == hi ? j : -j;
}

static float face_jump(const or_sim* s, int c, int n, float open) {
    if (open <= 0.0f) return 0.0f;
    int lc = s->cell[c].w > 0.0f, ln = s->cell[n].w > 0.0f;
    if (lc != ln) return air_jump(s, c, n);
    if (lc && ln) return pair_jump(s, c, n);
    return 0.0f;
}

static void stage_surf_apply(or_sim* s) {
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < s->W; cx++) {
            int i = cidx(s, cx, cy);
            v2 pj = {0.0f, 0.0f};
            if (cx > 0 && cy > 0) {
                v4 sol = s->solid[i];
                pj.x = face_jump(s, i, i - 1, sol.x);
                pj.y = face_jump(s, i, i - s->W, sol.y);
            }
            s->pjump[i] = pj;
        }
}

/* ---- explicit viscosity: visc_setup.comp + visc_apply.comp mode 2 ---------------------- */

static int live_u(const or_sim* s, int i) { return face_live(s->vel[i].z); }
static int live_v(const or_sim* s, int i) { return face_live(s->vel[i].w); }
static int closed_u(const or_sim* s, int i) { return face_closed(s->vel[i].z); }
static int closed_v(const or_sim* s, int i) { return face_closed(s->vel[i].w); }
static float vel_u(const or_sim* s, int i) { return live_u(s, i) ? s->vel[i].x : 0.0f; }
static float vel_v(const or_sim* s, int i) { return live_v(s, i) ? s->vel[i].y : 0.0f; }

static v4 strain_of_vel(const or_sim* s, int i, float wx, float wy) {
    int W = s->W;
    v4 o = {0, 0, 0, 0};
    if (wx > 0.0f) {
        o.x = wx * (vel_u(s, i + 1) - vel_u(s, i));
        o.y = wx * (vel_v(s, i + W) - vel_v(s, i));
    }
    if (wy > 0.0f) {
        float ua = vel_u(s, i), ub = vel_u(s, i - W), va = vel_v(s, i), vb = vel_v(s, i - 1);
        if (closed_u(s, i) && live_u(s, i - W)) ua = -ub;
        if (closed_u(s, i - W) && live_u(s, i)) ub = -ua;
        if (closed_v(s, i) && live_v(s, i - 1)) va = -vb;
        if (closed_v(s, i - 1) && live_v(s, i)) vb = -va;
        o.z = wy * 0.5f * ((ua - ub) + (va - vb));
    }
    return o;
}

static void visc_setup(or_sim* s, int mark);

static void stage_visc_explicit(or_sim* s) {
    int W = s->W, H = s->H;
    visc_setup(s, 0);
    /* visc_apply.comp mode 2 (visc_apply.comp:16-48). */
    for (int cy = 1; cy < H - 1; cy++)
        for (int cx = 1; cx < W - 1; cx++) {
            int i = cidx(s, cx, cy);
            float gx = 0.0f, gy = 0.0f;
            if (live_u(s, i)) {
                float c1 = closed_u(s, i - W) ? 1.0f : 0.5f, c2 = closed_u(s, i + W) ? -1.0f : -0.5f;
                gx = s->vs[i - 1].x - s->vs[i].x + s->vs[i].z * c1 + s->vs[i + W].z * c2;
            }
            if (live_v(s, i)) {
                float d1 = closed_v(s, i - 1) ? 1.0f : 0.5f, d2 = closed_v(s, i + 1) ? -1.0f : -0.5f;
                gy = s->vs[i - W].y - s->vs[i].y + s->vs[i].z * d1 + s->vs[i + 1].z * d2;
            }
            v4 cf = s->coef[i];
            if (live_u(s, i)) s->vel[i].x -= gx / cf.z;
            if (live_v(s, i)) s->vel[i].y -= gy / cf.w;
        }
}

/* ---- rhs.comp (rhs.comp:19-93) ------------------------------------------------------- */

static void stage_rhs(or_sim* s) {
    const sp_gparams* P = &s->P;
    int W = s->W, H = s->H;
    int tx = (W + 15) >> 4;
    memset(s->ptile, 0, sizeof(uint32_t) * (size_t)tx * ((H + 15) >> 4));
    float dt = s->S.dt;
    float corr_gain = fminf_(1.0f - expf(-dt / P->tol[3]), P->density_k);
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            v4 me = s->cell[i];
            if (me.w <= 0.0f || cx == 0 || cy == 0 || cx == W - 1 || cy == H - 1) {
                s->mga[i] = (v4){0, 0, 0, 0};
                s->mgd[i] = 0.0f;
                s->prhs[i] = (v4){0, 0, 0, 0};
                continue;
            }
            v4 sol = s->solid[i];
            v4 cf = s->coef[i];
            float aL = sol.x, aB = sol.y, aR = s->solid[i + 1].x, aT = s->solid[i + W].y;
            v4 w = {aL * cf.x, aR * s->coef[i + 1].x, aB * cf.y, aT * s->coef[i + W].y};
            float diag = w.x + w.y + w.z + w.w;
            uint32_t rg = s->region[i];
            int sealed = (s->region_air & (1u << (rg & SP_REGION_MASK))) == 0u;
            if ((rg & SP_REGION_VENT) != 0u && sealed && P->sealed_vent > 0.0f) diag += 1.0f / fmaxf_(me.y, 0.05f);
            v4 vme = s->vel[i];
            float div = aR * s->vel[i + 1].x - aL * vme.x + aT * s->vel[i + W].y - aB * vme.y;
            int lR = s->cell[i + 1].w > 0.0f, lT = s->cell[i + W].w > 0.0f;
            float jump = w.x * s->pjump[i].x + w.y * (lR ? -s->pjump[i + 1].x : s->pjump[i + 1].x)
                + w.z * s->pjump[i].y + w.w * (lT ? -s->pjump[i + W].y : s->pjump[i + W].y);
            float err = clampf(s->comp[i].w - SP_REST_PHI, -0.5f, 0.5f);
            {
                int interior = (aL <= 0.0f || s->cell[i - 1].w > 0.0f) && (aR <= 0.0f || s->cell[i + 1].w > 0.0f)
                    && (aB <= 0.0f || s->cell[i - W].w > 0.0f) && (aT <= 0.0f || s->cell[i + W].w > 0.0f);
                int pressed = s->pres[i].x * (dt / s->S.dt_prev) > 0.5f * me.y * P->gravity * dt;
                if (interior && !pressed) err = fmaxf_(err - P->corr[1], 0.0f);
                else if (!interior && err < 0.0f) err = 0.0f;
            }
            float sv = corr_gain * err;
            if (sealed && P->sealed_vent > 0.0f) sv = 0.0f;
            if (diag > 1e-9f) s->ptile[(cy >> 4) * tx + (cx >> 4)] = 1u;
            s->mga[i] = diag > 1e-9f ? w : (v4){0, 0, 0, 0};
            s->mgd[i] = diag > 1e-9f ? diag : 0.0f;
            s->prhs[i] = (v4){-div + jump, sv / dt, 0.0f, 0.0f};
        }
}

/* ---- pressure: multigrid-preconditioned CG (mod.rs:2929-3010) ------------------------- */

static const int NBX[4] = {-1, 1, 0, 0}, NBY[4] = {0, 0, -1, 1};

void or_levels(int32_t W, int32_t H, sp_levels_info* L) {
    memset(L, 0, sizeof(*L));
    int n = 0;
    L->dims[0][0] = W;
    L->dims[0][1] = H;
    n = 1;
    while (L->dims[n - 1][0] * L->dims[n - 1][1] > SP_COARSEST_CELLS && n < SP_MAX_LEVELS) {
        L->dims[n][0] = (L->dims[n - 1][0] + 1) / 2;
        L->dims[n][1] = (L->dims[n - 1][1] + 1) / 2;
        n++;
    }
    L->n = n;
    int total = 0;
    for (int l = 0; l < n; l++) {
        L->offs[l] = total;
        total += L->dims[l][0] * L->dims[l][1];
    }
    L->total = total;
    int tail = n - 1;
    for (int l = 1; l < n; l++)
        if (L->dims[l][0] * L->dims[l][1] <= SP_TAIL_CELLS) {
            tail = l;
            break;
        }
    if (tail < 1) tail = 1;
    if (tail > n - 1) tail = n - 1;
    L->tail = tail;
    int tiles = 0;
    for (int l = 0; l <= tail; l++) {
        L->tile_offs[l] = tiles;
        tiles += ((L->dims[l][0] + 15) / 16) * ((L->dims[l][1] + 15) / 16);
    }
    L->tiles = tiles;
}

typedef struct { int off, w, h; } lvl;
static lvl level_of(const or_sim* s, int l) {
    lvl r = {s->lv.offs[l], s->lv.dims[l][0], s->lv.dims[l][1]};
    return r;
}
static int inside(lvl L, int x, int y) { return x >= 0 && y >= 0 && x < L.w && y < L.h; }
static int lat(lvl L, int x, int y) { return L.off + y * L.w + x; }
static int unknown(const or_sim* s, lvl L, int x, int y) { return inside(L, x, y) && s->mgd[lat(L, x, y)] > 0.0f; }

/* Galerkin coarse matrix of coarse cell (cx, cy) (mg_build.comp:43-65, mg_tail.comp:44-70). */
static void build_cell(or_sim* s, lvl F, lvl C, int cx, int cy) {
    float D = 0.0f;
    v4 A = {0, 0, 0, 0};
    for (int k = 0; k < 4; k++) {
        int x = 2 * cx + (k & 1), y = 2 * cy + (k >> 1);
        if (!unknown(s, F, x, y)) continue;
        int i = lat(F, x, y);
        D += s->mgd[i];
        v4 w = s->mga[i];
        for (int d = 0; d < 4; d++) {
            int nx = x + NBX[d], ny = y + NBY[d];
            if (!unknown(s, F, nx, ny)) continue;
            float wd = v4c(w, d);
            if ((nx >> 1) == cx && (ny >> 1) == cy) D -= wd;
            else if (d == 0) A.x += wd;
            else if (d == 1) A.y += wd;
            else if (d == 2) A.z += wd;
            else A.w += wd;
        }
    }
    int o = lat(C, cx, cy);
    s->mgd[o] = D > 1e-7f ? D : 0.0f;
    s->mga[o] = D > 1e-7f ? A : (v4){0, 0, 0, 0};
}

static void stage_mg_build(or_sim* s) {
    const sp_levels_info* L = &s->lv;
    for (int l = 1; l <= L->tail; l++) {
        lvl F = level_of(s, l - 1), C = level_of(s, l);
        /* Tile flags: a coarse tile is active when one of its four fine tiles is. */
        int tfx = (F.w + 15) >> 4, tfy = (F.h + 15) >> 4;
        int tcx = (C.w + 15) >> 4, tcy = (C.h + 15) >> 4;
        for (int ty = 0; ty < tcy; ty++)
            for (int tx = 0; tx < tcx; tx++) {
                uint32_t on = 0;
                for (int k = 0; k < 4; k++) {
                    int fx = 2 * tx + (k & 1), fy = 2 * ty + (k >> 1);
                    if (fx < tfx && fy < tfy) on |= s->ptile[L->tile_offs[l - 1] + fy * tfx + fx];
                }
                s->ptile[L->tile_offs[l] + ty * tcx + tx] = on ? 1u : 0u;
            }
        for (int cy = 0; cy < C.h; cy++)
            for (int cx = 0; cx < C.w; cx++) build_cell(s, F, C, cx, cy);
    }
    for (int l = L->tail; l < L->n - 1; l++) {
        lvl F = level_of(s, l), C = level_of(s, l + 1);
        for (int cy = 0; cy < C.h; cy++)
            for (int cx = 0; cx < C.w; cx++) build_cell(s, F, C, cx, cy);
    }
}

/* black cell after the first red half-sweep from zero (mg_down.comp:23-34) */
static v2 down_black(const or_sim* s, lvl F, int x, int y) {
    int i = lat(F, x, y);
    v4 w = s->mga[i];
    v2 sum = s->mgr[i];
    for (int d = 0; d < 4; d++) {
        int mx = x + NBX[d], my = y + NBY[d];
        if (!unknown(s, F, mx, my)) continue;
        int j = lat(F, mx, my);
        float wd = v4c(w, d);
        sum.x += wd * s->mgr[j].x / s->mgd[j];
        sum.y += wd * s->mgr[j].y / s->mgd[j];
    }
    float D = s->mgd[i];
    return (v2){sum.x / D, sum.y / D};
}

static v2 down_z(const or_sim* s, lvl F, int x, int y) {
    if (!unknown(s, F, x, y)) return (v2){0, 0};
    int i = lat(F, x, y);
    if (((x + y) & 1) == 0) return (v2){s->mgr[i].x / s->mgd[i], s->mgr[i].y / s->mgd[i]};
    return down_black(s, F, x, y);
}

/* mg_down.comp:36-69 */
static void mg_down(or_sim* s, int l) {
    lvl F = level_of(s, l), C = level_of(s, l + 1);
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) s->mgx[lat(F, x, y)] = down_z(s, F, x, y);
    for (int cy = 0; cy < C.h; cy++)
        for (int cx = 0; cx < C.w; cx++) {
            v2 sum = {0, 0};
            for (int k = 0; k < 2; k++) {
                int qx = 2 * cx + k, qy = 2 * cy + k;
                if (!unknown(s, F, qx, qy)) continue;
                v4 w = s->mga[lat(F, qx, qy)];
                for (int d = 0; d < 4; d++) {
                    int nx = qx + NBX[d], ny = qy + NBY[d];
                    if (!unknown(s, F, nx, ny)) continue;
                    v2 zn = s->mgx[lat(F, nx, ny)];
                    float wd = v4c(w, d);
                    sum.x += wd * zn.x;
                    sum.y += wd * zn.y;
                }
            }
            s->mgr[lat(C, cx, cy)] = sum;
        }
}

/* black cell of the way up (mg_up.comp:26-36) */
static v2 up_black(const or_sim* s, lvl F, lvl C, int x, int y) {
    int i = lat(F, x, y);
    v4 w = s->mga[i];
    v2 sum = s->mgr[i];
    for (int d = 0; d < 4; d++) {
        int mx = x + NBX[d], my = y + NBY[d];
        if (!unknown(s, F, mx, my)) continue;
        float wd = v4c(w, d);
        v2 xm = s->mgx[lat(F, mx, my)];
        v2 zc = s->mgz[lat(C, mx >> 1, my >> 1)];
        sum.x += wd * (xm.x + SP_MG_OMEGA * zc.x);
        sum.y += wd * (xm.y + SP_MG_OMEGA * zc.y);
    }
    float D = s->mgd[i];
    return (v2){sum.x / D, sum.y / D};
}

/* mg_up.comp:38-75; on level 0 returns the r.z and r.r sums (FP64). */
static void mg_up(or_sim* s, int l, double sums[4]) {
    lvl F = level_of(s, l), C = level_of(s, l + 1);
    /* the black cells first (their values only feed red neighbours and the output) */
    v2* nb = (v2*)calloc((size_t)F.w * F.h, sizeof(v2));
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++)
            if (unknown(s, F, x, y) && ((x + y) & 1)) nb[y * F.w + x] = up_black(s, F, C, x, y);
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            int i = lat(F, x, y);
            v2 z = {0, 0};
            if (unknown(s, F, x, y)) {
                v2 r = s->mgr[i];
                if (((x + y) & 1) == 0) {
                    v4 w = s->mga[i];
                    v2 sum = r;
                    for (int d = 0; d < 4; d++) {
                        int nx = x + NBX[d], ny = y + NBY[d];
                        if (!unknown(s, F, nx, ny)) continue;
                        float wd = v4c(w, d);
                        v2 zn = nb[ny * F.w + nx];
                        sum.x += wd * zn.x;
                        sum.y += wd * zn.y;
                    }
                    float D = s->mgd[i];
                    z.x = sum.x / D;
                    z.y = sum.y / D;
                } else {
                    z = nb[y * F.w + x];
                }
                if (sums) {
                    sums[0] += (double)r.x * (double)z.x;
                    sums[1] += (double)r.y * (double)z.y;
                    sums[2] += (double)r.x * (double)r.x;
                    sums[3] += (double)r.y * (double)r.y;
                }
            }
            s->mgz[i] = z;
        }
    free(nb);
}

/* mg_tail.comp mode 0 (mg_tail.comp:159-225): levels first..last in shared arrays. */
static void mg_tail(or_sim* s) {
    const sp_levels_info* L = &s->lv;
    int first = L->tail, last = L->n - 1;
    int BASE = L->offs[first];
    int total = L->offs[last] + L->dims[last][0] * L->dims[last][1] - BASE;
    int NF = L->dims[first][0] * L->dims[first][1];
    v2* sx = (v2*)calloc((size_t)total, sizeof(v2));
    float* sd = (float*)calloc((size_t)total, sizeof(float));
    v2* sr = (v2*)calloc((size_t)(total > NF ? total - NF : 1), sizeof(v2));
    for (int t = 0; t < total; t++) sd[t] = s->mgd[BASE + t];
#define RHS_AT(sl) ((sl) < NF ? s->mgr[BASE + (sl)] : sr[(sl) - NF])
#define SAT(Lv, x, y) ((Lv).off - BASE + (y) * (Lv).w + (x))
#define SUNK(Lv, x, y) (inside(Lv, x, y) && sd[SAT(Lv, x, y)] > 0.0f)
    /* one half-sweep of one colour; colours are independent within a half-sweep */
#define SWEEP(Lv, colour, zero)                                                               \
    do {                                                                                      \
        for (int t = 0; t < (Lv).w * (Lv).h; t++) {                                           \
            int x = t % (Lv).w, y = t / (Lv).w;                                               \
            int sl = (Lv).off - BASE + t;                                                     \
            float D = sd[sl];                                                                 \
            if (D <= 0.0f) {                                                                  \
                if (zero) sx[sl] = (v2){0, 0};                                                \
                continue;                                                                     \
            }                                                                                 \
            if (((x + y) & 1) != (colour)) {                                                  \
                if (zero) sx[sl] = (v2){0, 0};                                                \
                continue;                                                                     \
            }                                                                                 \
            v2 sum = RHS_AT(sl);                                                              \
            if (!(zero)) {                                                                    \
                v4 w = s->mga[BASE + sl];                                                     \
                for (int d = 0; d < 4; d++) {                                                 \
                    int nx = x + NBX[d], ny = y + NBY[d];                                     \
                    if (SUNK(Lv, nx, ny)) {                                                   \
                        float wd = v4c(w, d);                                                 \
                        v2 xn = sx[SAT(Lv, nx, ny)];                                          \
                        sum.x += wd * xn.x;                                                   \
                        sum.y += wd * xn.y;                                                   \
                    }                                                                         \
                }                                                                             \
            }                                                                                 \
            sx[sl] = (v2){sum.x / D, sum.y / D};                                              \
        }                                                                                     \
    } while (0)
    for (int l = first; l < last; l++) {
        lvl F = level_of(s, l), C = level_of(s, l + 1);
        SWEEP(F, 0, 1);
        SWEEP(F, 1, 0);
        for (int t = 0; t < C.w * C.h; t++) {
            int cx = t % C.w, cy = t / C.w;
            v2 sum = {0, 0};
            for (int k = 0; k < 4; k++) {
                int x = 2 * cx + (k & 1), y = 2 * cy + (k >> 1);
                if (!SUNK(F, x, y)) continue;
                int sl = SAT(F, x, y);
                v4 w = s->mga[BASE + sl];
                v2 az = {sd[sl] * sx[sl].x, sd[sl] * sx[sl].y};
                for (int d = 0; d < 4; d++) {
                    int nx = x + NBX[d], ny = y + NBY[d];
                    if (SUNK(F, nx, ny)) {
                        float wd = v4c(w, d);
                        v2 xn = sx[SAT(F, nx, ny)];
                        az.x -= wd * xn.x;
                        az.y -= wd * xn.y;
                    }
                }
                v2 r = RHS_AT(sl);
                sum.x += r.x - az.x;
                sum.y += r.y - az.y;
            }
            sr[C.off - BASE + t - NF] = sum;
        }
    }
    {
        lvl E = level_of(s, last);
        SWEEP(E, 0, 1);
        for (int h = 1; h < SP_TAIL_HALF_SWEEPS; h++) SWEEP(E, h & 1, 0);
    }
    for (int l = last - 1; l >= first; l--) {
        lvl F = level_of(s, l), C = level_of(s, l + 1);
        /* coarse-grid correction fused with the black half-sweep (mg_tail.comp:141-157) */
        for (int t = 0; t < F.w * F.h; t++) {
            int x = t % F.w, y = t / F.w;
            if (((x + y) & 1) == 0) continue;
            int sl = SAT(F, x, y);
            float D = sd[sl];
            if (D <= 0.0f) continue;
            v4 w = s->mga[BASE + sl];
            v2 sum = RHS_AT(sl);
            for (int d = 0; d < 4; d++) {
                int nx = x + NBX[d], ny = y + NBY[d];
                if (!SUNK(F, nx, ny)) continue;
                float wd = v4c(w, d);
                v2 xn = sx[SAT(F, nx, ny)];
                v2 zc = sx[SAT(C, nx >> 1, ny >> 1)];
                sum.x += wd * (xn.x + SP_MG_OMEGA * zc.x);
                sum.y += wd * (xn.y + SP_MG_OMEGA * zc.y);
            }
            sx[sl] = (v2){sum.x / D, sum.y / D};
        }
        SWEEP(F, 0, 0);
    }
    for (int t = 0; t < NF; t++) s->mgz[BASE + t] = sx[t];
#undef RHS_AT
#undef SAT
#undef SUNK
#undef SWEEP
    free(sx);
    free(sd);
    free(sr);
}

static void vcycle(or_sim* s, double sums[4]) {
    for (int l = 0; l < s->lv.tail; l++) mg_down(s, l);
    mg_tail(s);
    for (int l = s->lv.tail - 1; l >= 0; l--) mg_up(s, l, l == 0 ? sums : NULL);
}

/* cg_scale.comp:14-27. Mode 1 (a second solve follows the viscosity, with its own warm
 * start): the first solve starts from last substep's P1 = P - P2. */
static void cg_scale(or_sim* s) {
    lvl F = level_of(s, 0);
    int tx = (F.w + 15) >> 4;
    float ratio = s->S.dt / s->S.dt_prev;
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            if (s->ptile[(y >> 4) * tx + (x >> 4)] == 0u) continue;
            int i = lat(F, x, y);
            if (s->mgd[i] > 0.0f) {
                float p = s->pres[i].x;
                if (s->P.v_p2warm) p -= s->p2keep[i];
                else s->p2keep[i] = 0.0f;
                s->pres[i].x = p * ratio;
            } else {
                s->pres[i] = (v2){0, 0};
            }
        }
}

/* cg_init.comp:11-37: r = b - A x, and b.b */
static void cg_init(or_sim* s, double bb[2]) {
    lvl F = level_of(s, 0);
    bb[0] = bb[1] = 0.0;
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            int i = lat(F, x, y);
            float D = s->mgd[i];
            if (D > 0.0f) {
                v2 xv = s->pres[i];
                v4 w = s->mga[i];
                v2 ax = {D * xv.x, D * xv.y};
                for (int d = 0; d < 4; d++) {
                    int nx = x + NBX[d], ny = y + NBY[d];
                    if (unknown(s, F, nx, ny)) {
                        float wd = v4c(w, d);
                        v2 pn = s->pres[lat(F, nx, ny)];
                        ax.x -= wd * pn.x;
                        ax.y -= wd * pn.y;
                    }
                }
                v4 b = s->prhs[i];
                s->mgr[i] = (v2){b.x - ax.x, b.y - ax.y};
                bb[0] += (double)b.x * (double)b.x;
                bb[1] += (double)b.y * (double)b.y;
            } else {
                s->mgr[i] = (v2){0, 0};
            }
        }
}

static v2 direction(const or_sim* s, int i, float bx, float by) {
    v2 z = s->mgz[i];
    if (bx == 0.0f && by == 0.0f) return z;
    v2 po = s->cgp[i];
    return (v2){bx != 0.0f ? z.x + bx * po.x : z.x, by != 0.0f ? z.y + by * po.y : z.y};
}

/* The CG iterations of a pressure solve from pres (record_cg, mod.rs:3012-3031). second:
 * the solve after the viscosity (cg_reduce modes 8 and 9: the density correction counts as
 * solved). */
static void pressure_iterate(or_sim* s, int second) {
    const sp_gparams* P = &s->P;
    sp_solve* K = &s->K;
    lvl F = level_of(s, 0);
    int N = F.w * F.h;
    double bb[2];
    cg_init(s, bb);
    double a[4] = {0, 0, 0, 0};
    vcycle(s, a);
    /* cg_reduce mode 0 (reduce_tail.glsl:61-78) */
    double tol2x = (double)P->tol[0] * (double)P->tol[0], tol2y = (double)P->tol[1] * (double)P->tol[1];
    K->bb[0] = bb[0];
    K->bb[1] = bb[1];
    K->rz[0] = a[0];
    K->rz[1] = a[1];
    K->rr[0] = a[2];
    K->rr[1] = a[3];
    K->beta[0] = K->beta[1] = 0.0;
    K->alpha[0] = K->alpha[1] = 0.0;
    K->iters[0] = K->iters[1] = 0;
    K->done[0] = a[2] <= tol2x * bb[0];
    K->done[1] = second ? 1u : (a[3] <= tol2y * bb[1]);
    K->solves += 1;
    v2* np = (v2*)malloc(sizeof(v2) * (size_t)N);
    uint32_t k = 0;
    for (; k < (uint32_t)P->p_max_iters && !(K->done[0] && K->done[1]); k++) {
        /* cg_apply.comp:15-51, alpha (reduce_tail.glsl:79-87) */
        float bx = (float)K->beta[0], by = (float)K->beta[1];
        double pq[2] = {0, 0};
        for (int y = 0; y < F.h; y++)
            for (int x = 0; x < F.w; x++) {
                int i = lat(F, x, y);
                float D = s->mgd[i];
                if (D > 0.0f) {
                    v2 p = direction(s, i, bx, by);
                    v4 w = s->mga[i];
                    v2 q = {D * p.x, D * p.y};
                    for (int d = 0; d < 4; d++) {
                        int nx = x + NBX[d], ny = y + NBY[d];
                        if (unknown(s, F, nx, ny)) {
                            float wd = v4c(w, d);
                            v2 pn = direction(s, lat(F, nx, ny), bx, by);
                            q.x -= wd * pn.x;
                            q.y -= wd * pn.y;
                        }
                    }
                    np[i] = p;
                    s->cgq[i] = q;
                    pq[0] += (double)p.x * (double)q.x;
                    pq[1] += (double)p.y * (double)q.y;
                } else {
                    np[i] = (v2){0, 0};
                }
            }
        memcpy(s->cgp, np, sizeof(v2) * (size_t)N);
        K->alpha[0] = (K->done[0] == 0 && pq[0] > 0.0) ? K->rz[0] / pq[0] : 0.0;
        K->alpha[1] = (K->done[1] == 0 && pq[1] > 0.0) ? K->rz[1] / pq[1] : 0.0;
        /* cg_update.comp:11-20 */
        float ax = (float)K->alpha[0], ay = (float)K->alpha[1];
        for (int i = 0; i < N; i++) {
            if (s->mgd[i] <= 0.0f) continue;
            s->pres[i].x += ax * s->cgp[i].x;
            s->pres[i].y += ay * s->cgp[i].y;
            s->mgr[i].x -= ax * s->cgq[i].x;
            s->mgr[i].y -= ay * s->cgq[i].y;
        }
        /* V-cycle, then cg_reduce mode 2 (reduce_tail.glsl:88-104) */
        double r[4] = {0, 0, 0, 0};
        vcycle(s, r);
        uint32_t was0 = K->done[0], was1 = K->done[1];
        K->iters[0] += 1u - was0;
        K->iters[1] += 1u - was1;
        uint32_t now0 = r[2] <= tol2x * K->bb[0], now1 = r[3] <= tol2y * K->bb[1];
        uint32_t d0 = was0 | now0, d1 = was1 | now1;
        K->beta[0] = (d0 == 0 && K->rz[0] > 0.0) ? r[0] / K->rz[0] : 0.0;
        K->beta[1] = (d1 == 0 && K->rz[1] > 0.0) ? r[1] / K->rz[1] : 0.0;
        K->rz[0] = r[0];
        K->rz[1] = r[1];
        if (!was0) K->rr[0] = r[2];
        if (!was1) K->rr[1] = r[3];
        K->done[0] = d0;
        K->done[1] = d1;
    }
    free(np);
    /* cg_reduce mode 3 / 9 (reduce_tail.glsl:105-116) */
    uint32_t it = K->iters[0] > K->iters[1] ? K->iters[0] : K->iters[1];
    if (second) {
        if (K->iters[0] > K->p2_it_max) K->p2_it_max = K->iters[0];
        if (K->done[0] == 0) K->p2_unconv += 1;
        K->p2_iters_last = K->iters[0];
    } else {
        if (it > K->it_max) K->it_max = it;
        if (K->done[0] == 0) K->unconv += 1;
        K->iters_last = k;
        K->iters_sum += k;
    }
    float rel = K->bb[0] > 0.0 ? (float)sqrt(K->rr[0] / K->bb[0]) : 0.0f;
    if (rel > K->res_max) K->res_max = rel;
}

/* record_pressure (mod.rs:2956-2990), the first solve. */
static void stage_pressure(or_sim* s) {
    stage_mg_build(s);
    cg_scale(s);
    pressure_iterate(s, 0);
}

/* ---- project.comp (project.comp:26-58) ------------------------------------------------ */

static float difference(int lc, int ln, float pc, float pn, float j) {
    if (lc && ln) return pc - pn - j;
    return (lc ? pc : j) - (ln ? pn : j);
}

static void stage_project(or_sim* s) {
    float m = u2f(s->S.umax_grid);
    for (int cy = 1; cy < s->H; cy++)
        for (int cx = 1; cx < s->W; cx++) {
            int i = cidx(s, cx, cy);
            v4 cf = s->coef[i];
            v4 me = s->vel[i];
            if (s->mgd[i] > 0.0f) s->prhs[i].z = s->pres[i].x;
            if (cf.x != 0.0f || cf.y != 0.0f) {
                v2 p = s->pres[i];
                int liq = s->cell[i].w > 0.0f;
                if (face_live(me.z) && cf.x != 0.0f) {
                    int n = i - 1;
                    me.x -= cf.x * difference(liq, s->cell[n].w > 0.0f, p.x, s->pres[n].x, s->pjump[i].x);
                    me.z = -cf.x * (p.y - s->pres[n].y);
                }
                if (face_live(me.w) && cf.y != 0.0f) {
                    int n = i - s->W;
                    me.y -= cf.y * difference(liq, s->cell[n].w > 0.0f, p.x, s->pres[n].x, s->pjump[i].y);
                    me.w = -cf.y * (p.y - s->pres[n].y);
                }
                s->vel[i] = me;
            }
            float mm = 0.0f;
            if (face_live(me.z) && cf.x != 0.0f) mm = fabsf(me.x);
            if (face_live(me.w) && cf.y != 0.0f) mm = fmaxf_(mm, fabsf(me.y));
            if (mm > m) m = mm;
        }
    s->S.umax_grid = f2u(m);
}

/* ---- implicit viscosity (visc.glsl, visc_setup/split/diag/strain/apply/iter/update.comp,
 *      vmg.glsl, vmg_build/down/up.comp, visc_mgu/mgz.comp, reduce_tail.glsl modes 4-7,
 *      10-13) ------------------------------------------------------------------------------
 *
 * The reference runs the viscosity multigrid tile by tile, each 16 x 16 tile smoothing in
 * shared memory a region of the tile plus 2 NU (+1) cells, every half-sweep on a ring one
 * cell narrower than the last (vmg.glsl:113-134). That is the same arithmetic as sweeping the
 * whole level (the cells each ring needs were updated exactly by the half-sweep before), so
 * the restatement sweeps whole levels, reading zero outside the level, and writes back only
 * the cells of active tiles, as the tiles do. */

static int vtile_of(const or_sim* s, int x, int y) { return (y >> 4) * ((s->W + 15) >> 4) + (x >> 4); }
static int v_on(const or_sim* s, int x, int y) { return s->vtiles[vtile_of(s, x, y)] != 0u; }
static int v_interior(const or_sim* s, int x, int y) { return x >= 1 && y >= 1 && x < s->W - 1 && y < s->H - 1; }
/* GLSL max(x, y): y if x < y, otherwise x. */
static float gmax(float x, float y) { return x < y ? y : x; }

/* Per-cell FP64 dot products of two float2 (the build's precision for the solve scalars). */
static double dot2(v2 a, v2 b) { return (double)a.x * (double)b.x + (double)a.y * (double)b.y; }

/* visc_setup.comp:14-58 with the tile marks (mark != 0): weights, strain of the velocity. */
static void visc_setup(or_sim* s, int mark) {
    int W = s->W, H = s->H;
    float dt = s->S.dt;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            float wc = 0.0f, wn = 0.0f;
            if (v_interior(s, cx, cy)) {
                v4 me = s->cell[i];
#define OK_U(j) (live_u(s, j) || closed_u(s, j))
#define OK_V(j) (live_v(s, j) || closed_v(s, j))
                if (me.w > 0.0f && OK_U(i) && OK_U(i + 1) && OK_V(i) && OK_V(i + W)) wc = 2.0f * dt * me.z;
                int cs4[4] = {i - W - 1, i - W, i - 1, i};
                int nliq = 0, any_air = 0;
                float mu = 0.0f;
                for (int k = 0; k < 4; k++) {
                    v4 q = s->cell[cs4[k]];
                    if (q.w > 0.0f) {
                        nliq++;
                        mu += q.z;
                    } else if (s->solid[cs4[k]].w > 0.02f) {
                        any_air = 1;
                    }
                }
                if (nliq > 0 && !any_air && OK_U(i) && OK_U(i - W) && OK_V(i) && OK_V(i - 1))
                    wn = 4.0f * dt * (mu / (float)nliq) * ((float)nliq / 4.0f);
#undef OK_U
#undef OK_V
            }
            s->vw[i].x = wc;
            s->vw[i].y = wn;
        }
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            s->vs[i] = strain_of_vel(s, i, s->vw[i].x, s->vw[i].y);
            if (!mark) continue;
            v4 me = s->cell[i];
            if (me.w > 0.0f && dt * me.z / fmaxf_(me.y, 0.05f) >= s->P.corr[3]) {
                for (int b = -1; b <= 1; b++)
                    for (int a = -1; a <= 1; a++)
                        s->vtiles[vtile_of(s, clampi(cx + a, 0, W - 1), clampi(cy + b, 0, H - 1))] = 1u;
            }
        }
}

/* G^T of a weighted strain field st at the faces of cell i (visc_apply.comp:16-28). */
static v2 gts_of(const or_sim* s, const v4* st, int i) {
    int W = s->W;
    v2 g = {0.0f, 0.0f};
    if (live_u(s, i)) {
        float c1 = closed_u(s, i - W) ? 1.0f : 0.5f, c2 = closed_u(s, i + W) ? -1.0f : -0.5f;
        g.x = st[i - 1].x - st[i].x + st[i].z * c1 + st[i + W].z * c2;
    }
    if (live_v(s, i)) {
        float d1 = closed_v(s, i - 1) ? 1.0f : 0.5f, d2 = closed_v(s, i + 1) ? -1.0f : -0.5f;
        g.y = st[i - W].y - st[i].y + st[i].z * d1 + st[i + 1].z * d2;
    }
    return g;
}

/* visc_split.comp:24-38: the explicit step in the inactive tiles. */
static void visc_split(or_sim* s) {
    for (int cy = 1; cy < s->H - 1; cy++)
        for (int cx = 1; cx < s->W - 1; cx++) {
            if (v_on(s, cx, cy)) continue;
            int i = cidx(s, cx, cy);
            v2 g = gts_of(s, s->vs, i);
            v4 cf = s->coef[i];
            if (live_u(s, i)) s->vel[i].x -= g.x / cf.z;
            if (live_v(s, i)) s->vel[i].y -= g.y / cf.w;
        }
}

/* Strain weights as the implicit system sees them (visc_diag.comp:75-81). */
static v2 v_wts(const or_sim* s, int x, int y) {
    if (!v_on(s, x, y)) return (v2){0.0f, 0.0f};
    v4 w = s->vw[cidx(s, x, y)];
    return (v2){w.x, w.y};
}
static int unk_u(const or_sim* s, int x, int y) { return v_interior(s, x, y) && v_on(s, x, y) && live_u(s, cidx(s, x, y)); }
static int unk_v(const or_sim* s, int x, int y) { return v_interior(s, x, y) && v_on(s, x, y) && live_v(s, cidx(s, x, y)); }

/* The couplings of A at cell (x, y) (visc_diag.comp:83-117). */
static void couplings(const or_sim* s, int x, int y, v4* a, v4* b) {
    int W = s->W, i = cidx(s, x, y);
    int u = unk_u(s, x, y), v = unk_v(s, x, y);
    *a = (v4){0, 0, 0, 0};
    *b = (v4){0, 0, 0, 0};
    if (u && unk_u(s, x - 1, y)) a->x = -v_wts(s, x - 1, y).x;
    if (v && unk_v(s, x, y - 1)) a->w = -v_wts(s, x, y - 1).x;
    float wn = v_wts(s, x, y).y;
    if (wn > 0.0f) {
        float ua = u ? (closed_u(s, i - W) ? 1.0f : 0.5f) : 0.0f;
        float ub = unk_u(s, x, y - 1) ? -(closed_u(s, i) ? 1.0f : 0.5f) : 0.0f;
        float va = v ? (closed_v(s, i - 1) ? 1.0f : 0.5f) : 0.0f;
        float vb = unk_v(s, x - 1, y) ? -(closed_v(s, i) ? 1.0f : 0.5f) : 0.0f;
        a->y = wn * ua * ub;
        a->z = wn * va * vb;
        b->x = wn * ua * va;
        b->y = wn * ua * vb;
    }
    float wu = v_wts(s, x, y + 1).y;
    if (wu > 0.0f && u) {
        float ub = -(closed_u(s, i + W) ? 1.0f : 0.5f);
        float va = unk_v(s, x, y + 1) ? (closed_v(s, i + W - 1) ? 1.0f : 0.5f) : 0.0f;
        float vb = unk_v(s, x - 1, y + 1) ? -(closed_v(s, i + W) ? 1.0f : 0.5f) : 0.0f;
        b->z = wu * ub * va;
        b->w = wu * ub * vb;
    }
}

/* visc_diag.comp:121-184: the Jacobi diagonal, the warm start, the clearing of inactive
 * tiles and (multigrid) the matrix of level 0. */
static void visc_diag(or_sim* s) {
    int W = s->W, H = s->H;
    int mg = s->P.v_precond != 0, warm = s->P.v_warm != 0;
    int ntx = (W + 15) >> 4, nt = ntx * ((H + 15) >> 4);
    uint8_t* was = (uint8_t*)calloc((size_t)nt, 1);
    if (mg)
        for (int t = 0; t < nt; t++) {
            was[t] = s->vtile[t] != 0u;
            s->vtile[t] = s->vtiles[t];
        }
    float ratio = s->S.dt / s->S.dt_prev;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            int on = v_on(s, cx, cy);
            v2 d = {1.0f, 1.0f};
            int lu = live_u(s, i), lv = live_v(s, i);
            if (v_interior(s, cx, cy)) {
                v4 cf = s->coef[i];
                v2 me = v_wts(s, cx, cy);
                if (lu) {
                    float c1 = closed_u(s, i - W) ? 1.0f : 0.5f, c2 = closed_u(s, i + W) ? 1.0f : 0.5f;
                    d.x = cf.z + v_wts(s, cx - 1, cy).x + me.x + me.y * c1 * c1 + v_wts(s, cx, cy + 1).y * c2 * c2;
                }
                if (lv) {
                    float d1 = closed_v(s, i - 1) ? 1.0f : 0.5f, d2 = closed_v(s, i + 1) ? 1.0f : 0.5f;
                    d.y = cf.w + v_wts(s, cx, cy - 1).x + me.x + me.y * d1 * d1 + v_wts(s, cx + 1, cy).y * d2 * d2;
                }
            }
            v4 o = s->vw[i];
            o.z = d.x;
            o.w = d.y;
            v2 dv = {0.0f, 0.0f};
            if (warm && on && v_interior(s, cx, cy)) {
                dv.x = s->vd[i].x * ratio * (lu ? 1.0f : 0.0f);
                dv.y = s->vd[i].y * ratio * (lv ? 1.0f : 0.0f);
            }
            s->vd[i] = dv;
            if (!on) {
                o.x = o.y = 0.0f;
                s->vs[i] = (v4){0, 0, 0, 0};
                s->vr[i] = s->vp[i] = s->vp2[i] = s->vq[i] = (v2){0, 0};
            }
            if (mg && (on || was[vtile_of(s, cx, cy)])) {
                v4 a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
                v2 dm = {0.0f, 0.0f};
                if (on && v_interior(s, cx, cy)) {
                    couplings(s, cx, cy, &a, &b);
                    dm.x = d.x * (lu ? 1.0f : 0.0f);
                    dm.y = d.y * (lv ? 1.0f : 0.0f);
                }
                if (!on) s->vmz[i] = s->vaz[i] = s->vr2[i] = s->vq2[i] = (v2){0, 0};
                s->vma[i] = a;
                s->vmb[i] = b;
                s->vmd[i] = dm;
            }
            /* written after the neighbours' weights are read: o.xy of an active tile does
             * not change and an inactive tile's weights read as zero either way */
            s->vw[i] = o;
        }
    free(was);
}

static int vflag(const or_sim* s, int l, int slot) {
    return l == 0 ? s->vtiles[slot] != 0u : s->vtile[s->lv.tile_offs[l] + slot] != 0u;
}

/* vmg_build.comp:187-240: level l of the viscosity multigrid from level l - 1. A tile that
 * is not active reads as zero (the reference clears it once when it goes inactive). */
static void vmg_build(or_sim* s, int l) {
    lvl F = level_of(s, l - 1), C = level_of(s, l);
    static const int UVX[4] = {0, -1, 0, -1}, UVY[4] = {0, 0, 1, 1};
    int tfx = (F.w + 15) >> 4, tfy = (F.h + 15) >> 4;
    int tcx = (C.w + 15) >> 4, tcy = (C.h + 15) >> 4;
    for (int ty = 0; ty < tcy; ty++)
        for (int tx = 0; tx < tcx; tx++) {
            uint32_t on = 0;
            for (int k = 0; k < 4; k++) {
                int fx = 2 * tx + (k & 1), fy = 2 * ty + (k >> 1);
                if (fx < tfx && fy < tfy) on |= (uint32_t)vflag(s, l - 1, fy * tfx + fx);
            }
            s->vtile[s->lv.tile_offs[l] + ty * tcx + tx] = on ? 1u : 0u;
        }
    for (int cy = 0; cy < C.h; cy++)
        for (int cx = 0; cx < C.w; cx++) {
            int o = lat(C, cx, cy);
            if (!vflag(s, l, (cy >> 4) * tcx + (cx >> 4))) {
                s->vma[o] = s->vmb[o] = (v4){0, 0, 0, 0};
                s->vmd[o] = (v2){0, 0};
                continue;
            }
            v2 D = {0.0f, 0.0f};
            v4 A = {0, 0, 0, 0}, B = {0, 0, 0, 0};
            for (int k = 0; k < 4; k++) {
                int ax = k & 1, ay = k >> 1;
                int x = 2 * cx + ax, y = 2 * cy + ay;
                if (!inside(F, x, y)) continue;
                int i = lat(F, x, y);
                v4 a = s->vma[i], b = s->vmb[i];
                D.x += s->vmd[i].x;
                D.y += s->vmd[i].y;
                if (ax == 1) {
                    D.x += 2.0f * a.x;
                    D.y += 2.0f * a.z;
                } else {
                    A.x += a.x;
                    A.z += a.z;
                }
                if (ay == 1) {
                    D.x += 2.0f * a.y;
                    D.y += 2.0f * a.w;
                } else {
                    A.y += a.y;
                    A.w += a.w;
                }
                for (int m = 0; m < 4; m++) {
                    int rx = ((x + UVX[m]) >> 1) - cx, ry = ((y + UVY[m]) >> 1) - cy;
                    int sl = (rx == 0 ? 0 : 1) + (ry == 0 ? 0 : 2);
                    float bm = v4c(b, m);
                    if (sl == 0) B.x += bm;
                    else if (sl == 1) B.y += bm;
                    else if (sl == 2) B.z += bm;
                    else B.w += bm;
                }
            }
            s->vma[o] = A;
            s->vmb[o] = B;
            s->vmd[o] = (v2){gmax(D.x, 0.0f), gmax(D.y, 0.0f)};
        }
}

/* One level of the viscosity multigrid in whole-level arrays (vmg.glsl:64-134). */
typedef struct {
    lvl L;
    v4 *a, *b;
    v2 *d, *r, *x;
} vlev;

static v4 va_(const vlev* V, int x, int y) { return inside(V->L, x, y) ? V->a[lat(V->L, x, y) - V->L.off] : (v4){0, 0, 0, 0}; }
static v4 vb_(const vlev* V, int x, int y) { return inside(V->L, x, y) ? V->b[lat(V->L, x, y) - V->L.off] : (v4){0, 0, 0, 0}; }
static v2 vx_(const vlev* V, const v2* xs, int x, int y) { return inside(V->L, x, y) ? xs[lat(V->L, x, y) - V->L.off] : (v2){0, 0}; }

/* Off-diagonal part of A x at (x, y) (vmg.glsl:69-78). */
static v2 v_offd(const vlev* V, const v2* xs, int x, int y) {
    v4 a = va_(V, x, y), b = vb_(V, x, y);
    v2 xl = vx_(V, xs, x - 1, y), xr = vx_(V, xs, x + 1, y), xd = vx_(V, xs, x, y - 1), xu = vx_(V, xs, x, y + 1);
    v2 xc = vx_(V, xs, x, y), xlu = vx_(V, xs, x - 1, y + 1), xrd = vx_(V, xs, x + 1, y - 1);
    v4 ar = va_(V, x + 1, y), au = va_(V, x, y + 1), br = vb_(V, x + 1, y), bd = vb_(V, x, y - 1), brd = vb_(V, x + 1, y - 1);
    float u = a.x * xl.x + ar.x * xr.x + a.y * xd.x + au.y * xu.x + b.x * xc.y + b.y * xl.y + b.z * xu.y + b.w * xlu.y;
    float v = a.z * xl.y + ar.z * xr.y + a.w * xd.y + au.w * xu.y + b.x * xc.x + br.y * xr.x + bd.z * xd.x + brd.w * xrd.x;
    return (v2){u, v};
}

/* A half-sweep of one colour over the level: every update reads the values from before it
 * (vmg.glsl:119-134). */
static void v_half_sweep(vlev* V, int red, v2* tmp) {
    int n = V->L.w * V->L.h;
    memcpy(tmp, V->x, sizeof(v2) * (size_t)n);
    for (int y = 0; y < V->L.h; y++)
        for (int x = 0; x < V->L.w; x++) {
            if ((((x + y) & 1) == 0) != red) continue;
            int k = y * V->L.w + x;
            v2 d = V->d[k];
            v2 o = v_offd(V, tmp, x, y);
            v2 sres = {V->r[k].x - o.x, V->r[k].y - o.y};
            V->x[k] = (v2){d.x > 0.0f ? sres.x / d.x : 0.0f, d.y > 0.0f ? sres.y / d.y : 0.0f};
        }
}

static vlev vlevel(or_sim* s, int l, v2* rhs, v2* xs) {
    vlev V;
    V.L = level_of(s, l);
    V.a = s->vma + V.L.off;
    V.b = s->vmb + V.L.off;
    V.d = s->vmd + V.L.off;
    V.r = rhs;
    V.x = xs;
    return V;
}

/* vmg.glsl:140-168 (way_down) on level l with right-hand side rhs (level-local array):
 * pre-smoothing from zero, vmx of the active tiles, and the restricted residual into vmr of
 * level l + 1 unless l is the coarsest level of the cycle. */
static void vmg_down(or_sim* s, int l, int coarsest, v2* rhs) {
    lvl F = level_of(s, l);
    int n = F.w * F.h, halo = 2 * s->P.v_mg_sweeps;
    v2* xs = (v2*)calloc((size_t)n, sizeof(v2));
    v2* tmp = (v2*)malloc(sizeof(v2) * (size_t)n);
    vlev V = vlevel(s, l, rhs, xs);
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            int k = y * F.w + x;
            v2 d = V.d[k], r = rhs[k];
            v2 z = {d.x > 0.0f ? r.x / d.x : 0.0f, d.y > 0.0f ? r.y / d.y : 0.0f};
            xs[k] = ((x + y) & 1) == 0 ? z : (v2){0.0f, 0.0f};
        }
    for (int h = 1; h < halo; h++) v_half_sweep(&V, (h & 1) == 0, tmp);
    int tx = (F.w + 15) >> 4;
    v2* res = tmp;
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            int k = y * F.w + x;
            v2 d = V.d[k], o = v_offd(&V, xs, x, y);
            res[k].x = (rhs[k].x - d.x * xs[k].x - o.x) * (d.x > 0.0f ? 1.0f : 0.0f);
            res[k].y = (rhs[k].y - d.y * xs[k].y - o.y) * (d.y > 0.0f ? 1.0f : 0.0f);
            if (vflag(s, l, (y >> 4) * tx + (x >> 4))) s->vmx[F.off + k] = xs[k];
        }
    if (!coarsest) {
        lvl C = level_of(s, l + 1);
        for (int cy = 0; cy < C.h; cy++)
            for (int cx = 0; cx < C.w; cx++) {
                int x = 2 * cx, y = 2 * cy;
                if (!vflag(s, l, (y >> 4) * tx + (x >> 4))) continue;
                v2 r00 = vx_(&V, res, x, y), r10 = vx_(&V, res, x + 1, y), r01 = vx_(&V, res, x, y + 1), r11 = vx_(&V, res, x + 1, y + 1);
                s->vmr[lat(C, cx, cy)] = (v2){r00.x + r10.x + r01.x + r11.x, r00.y + r10.y + r01.y + r11.y};
            }
    }
    free(xs);
    free(tmp);
}

/* The start of a way up: x = vmx (+ the coarse correction), masked to the unknowns, then
 * NU sweeps black then red (vmg_up.comp:269-290, visc_mgz.comp:94-105). Returns x. */
static v2* vmg_up_x(or_sim* s, int l, int coarsest, v2* rhs, vlev* V) {
    lvl F = level_of(s, l);
    int n = F.w * F.h, halo = 2 * s->P.v_mg_sweeps;
    v2* xs = (v2*)calloc((size_t)n, sizeof(v2));
    v2* tmp = (v2*)malloc(sizeof(v2) * (size_t)n);
    *V = vlevel(s, l, rhs, xs);
    lvl C = level_of(s, l + 1 < s->lv.n ? l + 1 : s->lv.n - 1);
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            int k = y * F.w + x;
            v2 xv = s->vmx[F.off + k];
            if (!coarsest) {
                v2 zc = s->vmz[lat(C, x >> 1, y >> 1)];
                xv.x += zc.x;
                xv.y += zc.y;
            }
            v2 d = V->d[k];
            xs[k] = (v2){xv.x * (d.x > 0.0f ? 1.0f : 0.0f), xv.y * (d.y > 0.0f ? 1.0f : 0.0f)};
        }
    for (int h = 1; h <= halo; h++) v_half_sweep(V, (h & 1) == 0, tmp);
    free(tmp);
    return xs;
}

/* vmg_up.comp:269-290 on a coarse level l. */
static void vmg_up(or_sim* s, int l, int coarsest) {
    lvl F = level_of(s, l);
    vlev V;
    v2* xs = vmg_up_x(s, l, coarsest, s->vmr + F.off, &V);
    int tx = (F.w + 15) >> 4;
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++)
            if (vflag(s, l, (y >> 4) * tx + (x >> 4))) s->vmz[F.off + y * F.w + x] = xs[y * F.w + x];
    free(xs);
}

/* The coarse levels of a V-cycle, down and up (record_visc_coarse, mod.rs:2838-2853). */
static void visc_coarse(or_sim* s) {
    int n = s->P.v_mg_levels;
    for (int l = 1; l <= n; l++) vmg_down(s, l, l == n, s->vmr + s->lv.offs[l]);
    for (int l = n; l >= 1; l--) vmg_up(s, l, l == n);
}

/* visc_mgz.comp:85-120: the way up on level 0 from residual r; z, A z and the sums
 * r.z, z.Az, z.q (q: this iteration's A p). */
static void visc_mgz(or_sim* s, const v2* r, const v2* q, double sums[3]) {
    lvl F = level_of(s, 0);
    vlev V;
    v2* rl = (v2*)malloc(sizeof(v2) * (size_t)s->N);
    memcpy(rl, r, sizeof(v2) * (size_t)s->N);
    v2* xs = vmg_up_x(s, 0, 0, rl, &V);
    sums[0] = sums[1] = sums[2] = 0.0;
    for (int y = 0; y < F.h; y++)
        for (int x = 0; x < F.w; x++) {
            if (!v_on(s, x, y)) continue;
            int i = y * F.w + x;
            v2 z = xs[i], d = V.d[i], o = v_offd(&V, xs, x, y);
            v2 az = {d.x * z.x + o.x, d.y * z.y + o.y};
            s->vmz[i] = z;
            s->vaz[i] = az;
            sums[0] += dot2(rl[i], z);
            sums[1] += dot2(z, az);
            sums[2] += dot2(z, q[i]);
        }
    free(xs);
    free(rl);
}

/* visc_mgu.comp:28-65: the update of the preconditioned iteration, then the way down on
 * level 0 from the new residual. r_in / q_in: last iteration's; r_out / q_out: this one's. */
static void visc_mgu(or_sim* s, const v2* r_in, const v2* q_in, v2* r_out, v2* q_out, double sums[2]) {
    int W = s->W, H = s->H, N = s->N;
    sp_solve* K = &s->K;
    float a = (float)K->v_alpha, b = (float)K->v_beta;
    v2* rn = (v2*)malloc(sizeof(v2) * (size_t)N);
    for (int i = 0; i < N; i++) {
        v2 qn = b == 0.0f ? s->vaz[i] : (v2){s->vaz[i].x + b * q_in[i].x, s->vaz[i].y + b * q_in[i].y};
        rn[i] = (v2){r_in[i].x - a * qn.x, r_in[i].y - a * qn.y};
    }
    sums[0] = sums[1] = 0.0;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            if (!v_on(s, cx, cy)) continue;
            int i = cidx(s, cx, cy);
            v2 z = s->vmz[i];
            v2 p = b == 0.0f ? z : (v2){z.x + b * s->vp[i].x, z.y + b * s->vp[i].y};
            v2 q = b == 0.0f ? s->vaz[i] : (v2){s->vaz[i].x + b * q_in[i].x, s->vaz[i].y + b * q_in[i].y};
            v2 r = {r_in[i].x - a * q.x, r_in[i].y - a * q.y};
            if (v_interior(s, cx, cy)) {
                if (face_live(s->vel[i].z)) s->vel[i].x += a * p.x;
                if (face_live(s->vel[i].w)) s->vel[i].y += a * p.y;
                s->vd[i].x += a * p.x;
                s->vd[i].y += a * p.y;
            }
            s->vp[i] = p;
            q_out[i] = q;
            r_out[i] = r;
            sums[0] += dot2(r, r);
            sums[1] += dot2(p, q);
        }
    vmg_down(s, 0, 0, rn);
    free(rn);
}

/* The strain of a search direction p (cell values, zero outside the interior) at every cell
 * (visc_iter.comp:55-75). */
static void strain_of_dir(const or_sim* s, const v2* p, v4* out) {
    int W = s->W;
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            v4 w = s->vw[i], o = {0, 0, 0, 0};
#define PU(j) (live_u(s, j) ? p[j].x : 0.0f)
#define PV(j) (live_v(s, j) ? p[j].y : 0.0f)
            if (w.x > 0.0f) {
                o.x = w.x * (PU(i + 1) - PU(i));
                o.y = w.x * (PV(i + W) - PV(i));
            }
            if (w.y > 0.0f) {
                float ua = PU(i), ub = PU(i - W), va = PV(i), vb = PV(i - 1);
                if (closed_u(s, i) && live_u(s, i - W)) ua = -ub;
                if (closed_u(s, i - W) && live_u(s, i)) ub = -ua;
                if (closed_v(s, i) && live_v(s, i - 1)) va = -vb;
                if (closed_v(s, i - 1) && live_v(s, i)) vb = -va;
                o.z = w.y * 0.5f * ((ua - ub) + (va - vb));
            }
#undef PU
#undef PV
            out[i] = o;
        }
}

/* visc_iter.comp: with warm != 0 the warm start (x = u* + d, r -= A d; sums r.z, r.r);
 * otherwise p = z + beta p_old into p_new, q = A p (sums p.q). */
static void visc_iter(or_sim* s, int warm, const v2* p_old, v2* p_new, double sums[2]) {
    int W = s->W, H = s->H, N = s->N;
    float beta = (float)s->K.v_beta;
    v2* pd = (v2*)calloc((size_t)N, sizeof(v2));
    v4* st = (v4*)malloc(sizeof(v4) * (size_t)N);
    for (int cy = 1; cy < H - 1; cy++)
        for (int cx = 1; cx < W - 1; cx++) {
            int i = cidx(s, cx, cy);
            if (warm) {
                pd[i] = s->vd[i];
            } else {
                v4 vw = s->vw[i];
                v2 z = {s->vr[i].x / vw.z, s->vr[i].y / vw.w};
                pd[i] = beta == 0.0f ? z : (v2){z.x + beta * p_old[i].x, z.y + beta * p_old[i].y};
            }
        }
    strain_of_dir(s, pd, st);
    sums[0] = sums[1] = 0.0;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            if (!v_on(s, cx, cy)) continue;
            int i = cidx(s, cx, cy);
            if (!warm) p_new[i] = pd[i];
            if (!v_interior(s, cx, cy)) {
                if (!warm) s->vq[i] = (v2){0, 0};
                continue;
            }
            v2 g = gts_of(s, st, i);
            float lx = live_u(s, i) ? 1.0f : 0.0f, ly = live_v(s, i) ? 1.0f : 0.0f;
            v4 cf = s->coef[i];
            v2 p = {pd[i].x * lx, pd[i].y * ly};
            v2 q = {(cf.z * p.x + g.x) * lx, (cf.w * p.y + g.y) * ly};
            if (warm) {
                s->vel[i].x += p.x;
                s->vel[i].y += p.y;
                v2 r = {s->vr[i].x - q.x, s->vr[i].y - q.y};
                s->vr[i] = r;
                v2 z = {r.x / s->vw[i].z, r.y / s->vw[i].w};
                sums[0] += dot2(r, z);
                sums[1] += dot2(r, r);
            } else {
                s->vq[i] = q;
                sums[0] += dot2(p, q);
            }
        }
    free(pd);
    free(st);
}

/* visc_update.comp:11-36 with this iteration's direction p. */
static void visc_update(or_sim* s, const v2* p, double sums[2]) {
    float a = (float)s->K.v_alpha;
    sums[0] = sums[1] = 0.0;
    for (int cy = 1; cy < s->H - 1; cy++)
        for (int cx = 1; cx < s->W - 1; cx++) {
            if (!v_on(s, cx, cy)) continue;
            int i = cidx(s, cx, cy);
            int lu = live_u(s, i), lv = live_v(s, i);
            if (lu) s->vel[i].x += a * p[i].x;
            if (lv) s->vel[i].y += a * p[i].y;
            s->vd[i].x += a * (lu ? p[i].x : 0.0f);
            s->vd[i].y += a * (lv ? p[i].y : 0.0f);
            v2 r = {s->vr[i].x - a * s->vq[i].x, s->vr[i].y - a * s->vq[i].y};
            s->vr[i] = r;
            v2 z = {r.x / s->vw[i].z, r.y / s->vw[i].w};
            sums[0] += dot2(r, z);
            sums[1] += dot2(r, r);
        }
}

/* visc_apply.comp:30-65 mode 0: r = -G^T W G u*, p = z = r / diag; sums r.z, r.r. */
static void visc_start(or_sim* s, double sums[2]) {
    sums[0] = sums[1] = 0.0;
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < s->W; cx++) {
            if (!v_on(s, cx, cy)) continue;
            int i = cidx(s, cx, cy);
            if (!v_interior(s, cx, cy)) {
                s->vr[i] = s->vp[i] = (v2){0, 0};
                continue;
            }
            v2 g = gts_of(s, s->vs, i);
            float lx = live_u(s, i) ? 1.0f : 0.0f, ly = live_v(s, i) ? 1.0f : 0.0f;
            v2 r = {-g.x * lx, -g.y * ly};
            v2 z = {r.x / s->vw[i].z, r.y / s->vw[i].w};
            s->vr[i] = r;
            s->vp[i] = z;
            sums[0] += dot2(r, z);
            sums[1] += dot2(r, r);
        }
}

/* record_viscosity_split (mod.rs:2812-2833). */
static void stage_visc_split(or_sim* s) {
    int nt = ((s->W + 15) >> 4) * ((s->H + 15) >> 4);
    memset(s->vtiles, 0, sizeof(uint32_t) * (size_t)nt);
    visc_setup(s, 1);
    visc_split(s);
    visc_diag(s);
    if (s->P.v_precond)
        for (int l = 1; l <= s->P.v_mg_levels; l++) vmg_build(s, l);
}

/* record_viscosity_solve (mod.rs:2867-2927) with the scalar steps of reduce_tail.glsl
 * (modes 4, 10, 11, 12, 13 / 5, 6, and 7 at the end) in FP64. */
static void stage_visc_solve(or_sim* s, int restrain) {
    sp_solve* K = &s->K;
    const double tol2 = (double)s->P.tol[2] * (double)s->P.tol[2];
    if (restrain)
        for (int cy = 0; cy < s->H; cy++)
            for (int cx = 0; cx < s->W; cx++)
                if (v_on(s, cx, cy)) {
                    int i = cidx(s, cx, cy);
                    s->vs[i] = strain_of_vel(s, i, s->vw[i].x, s->vw[i].y);
                }
    double a2[2];
    visc_start(s, a2);
    /* mode 4 */
    K->v_rz = a2[0];
    K->v_r0 = a2[1];
    K->v_rr = a2[1];
    K->v_beta = 0.0;
    K->v_alpha = 0.0;
    K->v_iters = 0;
    K->v_done = (a2[1] <= 1e-30 || a2[0] <= 0.0) ? 1u : 0u;
    if (s->P.v_warm && !K->v_done) {
        visc_iter(s, 1, NULL, NULL, a2);
        /* mode 10 */
        K->v_rz = a2[0];
        K->v_rr = a2[1];
        K->v_beta = 0.0;
        K->v_done = (a2[1] <= tol2 * K->v_r0 || a2[0] <= 0.0) ? 1u : 0u;
    }
    uint32_t k = 0;
    uint32_t cap = (uint32_t)s->P.v_max_iters;
    if (s->P.v_precond) {
        /* The residual ping-pongs between vr and vr2, A p between vq and vq2. */
        v2 *rb[2] = {s->vr, s->vr2}, *qb[2] = {s->vq, s->vq2};
        if (!K->v_done) {
            vmg_down(s, 0, 0, s->vr);
            visc_coarse(s);
            double a3[3];
            visc_mgz(s, s->vr, s->vq, a3);
            /* mode 11 */
            K->v_rz = a3[0];
            K->v_beta = 0.0;
            K->v_pq = a3[1];
            K->v_alpha = a3[1] > 0.0 ? a3[0] / a3[1] : 0.0;
            if (a3[0] <= 0.0 || a3[1] <= 0.0) K->v_done = 1u;
        }
        for (; k < cap && !K->v_done; k++) {
            int c = k & 1;
            double m[2];
            visc_mgu(s, rb[c], qb[c], rb[c ^ 1], qb[c ^ 1], m);
            /* mode 12 */
            uint32_t was = K->v_done;
            K->v_iters += 1u - was;
            uint32_t now = m[0] <= tol2 * K->v_r0;
            if (!was) K->v_rr = m[0];
            K->v_pq = m[1];
            K->v_done = was | now;
            if (K->v_done) {
                k++;
                break;
            }
            visc_coarse(s);
            double a3[3];
            visc_mgz(s, rb[c ^ 1], qb[c ^ 1], a3);
            /* mode 13 */
            double beta = (K->v_done == 0u && K->v_rz > 0.0) ? a3[0] / K->v_rz : 0.0;
            double pq = a3[1] + 2.0 * beta * a3[2] + beta * beta * K->v_pq;
            K->v_beta = beta;
            K->v_alpha = (K->v_done == 0u && pq > 0.0) ? a3[0] / pq : 0.0;
            K->v_rz = a3[0];
        }
    } else {
        v2 *pb[2] = {s->vp, s->vp2};
        for (; k < cap && !K->v_done; k++) {
            int c = k & 1;
            double m[2];
            visc_iter(s, 0, pb[c], pb[c ^ 1], m);
            /* mode 5 */
            K->v_alpha = (K->v_done == 0u && m[0] > 0.0) ? K->v_rz / m[0] : 0.0;
            visc_update(s, pb[c ^ 1], m);
            /* mode 6 */
            uint32_t was = K->v_done;
            K->v_iters += 1u - was;
            uint32_t now = m[1] <= tol2 * K->v_r0;
            uint32_t done = was | now;
            K->v_beta = (done == 0u && K->v_rz > 0.0) ? m[0] / K->v_rz : 0.0;
            K->v_rz = m[0];
            if (!was) K->v_rr = m[1];
            K->v_done = done;
        }
    }
    /* mode 7 */
    if (K->v_iters > K->v_it_max) K->v_it_max = K->v_iters;
    if (K->v_done == 0u) K->v_unconv += 1u;
    float rel = K->v_r0 > 0.0 ? (float)sqrt(K->v_rr / K->v_r0) : 0.0f;
    if (rel > K->v_res_max) K->v_res_max = rel;
    K->v_iters_last = K->v_iters;
}

/* rhs_visc.comp:143-168: b = -div(u) + A P1 for the second solve, and its start. */
static void stage_rhs_visc(or_sim* s) {
    int W = s->W;
    float ratio = s->S.dt / s->S.dt_prev;
    /* P1 is read from prhs.z of the neighbours, which this pass also rewrites (with the
     * same value): read it from a copy. */
    float* p1s = (float*)malloc(sizeof(float) * (size_t)s->N);
    for (int i = 0; i < s->N; i++) p1s[i] = s->prhs[i].z;
    for (int cy = 0; cy < s->H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            float D = s->mgd[i];
            if (D <= 0.0f) {
                s->prhs[i].z = 0.0f;
                continue;
            }
            v4 w = s->mga[i];
            float p1 = p1s[i];
            float ap = D * p1;
            if (s->mgd[i - 1] > 0.0f) ap -= w.x * p1s[i - 1];
            if (s->mgd[i + 1] > 0.0f) ap -= w.y * p1s[i + 1];
            if (s->mgd[i - W] > 0.0f) ap -= w.z * p1s[i - W];
            if (s->mgd[i + W] > 0.0f) ap -= w.w * p1s[i + W];
            float aL = s->solid[i].x, aB = s->solid[i].y, aR = s->solid[i + 1].x, aT = s->solid[i + W].y;
            v4 vme = s->vel[i];
            float div = aR * s->vel[i + 1].x - aL * vme.x + aT * s->vel[i + W].y - aB * vme.y;
            v4 r = s->prhs[i];
            s->prhs[i] = (v4){ap - div, r.y, p1, 0.0f};
            if (s->P.v_p2warm) s->pres[i].x = p1 + s->p2keep[i] * ratio;
        }
    free(p1s);
}

/* project_visc.comp:187-210: subtract the gradient of P2 = P - P1, keep P2. */
static void stage_project_visc(or_sim* s) {
    float m = u2f(s->S.umax_grid);
#define P2(j) (s->mgd[j] > 0.0f ? s->pres[j].x - s->prhs[j].z : 0.0f)
    for (int cy = 1; cy < s->H; cy++)
        for (int cx = 1; cx < s->W; cx++) {
            int i = cidx(s, cx, cy);
            s->p2keep[i] = P2(i);
            v4 cf = s->coef[i];
            v4 me = s->vel[i];
            int lu = face_live(me.z) && cf.x != 0.0f, lv = face_live(me.w) && cf.y != 0.0f;
            if (lu || lv) {
                float pc = P2(i);
                if (lu) me.x -= cf.x * (pc - P2(i - 1));
                if (lv) me.y -= cf.y * (pc - P2(i - s->W));
                s->vel[i] = me;
            }
            float mm = 0.0f;
            if (lu) mm = fabsf(me.x);
            if (lv) mm = fmaxf_(mm, fabsf(me.y));
            if (mm > m) m = mm;
        }
#undef P2
    s->S.umax_grid = f2u(m);
}

/* The viscosity before the first pressure solve (record_substep, mod.rs:2681-2686). */
static void stage_visc_before(or_sim* s) {
    if (s->P.visc_mode == 1) {
        stage_visc_explicit(s);
    } else if (s->P.visc_mode == 2) {
        stage_visc_split(s);
        if (!s->P.v_twice) stage_visc_solve(s, 0);
    }
}

/* ---- rigid bodies (rigid.comp) --------------------------------------------------------- */

/* shear_rate (common.glsl:527-540) */
static float uf_(const or_sim* s, int k) { return face_live(s->vel[k].z) ? s->vel[k].x : 0.0f; }
static float vf_(const or_sim* s, int k, float gv) { return face_live(s->vel[k].w) ? s->vel[k].y + gv : 0.0f; }
static float shear_rate(const or_sim* s, int i, int pre) {
    const int W = s->W;
    const float gv = pre ? s->P.gravity * s->S.dt : 0.0f;
    const float dxx = uf_(s, i + 1) - uf_(s, i);
    const float dyy = vf_(s, i + W, gv) - vf_(s, i, gv);
    const float duy = 0.25f * (uf_(s, i + W) + uf_(s, i + 1 + W) - uf_(s, i - W) - uf_(s, i + 1 - W));
    const float dvx = 0.25f * (vf_(s, i + 1, gv) + vf_(s, i + 1 + W, gv) - vf_(s, i - 1, gv) - vf_(s, i - 1 + W, gv));
    const float dxy = 0.5f * (duy + dvx);
    return sqrtf(2.0f * (dxx * dxx + dyy * dyy + 2.0f * dxy * dxy));
}

/* nn_eval (common.glsl:546-571): the factor on the Newtonian viscosity and the yield stress
 * of cell i's share of its non-Newtonian material at shear rate g */
static void nn_eval(const or_sim* s, int i, v4 q, float g, float* F, float* tau) {
    const sp_gmat* m = &s->mats[(int)q.y];
    const float n = m->rheo_p[1], lam = m->rheo_p[2], ratio = m->rheo_p[3];
    *F = 1.0f;
    *tau = 0.0f;
    if (m->rheo == 1) {
        *F = ratio + (1.0f - ratio) * powf(1.0f + (lam * g) * (lam * g), 0.5f * (n - 1.0f));
    } else if (m->rheo == 2) {
        *F = powf(fmaxf_(g, 1e-6f), n - 1.0f);
        *tau = m->rheo_p[0] * q.z;
    } else if (m->rheo == 3) {
        *F = fminf_(powf(1.0f + (lam * g) * (lam * g), 0.5f * (n - 1.0f)), ratio);
    }
    const float fr = m->gr[0];
    if (fr > 0.0f) *tau += fr * q.z * q.z * q.z * fmaxf_(s->pres[i].x, 0.0f) / fmaxf_(s->S.dt, 1e-6f);
    const int W = s->W;
    const float skin = m->temp[3];
    if (skin > 0.0f && (s->cell[i - 1].w <= 0.0f || s->cell[i + 1].w <= 0.0f || s->cell[i - W].w <= 0.0f || s->cell[i + W].w <= 0.0f))
        *tau += skin;
    *F = 1.0f + q.x * (*F - 1.0f);
    *tau *= q.x;
}

/* rheo.comp (rheo.comp:31-55) */
static void stage_rheo(or_sim* s) {
    const int W = s->W, H = s->H;
    for (int cy = 1; cy < H - 1; cy++)
        for (int cx = 1; cx < W - 1; cx++) {
            const int i = cy * W + cx;
            v4 me = s->cell[i];
            v4 q = s->rh[i];
            if (me.w <= 0.0f || q.x <= 0.0f) continue;
            const float g = shear_rate(s, i, 1);
            float F, tau;
            nn_eval(s, i, q, g, &F, &tau);
            const float eff = me.z * F + tau / fmaxf_(g, 1e-6f);
            const float capv = me.y * expf(s->P.ln_rigid);
            const float skin = q.x * s->mats[(int)q.y].temp[3];
            if (skin > 0.0f && me.y * g * g < skin &&
                (s->cell[i - 1].w <= 0.0f || s->cell[i + 1].w <= 0.0f || s->cell[i - W].w <= 0.0f || s->cell[i + W].w <= 0.0f)) {
                q.w = 1.0f;
                s->rh[i] = q;
            }
            me.z = fminf_(eff, capv);
            s->cell[i] = me;
        }
}

/* rigid (rigid.comp:34-47) */
static int rigid_cell(const or_sim* s, int cx, int cy, int i) {
    const v4 me = s->cell[i];
    if (me.w <= 0.0f) return 0;
    if (thermal(s) && s->therm[i].w > SP_RIGID_SHARE) return 1;
    const v4 q = s->rh[i];
    if (q.w > 0.5f) return 1;
    if (!(s->P.caps & SP_CAP_RHEO) || q.x <= 0.0f || cx < 1 || cy < 1 || cx >= s->W - 1 || cy >= s->H - 1) return 0;
    const float g = shear_rate(s, i, 0);
    float F, tau;
    nn_eval(s, i, q, g, &F, &tau);
    if (tau > 0.0f && me.z * g < tau) return 1;
    return s->mats[(int)q.y].rheo == 3 && me.z * F >= me.y * expf(s->P.ln_rigid);
}

static int rigid_root(const or_sim* s, int i) {
    while (s->rl[2 * i] != i) i = s->rl[2 * i];
    return i;
}

/* round(x) into the 64-bit sum at words w, w + 1 (add64, rigid.comp:92-99) */
static void rigid_add64(or_sim* s, uint32_t w, float x) {
    const float r = floorf(x + 0.5f);
    uint64_t v;
    memcpy(&v, s->rbody + w, 8);
    v += (uint64_t)(int64_t)r;
    memcpy(s->rbody + w, &v, 8);
}

/* rigid.comp modes 0-4 (rigid.comp:101-196). The union joins 8-connected rigid cells with the
 * smallest cell index as each body's root (what the reference's atomicMin union-find leaves);
 * bodies claim slots in the order of their roots (the reference's order of arrival, which
 * the sums and motions do not depend on). */
static void stage_rigid(or_sim* s, int mode) {
    const int W = s->W, H = s->H, N = s->N;
    if (mode == 0) {
        for (int i = 0; i < N; i++) {
            s->rl[2 * i] = rigid_cell(s, i % W, i / W, i) ? i : -1;
            s->rl[2 * i + 1] = -1;
        }
        s->rbody[0] = 0u;
        return;
    }
    if (mode == 1) {
        for (int cy = 0; cy < H; cy++)
            for (int cx = 0; cx < W; cx++) {
                const int i = cy * W + cx;
                if (s->rl[2 * i] < 0) continue;
                const int nb[4] = {cx > 0 ? i - 1 : -1, cy > 0 ? i - W : -1, cy > 0 && cx > 0 ? i - W - 1 : -1,
                                   cy > 0 && cx < W - 1 ? i - W + 1 : -1};
                for (int k = 0; k < 4; k++) {
                    if (nb[k] < 0 || s->rl[2 * nb[k]] < 0) continue;
                    int a = rigid_root(s, i), b = rigid_root(s, nb[k]);
                    if (a < b) s->rl[2 * b] = a;
                    else if (b < a) s->rl[2 * a] = b;
                }
            }
        return;
    }
    if (mode == 2) {
        for (int i = 0; i < N; i++)
            if (s->rl[2 * i] >= 0) s->rl[2 * i] = rigid_root(s, i);
        for (int i = 0; i < N; i++) {
            if (s->rl[2 * i] != i) continue;
            const uint32_t b = s->rbody[0]++;
            if (b < (uint32_t)SP_RIGID_BODIES) {
                memset(s->rbody + 16 + SP_BODY_WORDS * b, 0, 4 * SP_BODY_WORDS);
                s->rl[2 * i + 1] = (int)b;
            }
        }
        return;
    }
    if (mode == 3) {
        for (int cy = 0; cy < H; cy++)
            for (int cx = 0; cx < W; cx++) {
                const int i = cy * W + cx;
                const int l = s->rl[2 * i];
                if (l < 0) continue;
                const int b = s->rl[2 * l + 1];
                if (b < 0) continue;
                const v4 me = s->cell[i];
                const float m = me.y * fminf_(me.x, 1.0f) * SP_RIGID_FIX;
                const float dx = (float)(cx - l % W), dy = (float)(cy - l / W);
                const int last = N - 1;
                const v4 q = s->vel[i];
                const float ux = 0.5f * (q.x + s->vel[i + 1 < last ? i + 1 : last].x);
                const float uy = 0.5f * (q.y + s->vel[i + W < last ? i + W : last].y);
                const uint32_t o = 16u + (uint32_t)SP_BODY_WORDS * (uint32_t)b;
                rigid_add64(s, o, m);
                rigid_add64(s, o + 2u, m * dx);
                rigid_add64(s, o + 4u, m * dy);
                rigid_add64(s, o + 6u, m * ux);
                rigid_add64(s, o + 8u, m * uy);
                rigid_add64(s, o + 10u, m * (dx * uy - dy * ux));
                rigid_add64(s, o + 12u, m * (dx * dx + dy * dy + 1.0f / 6.0f));
                if (s->P.rigid_friction >= 0.0f) {
                    float gx, gy;
                    const float dist = sample_sdf(s, (float)cx + 0.5f, (float)cy + 0.5f, &gx, &gy);
                    if (dist < 1.0f && gx * gx + gy * gy > 1e-8f) {
                        const float gl = len2(gx, gy);
                        const float nx = gx / gl, ny = gy / gl;
                        const float F = SP_RIGID_FIX;
                        rigid_add64(s, o + 14u, F);
                        rigid_add64(s, o + 16u, F * nx);
                        rigid_add64(s, o + 18u, F * ny);
                        rigid_add64(s, o + 20u, F * dx);
                        rigid_add64(s, o + 22u, F * dy);
                        rigid_add64(s, o + 24u, F * dx * dx);
                        rigid_add64(s, o + 26u, F * dy * dy);
                        rigid_add64(s, o + 28u, F * dx * dy);
                    }
                }
            }
        return;
    }
    /* mode 4: faces between two cells of one body move with it */
    const float clump = SP_REST_PHI + fmaxf_(s->P.corr[1], 0.05f);
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            const int i = cy * W + cx;
            const int l = s->rl[2 * i];
            v4 q = s->vel[i];
            int changed = 0;
            float u[3];
            const int cl = s->cell[i].x > clump;
            if (l >= 0 && cx > 0 && s->rl[2 * (i - 1)] == l && face_live(q.z) && s->rl[2 * l + 1] >= 0 &&
                body_motion(s, (uint32_t)s->rl[2 * l + 1], l, (float)cx, (float)cy + 0.5f, u)) {
                q.x = u[0];
                if (!cl && s->cell[i - 1].x <= clump) q.z = 0.0f;
                changed = 1;
            }
            if (l >= 0 && cy > 0 && s->rl[2 * (i - W)] == l && face_live(q.w) && s->rl[2 * l + 1] >= 0 &&
                body_motion(s, (uint32_t)s->rl[2 * l + 1], l, (float)cx + 0.5f, (float)cy, u)) {
                q.y = u[1];
                if (!cl && s->cell[i - W].x <= clump) q.w = 0.0f;
                changed = 1;
            }
            if (changed) s->vel[i] = q;
        }
}

/* ---- public entry points -------------------------------------------------------------- */

/* ---- settle.comp ---------------------------------------------------------------------- */

static float fam_conc(const or_sim* s, int i, int f) {
    return f == 0 ? s->comp[i].z : s->mixb[(size_t)i * (s->P.n_fam - 1) + f - 1];
}

/* Flux (cell volumes / s) of family f's concentrate from cell up into cell dn below it: the
 * hindered flux v_t s_up (1 - s_dn) min(occupancies), limited to what up holds and what dn
 * has room for over the substep (settle.comp:21-30). */
static float settle_flux(const or_sim* s, int up, int dn, int f, float vt, float dt) {
    int nf = s->P.n_fam;
    float ou = s->famocc[(size_t)up * nf + f], od = s->famocc[(size_t)dn * nf + f];
    if (ou <= 1e-3f || od <= 1e-3f) return 0.0f;
    float su = fam_conc(s, up, f), sd = fam_conc(s, dn, f);
    if (su <= 0.0f || sd < 0.0f) return 0.0f;
    float F = vt * su * (1.0f - sd) * fminf_(ou, od);
    return fminf_(F, fminf_(su * ou, (1.0f - sd) * od) / dt);
}

/* Each cell's gain over its room and loss over its holding, per family (settle.comp:32-53).
 * Reads only comp, mixb, famocc and therm, which the pass does not write. */
static void stage_settle(or_sim* s) {
    const sp_gparams* P = &s->P;
    int W = s->W, H = s->H, nf = P->n_fam;
    float dt = s->S.dt;
    int th_on = thermal(s);
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            for (int f = 0; f < nf; f++) {
                int row = P->fams[f][0], dl = P->fams[f][1];
                float d = s->mats[row].grain;
                float occ = s->famocc[(size_t)i * nf + f];
                v2 ds = {0.0f, 0.0f};
                if (d > 0.0f && occ > 1e-3f && cy > 0 && cy < H - 1) {
                    float drho = s->mats[row].rho - s->mats[dl].rho;
                    float T = th_on ? s->therm[i].x : P->heat_x;
                    float vt = terminal(s, drho, s->mats[dl].rho, expf(row_lnu(s, dl, 1.0f, T, th_on)), d);
                    float gain = drho > 0.0f ? settle_flux(s, i + W, i, f, vt, dt) : settle_flux(s, i - W, i, f, vt, dt);
                    float loss = drho > 0.0f ? settle_flux(s, i, i - W, f, vt, dt) : settle_flux(s, i, i + W, f, vt, dt);
                    float sc = fam_conc(s, i, f);
                    ds = (v2){dt * (gain / fmaxf_(occ * (1.0f - sc), 1e-6f)), dt * (loss / fmaxf_(occ * sc, 1e-6f))};
                }
                s->dsettle[(size_t)i * nf + f] = ds;
            }
        }
}

/* ---- magnet.comp ---------------------------------------------------------------------- */

/* The field of the placed magnets at x (cells), each a point dipole in the plane, B0 on the
 * axis at the rim and constant inside it (common.glsl:390-401). */
static v2 magnet_field(const or_sim* s, float x, float y) {
    v2 B = {0.0f, 0.0f};
    for (int k = 0; k < (int)s->P.mag[0]; k++) {
        const float* mg = s->P.magnet[k];
        float mx = s->P.magnet_dir[k][0], my = s->P.magnet_dir[k][1];
        float dx = x - mg[0], dy = y - mg[1];
        float dl = len2(dx, dy);
        float r = fmaxf_(dl, mg[2]);
        float ex = mx, ey = my;
        if (dl > 1e-4f) {
            ex = dx / dl;
            ey = dy / dl;
        }
        float k3 = mg[3] * powf(mg[2] / r, 3.0f) * 0.5f;
        float me = 3.0f * (mx * ex + my * ey);
        B.x += k3 * (me * ex - mx);
        B.y += k3 * (me * ey - my);
    }
    return B;
}

/* The velocity change of the face between cell i and the cell n behind it: the magnetisable
 * share of the face (the mean of the two cells' susceptibilities where both are liquid, the
 * liquid cell's where one is air, over the tank's largest) times the difference of the
 * energy density over the face density (magnet.comp:63-67). */
static float mag_face_force(const or_sim* s, int i, int n, float rho) {
    int li = s->cell[i].w > 0.0f, ln = s->cell[n].w > 0.0f;
    float sh = (li == ln ? 0.5f * (s->magc[i].x + s->magc[n].x) : li ? s->magc[i].x : s->magc[n].x) / s->P.mag[1];
    return -s->S.dt * sh * (s->magc[i].y - s->magc[n].y) / fmaxf_(rho, 0.05f);
}

/* Mode 0: the magnetic energy density of a linear liquid that saturates, demagnetised as a
 * sphere (blended towards the shape's N by `demag`), per cell; mode 1: its force on the live
 * faces (magnet.comp:71-105). Mode 1 reads psi of neighbours and writes only velocities. */
static void stage_magnet(or_sim* s, int mode) {
    const sp_gparams* P = &s->P;
    int W = s->W, H = s->H;
    for (int cy = 0; cy < H; cy++)
        for (int cx = 0; cx < W; cx++) {
            int i = cidx(s, cx, cy);
            if (mode == 1) {
                if (cx < 1 || cy < 1) continue;
                v4 me = s->vel[i], cf = s->coef[i];
                if (face_live(me.z) && cf.x != 0.0f) me.x += mag_face_force(s, i, i - 1, cf.z);
                if (face_live(me.w) && cf.y != 0.0f) me.y += mag_face_force(s, i, i - W, cf.w);
                s->vel[i] = me;
                continue;
            }
            v2 B = magnet_field(s, (float)cx + 0.5f, (float)cy + 0.5f);
            float Bl = len2(B.x, B.y);
            float Hf = Bl / SP_MU0;
            float psi = 0.0f;
            float chi = P->mag[1];
            if (Hf > 0.0f && chi > 0.0f) {
                float N = 1.0f / 3.0f;
                if (P->demag > 0.0f) {
                    float gx = 0.0f, gy = 0.0f;
                    if (cx > 0 && cy > 0 && cx < W - 1 && cy < H - 1) {
                        gx = 0.5f * (s->surf[i + 1].x - s->surf[i - 1].x);
                        gy = 0.5f * (s->surf[i + W].x - s->surf[i - W].x);
                    }
                    float gl = len2(gx, gy);
                    float nh = gl > 1e-6f ? (gx / gl) * (B.x / Bl) + (gy / gl) * (B.y / Bl) : 0.0f;
                    N = mixf(1.0f / 3.0f, nh * nh, clampf(gl / 0.25f, 0.0f, 1.0f) * P->demag);
                }
                float ce = chi / (1.0f + N * chi);
                float Ms = P->mag[2] / SP_MU0;
                float xs = ce * Hf / Ms;
                /* x - ln(1 + x), by its series where float cancellation would eat it */
                float f = xs < 0.02f ? xs * xs * (0.5f - xs * (1.0f / 3.0f - 0.25f * xs)) : xs - logf(1.0f + xs);
                psi = -SP_MU0 * Ms * Ms / ce * f * P->mag[3];
            }
            s->magc[i].y = psi;
        }
}

/* ---- elastic.comp --------------------------------------------------------------------- */

/* tau_xy at the lower-left corner of cell i (elastic.comp:15-18) */
static float txy_node(const or_sim* s, int i) {
    int W = s->W;
    return 0.25f * (((s->estr[i].z + s->estr[i - 1].z) + s->estr[i - W].z) + s->estr[i - W - 1].z);
}

/* The divergence of the stress on the live faces, explicitly, cells 2 .. grid - 2
 * (elastic.comp:20-38). Each cell writes only its own faces from cell and estr, which the
 * pass does not change: the order of the cells does not matter. */
static void stage_elastic(or_sim* s) {
    int W = s->W, H = s->H;
    float dt = s->S.dt;
    for (int cy = 2; cy < H - 1; cy++)
        for (int cx = 2; cx < W - 1; cx++) {
            int i = cidx(s, cx, cy);
            v4 q = s->vel[i];
            v4 a = s->cell[i], l = s->cell[i - 1], b = s->cell[i - W];
            if (face_live(q.z)) {
                float rho = (a.x * a.y + l.x * l.y) / fmaxf_(a.x + l.x, 1e-6f);
                float f = (s->estr[i].x - s->estr[i - 1].x) + (txy_node(s, i + W) - txy_node(s, i));
                q.x += dt * f / fmaxf_(rho, 0.05f);
            }
            if (face_live(q.w)) {
                float rho = (a.x * a.y + b.x * b.y) / fmaxf_(a.x + b.x, 1e-6f);
                float f = (s->estr[i].y - s->estr[i - W].y) + (txy_node(s, i + 1) - txy_node(s, i));
                q.y += dt * f / fmaxf_(rho, 0.05f);
            }
            s->vel[i] = q;
        }
}

void or_stage(or_sim* s, int stage, int arg) {
    switch (stage) {
    case SP_STAGE_STEP: stage_step(s); break;
    case SP_STAGE_CLEAR_ACC: stage_clear_acc(s); break;
    case SP_STAGE_PARTICLES: stage_particles(s, arg); break;
    case SP_STAGE_GRID_PREP: stage_grid_prep(s); break;
    case SP_STAGE_COEF: stage_coef(s); break;
    case SP_STAGE_SURF:
        stage_surf_smooth(s);
        stage_surf_curv(s);
        stage_surf_apply(s);
        break;
    case SP_STAGE_VISC: stage_visc_before(s); break;
    case SP_STAGE_VISC_SPLIT: stage_visc_split(s); break;
    case SP_STAGE_VISC_SOLVE: stage_visc_solve(s, arg); break;
    case SP_STAGE_RHS_VISC: stage_rhs_visc(s); break;
    case SP_STAGE_PRESSURE2: pressure_iterate(s, 1); break;
    case SP_STAGE_PROJECT_VISC: stage_project_visc(s); break;
    case SP_STAGE_VISC_VCYCLE: {
        double a3[3];
        vmg_down(s, 0, 0, s->vr);
        visc_coarse(s);
        visc_mgz(s, s->vr, s->vq, a3);
        break;
    }
    case SP_STAGE_HEAT: stage_heat(s); break;
    case SP_STAGE_RHEO: stage_rheo(s); break;
    case SP_STAGE_ELASTIC: stage_elastic(s); break;
    case SP_STAGE_SETTLE: stage_settle(s); break;
    case SP_STAGE_MAGNET:
        if (arg == 0 || arg == 1) stage_magnet(s, arg);
        break;
    case SP_STAGE_RIGID:
        if (arg >= 0 && arg <= 4) stage_rigid(s, arg);
        break;
    case SP_STAGE_RHS: stage_rhs(s); break;
    case SP_STAGE_PRESSURE: stage_pressure(s); break;
    case SP_STAGE_PROJECT: stage_project(s); break;
    case SP_STAGE_SUBSTEP: or_substep(s, arg); break;
    case SP_STAGE_MG_BUILD: stage_mg_build(s); break;
    case SP_STAGE_VCYCLE: {
        double a[4] = {0, 0, 0, 0};
        vcycle(s, a);
        break;
    }
    default: break;
    }
}

void or_substep(or_sim* s, int gather) {
    stage_step(s);
    stage_clear_acc(s);
    stage_particles(s, gather);
    stage_grid_prep(s);
    if (s->P.caps & SP_CAP_RHEO) stage_rheo(s);
    if (s->P.caps & SP_CAP_ELASTIC) stage_elastic(s);
    if (s->P.caps & SP_CAP_SETTLE) stage_settle(s);
    stage_coef(s);
    stage_surf_smooth(s);
    stage_surf_curv(s);
    stage_surf_apply(s);
    if (s->P.caps & SP_CAP_MAGNET) {
        stage_magnet(s, 0);
        stage_magnet(s, 1);
    }
    if (thermal(s) && s->P.heat_rate > 0.0f) stage_heat(s);
    stage_visc_before(s);
    stage_rhs(s);
    stage_pressure(s);
    stage_project(s);
    if (s->P.visc_mode == 2 && s->P.v_twice) {
        stage_visc_solve(s, 1);
        stage_rhs_visc(s);
        pressure_iterate(s, 1);
        stage_project_visc(s);
    }
    if (s->P.caps & SP_CAP_RIGID)
        for (int mode = 0; mode < 5; mode++) stage_rigid(s, mode);
}

or_sim* or_create(int32_t W, int32_t H, uint32_t capacity) {
    or_sim* s = (or_sim*)calloc(1, sizeof(or_sim));
    s->W = W;
    s->H = H;
    s->N = W * H;
    s->cap = capacity;
    size_t N = (size_t)s->N;
    s->part = (sp_particle*)calloc(capacity ? capacity : 1, sizeof(sp_particle));
    s->acc = (int32_t*)calloc(N * SP_ACC_MAX, sizeof(int32_t));
    v4** f4[] = {&s->vel, &s->cell, &s->comp, &s->coef, &s->surf, &s->surf2, &s->sprop, &s->aux, &s->grho, &s->gbase, &s->solid, &s->prhs, &s->vw, &s->vs};
    for (size_t k = 0; k < sizeof(f4) / sizeof(f4[0]); k++) *f4[k] = (v4*)calloc(N, sizeof(v4));
    s->vel0 = (v2*)calloc(N, sizeof(v2));
    s->pjump = (v2*)calloc(N, sizeof(v2));
    s->pres = (v2*)calloc(N, sizeof(v2));
    s->sdf = (float*)calloc((size_t)(W + 1) * (H + 1), sizeof(float));
    s->mixb = (float*)calloc(N * SP_MAX_PRESENT, sizeof(float));
    s->p2keep = (float*)calloc(N, sizeof(float));
    s->region = (uint32_t*)calloc(N, sizeof(uint32_t));
    or_levels(W, H, &s->lv);
    size_t T = (size_t)s->lv.total;
    s->mga = (v4*)calloc(T, sizeof(v4));
    s->mgd = (float*)calloc(T, sizeof(float));
    s->mgx = (v2*)calloc(T, sizeof(v2));
    s->mgz = (v2*)calloc(T, sizeof(v2));
    s->mgr = (v2*)calloc(T, sizeof(v2));
    s->cgp = (v2*)calloc(N, sizeof(v2));
    s->cgq = (v2*)calloc(N, sizeof(v2));
    s->ptile = (uint32_t*)calloc((size_t)s->lv.tiles, sizeof(uint32_t));
    v2** f2[] = {&s->vr, &s->vp, &s->vp2, &s->vq, &s->vd, &s->vr2, &s->vq2, &s->vaz};
    for (size_t k = 0; k < sizeof(f2) / sizeof(f2[0]); k++) *f2[k] = (v2*)calloc(N, sizeof(v2));
    s->vtiles = (uint32_t*)calloc((size_t)(((W + 15) / 16) * ((H + 15) / 16)), sizeof(uint32_t));
    s->vtile = (uint32_t*)calloc((size_t)s->lv.tiles, sizeof(uint32_t));
    s->vma = (v4*)calloc(T, sizeof(v4));
    s->vmb = (v4*)calloc(T, sizeof(v4));
    v2** t2[] = {&s->vmd, &s->vmr, &s->vmx, &s->vmz};
    for (size_t k = 0; k < sizeof(t2) / sizeof(t2[0]); k++) *t2[k] = (v2*)calloc(T, sizeof(v2));
    s->therm = (v4*)calloc(N, sizeof(v4));
    s->dheat = (v4*)calloc(N, sizeof(v4));
    s->hcell = (int32_t*)calloc(2 * N, sizeof(int32_t));
    s->solid_t = (float*)calloc(N, sizeof(float));
    s->rl = (int32_t*)calloc(2 * (size_t)N, sizeof(int32_t));
    s->rbody = (uint32_t*)calloc(SP_RIGID_WORDS, sizeof(uint32_t));
    s->rh = (v4*)calloc(N, sizeof(v4));
    s->estr = (v4*)calloc(N, sizeof(v4));
    s->pstress = (v4*)calloc(capacity ? capacity : 1, sizeof(v4));
    s->famocc = (float*)calloc(N * SP_MAX_PRESENT, sizeof(float));
    s->dsettle = (v2*)calloc(N * SP_MAX_PRESENT, sizeof(v2));
    s->magc = (v2*)calloc(N, sizeof(v2));
    s->rmask = (uint32_t*)calloc(N, sizeof(uint32_t));
    for (size_t i = 0; i < N; i++) s->solid_t[i] = SP_NO_TEMP;
    return s;
}

void or_destroy(or_sim* s) {
    if (!s) return;
    free(s->part);
    free(s->acc);
    v4* f4[] = {s->vel, s->cell, s->comp, s->coef, s->surf, s->surf2, s->sprop, s->aux, s->grho, s->gbase, s->solid, s->prhs, s->vw, s->vs};
    for (size_t k = 0; k < sizeof(f4) / sizeof(f4[0]); k++) free(f4[k]);
    free(s->vel0);
    free(s->pjump);
    free(s->pres);
    free(s->sdf);
    free(s->mixb);
    free(s->p2keep);
    free(s->region);
    free(s->mga);
    free(s->mgd);
    free(s->mgx);
    free(s->mgz);
    free(s->mgr);
    free(s->cgp);
    free(s->cgq);
    free(s->ptile);
    v2* f2[] = {s->vr, s->vp, s->vp2, s->vq, s->vd, s->vr2, s->vq2, s->vaz, s->vmd, s->vmr, s->vmx, s->vmz};
    for (size_t k = 0; k < sizeof(f2) / sizeof(f2[0]); k++) free(f2[k]);
    free(s->vtiles);
    free(s->vtile);
    free(s->vma);
    free(s->vmb);
    free(s->therm);
    free(s->dheat);
    free(s->hcell);
    free(s->solid_t);
    free(s->rl);
    free(s->rbody);
    free(s->rh);
    free(s->estr);
    free(s->pstress);
    free(s->famocc);
    free(s->dsettle);
    free(s->magc);
    free(s->rmask);
    free(s);
}

void or_set_gparams(or_sim* s, const sp_gparams* P, const sp_gmat* mats) {
    s->P = *P;
    memcpy(s->mats, mats, sizeof(sp_gmat) * SP_MAX_MATERIALS);
}

/* temp: the temperature each cell's solid is held at, or NULL (insulating everywhere,
 * mod.rs:2442-2447). */
void or_set_solids(or_sim* s, const float* sdf, const float* solid, const uint32_t* region, const float* temp) {
    memcpy(s->sdf, sdf, sizeof(float) * (size_t)(s->W + 1) * (s->H + 1));
    memcpy(s->solid, solid, sizeof(v4) * (size_t)s->N);
    memcpy(s->region, region, sizeof(uint32_t) * (size_t)s->N);
    for (int i = 0; i < s->N; i++) s->solid_t[i] = temp ? temp[i] : SP_NO_TEMP;
}

void or_reset(or_sim* s, const sp_particle* p, uint32_t n) {
    if (n > s->cap) n = s->cap;
    memcpy(s->part, p, sizeof(sp_particle) * n);
    s->n_slots = n;
    s->P.n_slots = n;
    memset(s->pres, 0, sizeof(v2) * (size_t)s->N);
    memset(s->p2keep, 0, sizeof(float) * (size_t)s->N);
    memset(s->vd, 0, sizeof(v2) * (size_t)s->N);
    memset(&s->S, 0, sizeof(s->S));
    memset(s->dheat, 0, sizeof(v4) * (size_t)s->N);
    memset(s->ledger, 0, sizeof(s->ledger));
    memset(s->pstress, 0, sizeof(v4) * (s->cap ? s->cap : 1));
}

void or_begin_batch(or_sim* s, float target, uint32_t substeps) {
    s->S.frame_left = target > 0.0f ? target : 0.0f;
    s->S.batch_steps = substeps;
    s->S.batch_done = 0;
}

uint32_t or_n_slots(or_sim* s) { return s->n_slots; }
uint32_t or_last_iters(or_sim* s) { return s->K.iters_last; }

void or_air_velocity(or_sim* s, float x, float y, float lean, float* vel2, float* occ) {
    v2 v = air_velocity(s, x, y, lean, occ);
    vel2[0] = v.x;
    vel2[1] = v.y;
}

static void* field_ptr(or_sim* s, int field, size_t* bytes) {
    size_t N = (size_t)s->N, T = (size_t)s->lv.total;
    switch (field) {
    case SP_FIELD_PARTICLES: *bytes = sizeof(sp_particle) * s->n_slots; return s->part;
    case SP_FIELD_ACC: *bytes = sizeof(int32_t) * N * s->P.acc_stride; return s->acc;
    case SP_FIELD_VEL: *bytes = sizeof(v4) * N; return s->vel;
    case SP_FIELD_CELL: *bytes = sizeof(v4) * N; return s->cell;
    case SP_FIELD_COMP: *bytes = sizeof(v4) * N; return s->comp;
    case SP_FIELD_COEF: *bytes = sizeof(v4) * N; return s->coef;
    case SP_FIELD_SURF: *bytes = sizeof(v4) * N; return s->surf;
    case SP_FIELD_PJUMP: *bytes = sizeof(v2) * N; return s->pjump;
    case SP_FIELD_MGA: *bytes = sizeof(v4) * T; return s->mga;
    case SP_FIELD_MGD: *bytes = sizeof(float) * T; return s->mgd;
    case SP_FIELD_PRHS: *bytes = sizeof(v4) * N; return s->prhs;
    case SP_FIELD_PRES: *bytes = sizeof(v2) * N; return s->pres;
    case SP_FIELD_STEP: *bytes = sizeof(sp_stepbuf); return &s->S;
    case SP_FIELD_SOLVE: *bytes = sizeof(sp_solve); return &s->K;
    case SP_FIELD_SPROP: *bytes = sizeof(v4) * N; return s->sprop;
    case SP_FIELD_AUX: *bytes = sizeof(v4) * N; return s->aux;
    case SP_FIELD_GRHO: *bytes = sizeof(v4) * N; return s->grho;
    case SP_FIELD_VEL0: *bytes = sizeof(v2) * N; return s->vel0;
    case SP_FIELD_MGZ: *bytes = sizeof(v2) * T; return s->mgz;
    case SP_FIELD_MGR: *bytes = sizeof(v2) * T; return s->mgr;
    case SP_FIELD_PTILE: *bytes = sizeof(uint32_t) * (size_t)s->lv.tiles; return s->ptile;
    case SP_FIELD_REGION_AIR: *bytes = sizeof(uint32_t); return &s->region_air;
    case SP_FIELD_SURF2: *bytes = sizeof(v4) * N; return s->surf2;
    case SP_FIELD_VW: *bytes = sizeof(v4) * N; return s->vw;
    case SP_FIELD_VS: *bytes = sizeof(v4) * N; return s->vs;
    case SP_FIELD_GBASE: *bytes = sizeof(v4) * N; return s->gbase;
    case SP_FIELD_VR: *bytes = sizeof(v2) * N; return s->vr;
    case SP_FIELD_VP: *bytes = sizeof(v2) * N; return s->vp;
    case SP_FIELD_VP2: *bytes = sizeof(v2) * N; return s->vp2;
    case SP_FIELD_VQ: *bytes = sizeof(v2) * N; return s->vq;
    case SP_FIELD_VD: *bytes = sizeof(v2) * N; return s->vd;
    case SP_FIELD_VR2: *bytes = sizeof(v2) * N; return s->vr2;
    case SP_FIELD_VQ2: *bytes = sizeof(v2) * N; return s->vq2;
    case SP_FIELD_VAZ: *bytes = sizeof(v2) * N; return s->vaz;
    case SP_FIELD_VTILES: *bytes = sizeof(uint32_t) * (size_t)(((s->W + 15) / 16) * ((s->H + 15) / 16)); return s->vtiles;
    case SP_FIELD_VTILE: *bytes = sizeof(uint32_t) * (size_t)s->lv.tiles; return s->vtile;
    case SP_FIELD_VMA: *bytes = sizeof(v4) * T; return s->vma;
    case SP_FIELD_VMB: *bytes = sizeof(v4) * T; return s->vmb;
    case SP_FIELD_VMD: *bytes = sizeof(v2) * T; return s->vmd;
    case SP_FIELD_VMR: *bytes = sizeof(v2) * T; return s->vmr;
    case SP_FIELD_VMX: *bytes = sizeof(v2) * T; return s->vmx;
    case SP_FIELD_VMZ: *bytes = sizeof(v2) * T; return s->vmz;
    case SP_FIELD_P2KEEP: *bytes = sizeof(float) * N; return s->p2keep;
    case SP_FIELD_THERM: *bytes = sizeof(v4) * N; return s->therm;
    case SP_FIELD_DHEAT: *bytes = sizeof(v4) * N; return s->dheat;
    case SP_FIELD_HCELL: *bytes = sizeof(int32_t) * 2 * N; return s->hcell;
    case SP_FIELD_SOLID_T: *bytes = sizeof(float) * N; return s->solid_t;
    case SP_FIELD_LEDGER: *bytes = sizeof(s->ledger); return s->ledger;
    case SP_FIELD_RL: *bytes = 8 * (size_t)N; return s->rl;
    case SP_FIELD_RBODY: *bytes = 4 * (size_t)SP_RIGID_WORDS; return s->rbody;
    case SP_FIELD_RH: *bytes = sizeof(v4) * (size_t)N; return s->rh;
    case SP_FIELD_ESTR: *bytes = sizeof(v4) * (size_t)N; return s->estr;
    case SP_FIELD_PSTRESS: *bytes = sizeof(v4) * (size_t)s->n_slots; return s->pstress;
    case SP_FIELD_FAMOCC: *bytes = sizeof(float) * N * (size_t)(s->P.n_fam > 1 ? s->P.n_fam : 1); return s->famocc;
    case SP_FIELD_MAGC: *bytes = sizeof(v2) * N; return s->magc;
    case SP_FIELD_RMASK: *bytes = sizeof(uint32_t) * N; return s->rmask;
    case SP_FIELD_MGX: *bytes = sizeof(v2) * T; return s->mgx;
    case SP_FIELD_DSETTLE: *bytes = sizeof(v2) * N * (size_t)(s->P.n_fam > 1 ? s->P.n_fam : 1); return s->dsettle;
    default: *bytes = 0; return NULL;
    }
}

long or_get(or_sim* s, int field, void* out, size_t bytes) {
    size_t n;
    void* p = field_ptr(s, field, &n);
    if (!p || bytes < n) return -1;
    memcpy(out, p, n);
    return (long)n;
}

long or_put(or_sim* s, int field, const void* in, size_t bytes) {
    size_t n;
    void* p = field_ptr(s, field, &n);
    if (field == SP_FIELD_PARTICLES) {
        n = bytes < sizeof(sp_particle) * s->cap ? bytes : sizeof(sp_particle) * s->cap;
        s->n_slots = (uint32_t)(n / sizeof(sp_particle));
        s->P.n_slots = s->n_slots;
    }
    if (!p || bytes < n) return -1;
    memcpy(p, in, n);
    return (long)n;
}

/* sort.comp with a defined order: by cell_of(pos), stable by slot; dead slots dropped to
 * the end (sort.comp:28-68). */
void or_sort(or_sim* s) {
    size_t N = (size_t)s->N;
    uint32_t* count = (uint32_t*)calloc(N + 1, sizeof(uint32_t));
    int* cellof = (int*)malloc(sizeof(int) * (s->n_slots ? s->n_slots : 1));
    for (uint32_t id = 0; id < s->n_slots; id++) {
        cellof[id] = -1;
        if (s->part[id].kind == SP_KIND_DEAD) continue;
        int x = clampi((int)s->part[id].pos[0], 0, s->W - 1), y = clampi((int)s->part[id].pos[1], 0, s->H - 1);
        cellof[id] = cidx(s, x, y);
        count[cellof[id] + 1]++;
    }
    for (size_t c = 0; c < N; c++) count[c + 1] += count[c];
    sp_particle* out = (sp_particle*)calloc(s->cap ? s->cap : 1, sizeof(sp_particle));
    v4* sout = (v4*)calloc(s->cap ? s->cap : 1, sizeof(v4));
    uint32_t live = count[N];
    /* the stress moves with its particle (sort.comp:66) */
    for (uint32_t id = 0; id < s->n_slots; id++)
        if (cellof[id] >= 0) {
            sout[count[cellof[id]]] = s->pstress[id];
            out[count[cellof[id]]++] = s->part[id];
        }
    /* dead slots keep their (dead) contents at the end */
    uint32_t k = live;
    for (uint32_t id = 0; id < s->n_slots; id++)
        if (cellof[id] < 0) {
            sout[k] = s->pstress[id];
            out[k++] = s->part[id];
        }
    memcpy(s->part, out, sizeof(sp_particle) * s->n_slots);
    if (s->P.caps & SP_CAP_ELASTIC) memcpy(s->pstress, sout, sizeof(v4) * s->n_slots);
    free(out);
    free(sout);
    free(cellof);
    free(count);
}

/* The spawn step of record (mod.rs:2485-2503) and claim.comp with a defined order: pending
 * particles fill fresh slots first; the rest go to dead slots, the k-th dead slot in slot
 * order taking the k-th remaining particle, while there are any (claim.comp:10-17; the
 * reference's order is that of its atomics). Particles beyond the dead slots are dropped. */
void or_spawn(or_sim* s, const sp_particle* spawn, uint32_t n) {
    uint32_t fresh = s->cap - s->n_slots;
    if (fresh > n) fresh = n;
    for (uint32_t k = 0; k < fresh; k++) s->part[s->n_slots + k] = spawn[k];
    uint32_t old = s->n_slots;
    s->n_slots += fresh;
    s->P.n_slots = s->n_slots;
    uint32_t k = fresh;
    for (uint32_t id = 0; id < old && k < n; id++)
        if (s->part[id].kind == SP_KIND_DEAD) s->part[id] = spawn[k++];
}

void or_cell_counts(or_sim* s, int32_t x0, int32_t y0, int32_t w, int32_t h, uint32_t* out) {
    memset(out, 0, sizeof(uint32_t) * (size_t)(w > 0 && h > 0 ? w * h : 0));
    for (uint32_t id = 0; id < s->n_slots; id++) {
        if ((s->part[id].kind & 255u) == 0u) continue;
        int x = (int)floorf(s->part[id].pos[0]) - x0, y = (int)floorf(s->part[id].pos[1]) - y0;
        if (x < 0 || y < 0 || x >= w || y >= h) continue;
        out[y * w + x]++;
    }
}

void or_stats(or_sim* s, uint32_t* live, uint32_t* count, uint32_t* conc256) {
    *live = 0;
    memset(count, 0, sizeof(uint32_t) * SP_MAX_MATERIALS);
    memset(conc256, 0, sizeof(uint32_t) * SP_MAX_MATERIALS);
    for (uint32_t id = 0; id < s->n_slots; id++) {
        uint32_t kind = s->part[id].kind & 255u;
        if (kind == 0u) continue;
        (*live)++;
        count[kind - 1]++;
        conc256[kind - 1] += (uint32_t)(clampf(s->part[id].s, 0.0f, 1.0f) * 256.0f + 0.5f);
    }
}
