// dev.h — device-side helpers of the MI355X fluid step (gfx950, wave64).
//
// Float arithmetic follows the reference shaders' operation order and is compiled with
// -ffp-contract=off, so that quantities the reference makes order-independent (the
// fixed-point scatter sums, cell indices) come out bit-identical to a CPU evaluation of the
// same inputs. GLSL built-ins are spelled out with the semantics used throughout this
// build: mix(x, y, a) = x (1 - a) + y a, length = sqrt(x^2 + y^2), normalize = v / length.
// Square roots use sqrtf, which hipcc lowers to the correctly rounded llvm.sqrt.f32
// (__fsqrt_rn carries !fpmath 3.0, a 3-ulp approximation); '/' is correctly rounded.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/sprout_types.h"

#define SP_DEV static __device__ __forceinline__

// Device pointers of one tank, rebuilt at every kernel's entry from its TankRef.
struct Ctx {
    int W, H, N;
    const sp_gparams* P;
    const sp_gmat* mats;
    sp_stepbuf* S;
    sp_solve* K;
    sp_particle* part;
    int32_t* acc;
    float4 *vel, *cell, *comp, *coef, *surf, *surf2, *sprop, *aux, *grho, *gbase, *prhs, *vw, *vs;
    const float4* solid;
    float2 *vel0, *pjump, *pres;
    const float* sdf;
    float *mixb, *p2keep;
    const uint32_t* region;
    uint32_t* region_air;
    float4* mga;
    float* mgd;
    float2 *mgx, *mgz, *mgr, *cgp, *cgp2, *cgq;
    float2* r2;           // second level-0 residual buffer of the CG (alternates with mgr)
    uint32_t* ptile;
    double4* part_sums;   // per-block partial sums of the pressure CG's reductions (three kinds)
    const int4* lev;      // per level: offset, width, height, tile offset
    int n_levels, tail;
    // implicit viscosity (per cell of the grid; the multigrid per cell of every level)
    float2 *vr, *vp, *vp2, *vq, *vd, *vr2, *vq2, *vaz;
    uint32_t* vtiles;     // active viscosity tiles of the grid
    uint32_t* vtile;      // tile flags of the viscosity multigrid, per level
    float4 *vma, *vmb;
    float2 *vmd, *vmr, *vmx, *vmz;
    double4* vpart;       // per-tile partial sums of the viscosity solve
    // heat
    float4 *therm, *dheat;
    int2* hcell;          // per cell: heat capacity and heat sums of the particles in it
    float* solid_t;       // temperature a solid is held at (SP_NO_TEMP: insulating)
    uint32_t* ledger;     // SP_LEDGER_WORDS words: energy entries as 64-bit fixed point
    // rigid bodies (rigid.comp)
    int2* rl;             // per cell: root cell of its body (-1: not rigid); at a root, its slot
    uint32_t* rbody;      // SP_RIGID_WORDS: bodies claimed, then 64-bit sums per slot
    float4* rh;           // per cell: non-Newtonian share, its row, its concentration; w = 1 where stiff
    // viscoelasticity (elastic.comp)
    float4* estr;         // per cell: stress xx, yy, xy (cells^2/s^2), scattered by volume
    float4* pstress;      // per particle slot: its stress xx, yy, xy; moved with the particle by the sort
    // settling of suspensions (settle.comp), per cell and concentration family (cell * n_fam + f)
    float* famocc;        // the family's occupancy of the cell
    float2* dsettle;      // this substep's gain and loss factors of the family's concentration
    // magnetism (magnet.comp): per cell the susceptibility of its liquid (grid_prep) and the
    // magnetic energy density (the magnet pass, read by the next substep's grid_prep too)
    float2* magc;
    // reactions: per cell, bit j where the tank's j-th row makes up more than REACT_MIN
    uint32_t* rmask;
};

// A tank as a kernel argument: its arena (one allocation holding every buffer of the tank)
// and the sizes the buffers' offsets follow from. Kernels take the tanks of their launch by
// value in the kernel-argument segment and index them by blockIdx.z; pointers derived from
// a kernel argument are known to be global memory (global and scalar loads), which pointers
// read from ordinary memory are not.
struct TankRef {
    char* base;
    int W, H;
    uint32_t cap;        // particle slots
    int total;           // cells of every multigrid level
    int tiles;           // 16 x 16 tile flags of levels 0 .. tail
    int n_levels, tail;
};

#define SP_LAUNCH_TANKS 32  // most tanks of one launch (32 x 40 B of kernel arguments)
struct TankBatch {
    TankRef t[SP_LAUNCH_TANKS];
};

// The arena's layout: every buffer at a 256-byte aligned offset, in a fixed order. With
// base = nullptr, the returned `bytes` is the arena's size.
__host__ __device__ __forceinline__ Ctx ctx_of(const TankRef& r, size_t* bytes = nullptr) {
    Ctx c;
    size_t o = 0;
    auto take = [&](size_t n) {
        char* p = r.base + o;
        o += (n + 255) & ~(size_t)255;
        return p;
    };
    const size_t N = (size_t)r.W * r.H, T = (size_t)r.total;
    const size_t nb0 = (size_t)((r.W + 15) / 16) * ((r.H + 15) / 16);
    c.W = r.W;
    c.H = r.H;
    c.N = (int)N;
    c.n_levels = r.n_levels;
    c.tail = r.tail;
    c.P = (const sp_gparams*)take(sizeof(sp_gparams));
    c.mats = (const sp_gmat*)take(sizeof(sp_gmat) * SP_MAX_MATERIALS);
    c.S = (sp_stepbuf*)take(sizeof(sp_stepbuf));
    c.K = (sp_solve*)take(sizeof(sp_solve));
    c.lev = (const int4*)take(sizeof(int4) * SP_MAX_LEVELS);
    c.region_air = (uint32_t*)take(sizeof(uint32_t));
    c.part = (sp_particle*)take(sizeof(sp_particle) * (r.cap ? r.cap : 1));
    c.acc = (int32_t*)take(sizeof(int32_t) * N * SP_ACC_MAX);
    float4** f4[] = {&c.vel, &c.cell, &c.comp, &c.coef, &c.surf, &c.surf2, &c.sprop, &c.aux, &c.grho,
                     &c.gbase, &c.prhs, &c.vw, &c.vs};
    for (float4** f : f4) *f = (float4*)take(sizeof(float4) * N);
    c.solid = (const float4*)take(sizeof(float4) * N);
    float2** f2[] = {&c.vel0, &c.pjump, &c.pres, &c.cgp, &c.cgp2, &c.cgq, &c.r2, &c.vr, &c.vp, &c.vp2, &c.vq,
                     &c.vd, &c.vr2, &c.vq2, &c.vaz};
    for (float2** f : f2) *f = (float2*)take(sizeof(float2) * N);
    c.sdf = (const float*)take(sizeof(float) * (size_t)(r.W + 1) * (r.H + 1));
    c.mixb = (float*)take(sizeof(float) * N * SP_MAX_PRESENT);
    c.p2keep = (float*)take(sizeof(float) * N);
    c.region = (const uint32_t*)take(sizeof(uint32_t) * N);
    c.mga = (float4*)take(sizeof(float4) * T);
    c.mgd = (float*)take(sizeof(float) * T);
    float2** t2[] = {&c.mgx, &c.mgz, &c.mgr, &c.vmd, &c.vmr, &c.vmx, &c.vmz};
    for (float2** f : t2) *f = (float2*)take(sizeof(float2) * T);
    c.vma = (float4*)take(sizeof(float4) * T);
    c.vmb = (float4*)take(sizeof(float4) * T);
    c.ptile = (uint32_t*)take(sizeof(uint32_t) * (size_t)r.tiles);
    c.vtile = (uint32_t*)take(sizeof(uint32_t) * (size_t)r.tiles);
    c.vtiles = (uint32_t*)take(sizeof(uint32_t) * nb0);
    c.part_sums = (double4*)take(sizeof(double4) * 3 * nb0);
    c.vpart = (double4*)take(sizeof(double4) * nb0);
    c.therm = (float4*)take(sizeof(float4) * N);
    c.dheat = (float4*)take(sizeof(float4) * N);
    c.hcell = (int2*)take(sizeof(int2) * N);
    c.solid_t = (float*)take(sizeof(float) * N);
    c.ledger = (uint32_t*)take(sizeof(uint32_t) * SP_LEDGER_WORDS);
    c.rl = (int2*)take(sizeof(int2) * N);
    c.rbody = (uint32_t*)take(sizeof(uint32_t) * SP_RIGID_WORDS);
    c.rh = (float4*)take(sizeof(float4) * N);
    c.estr = (float4*)take(sizeof(float4) * N);
    c.pstress = (float4*)take(sizeof(float4) * (r.cap ? r.cap : 1));
    c.famocc = (float*)take(sizeof(float) * N * SP_MAX_PRESENT);
    c.dsettle = (float2*)take(sizeof(float2) * N * SP_MAX_PRESENT);
    c.magc = (float2*)take(sizeof(float2) * N);
    c.rmask = (uint32_t*)take(sizeof(uint32_t) * N);
    if (bytes) *bytes = o;
    return c;
}

SP_DEV float fmaxf_(float a, float b) { return a > b ? a : b; }
SP_DEV float fminf_(float a, float b) { return a < b ? a : b; }
SP_DEV float clampf_(float x, float lo, float hi) { return fminf_(fmaxf_(x, lo), hi); }
SP_DEV int clampi_(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
SP_DEV float mixf_(float x, float y, float a) { return x * (1.0f - a) + y * a; }
SP_DEV float len2_(float x, float y) { return sqrtf(x * x + y * y); }
SP_DEV float signf_(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
SP_DEV bool face_live(float s) { return fabsf(s) < 1e29f; }
SP_DEV bool face_closed(float s) { return s < -1e29f; }
SP_DEV float f4c(const float4& a, int k) { return k == 0 ? a.x : k == 1 ? a.y : k == 2 ? a.z : a.w; }

// Bilinear SDF and its gradient (air.glsl:13-22; rigid.comp:50-61).
SP_DEV float sample_sdf(const Ctx& c, float px, float py, float& gx, float& gy) {
    float x = clampf_(px, 0.0f, (float)c.W - 0.001f), y = clampf_(py, 0.0f, (float)c.H - 0.001f);
    int ix = (int)x, iy = (int)y;
    float fx = x - (float)ix, fy = y - (float)iy;
    int w1 = c.W + 1;
    int k = iy * w1 + ix;
    float d00 = c.sdf[k], d10 = c.sdf[k + 1], d01 = c.sdf[k + w1], d11 = c.sdf[k + w1 + 1];
    gx = mixf_(d10 - d00, d11 - d01, fy);
    gy = mixf_(d01 - d00, d11 - d10, fx);
    return mixf_(mixf_(d00, d10, fx), mixf_(d01, d11, fx), fy);
}

// ---- rheology (common.glsl:527-571) --------------------------------------------------------

// Shear rate sqrt(2 D:D) of cell i from the live faces around it; `pre`: from the velocity
// before the solve, which has not had gravity's step yet on v faces (shear_rate,
// common.glsl:527-540).
SP_DEV float shear_rate(const Ctx& c, int i, bool pre) {
    const int W = c.W;
    const float gv = pre ? c.P->gravity * c.S->dt : 0.0f;
    auto UF = [&](int k) { const float4 q = c.vel[k]; return face_live(q.z) ? q.x : 0.0f; };
    auto VF = [&](int k) { const float4 q = c.vel[k]; return face_live(q.w) ? q.y + gv : 0.0f; };
    const float dxx = UF(i + 1) - UF(i);
    const float dyy = VF(i + W) - VF(i);
    const float duy = 0.25f * (UF(i + W) + UF(i + 1 + W) - UF(i - W) - UF(i + 1 - W));
    const float dvx = 0.25f * (VF(i + 1) + VF(i + 1 + W) - VF(i - 1) - VF(i - 1 + W));
    const float dxy = 0.5f * (duy + dvx);
    return sqrtf(2.0f * (dxx * dxx + dyy * dyy + 2.0f * dxy * dxy));
}

// Non-Newtonian response of cell i (q = rh[i]) at shear rate g: the factor F on the
// Newtonian viscosity and the yield stress, both for the cell's share of the material
// (nn_eval, common.glsl:546-571).
SP_DEV void nn_eval(const Ctx& c, int i, float4 q, float g, float& F, float& tau) {
    const int r = (int)q.y;
    const sp_gmat& m = c.mats[r];
    const float n = m.rheo_p[1], lam = m.rheo_p[2], ratio = m.rheo_p[3];
    F = 1.0f;
    tau = 0.0f;
    if (m.rheo == 1) {
        F = ratio + (1.0f - ratio) * powf(1.0f + (lam * g) * (lam * g), 0.5f * (n - 1.0f));
    } else if (m.rheo == 2) {
        F = powf(fmaxf_(g, 1e-6f), n - 1.0f);
        tau = m.rheo_p[0] * q.z;
    } else if (m.rheo == 3) {
        F = fminf_(powf(1.0f + (lam * g) * (lam * g), 0.5f * (n - 1.0f)), ratio);
    }
    const float fr = m.gr[0];
    if (fr > 0.0f) tau += fr * q.z * q.z * q.z * fmaxf_(c.pres[i].x, 0.0f) / fmaxf_(c.S->dt, 1e-6f);
    const int W = c.W;
    const float skin = m.temp[3];
    if (skin > 0.0f && (c.cell[i - 1].w <= 0.0f || c.cell[i + 1].w <= 0.0f || c.cell[i - W].w <= 0.0f || c.cell[i + W].w <= 0.0f))
        tau += skin;
    F = 1.0f + q.x * (F - 1.0f);
    tau *= q.x;
}

// ---- rigid bodies (common.glsl:263-340) ---------------------------------------------------


// The 64-bit fixed-point sum at words w (low), w + 1 (high) of the body table, as a float
// (get64, common.glsl:277-279).
SP_DEV float get64(const Ctx& c, uint32_t w) {
    return ((float)(int32_t)c.rbody[w + 1] * 4294967296.0f + (float)c.rbody[w]) / SP_RIGID_FIX;
}

// The body's support on the solids it touches (body_contact, common.glsl:290-324): motion into
// the support is stopped, turning limited to what keeps its ends out of the solid (or tipping
// about its edge where the centre of mass overhangs), and Coulomb friction along it bounded
// by the normal speed taken away. Positions relative to the root cell.
SP_DEV void body_contact(const Ctx& c, uint32_t o, float2 cm, float2& U, float& w) {
    const float mu = c.P->rigid_friction;
    const float nc = get64(c, o + 14u);
    if (mu < 0.0f || nc < 0.5f) return;
    const float nsx = get64(c, o + 16u), nsy = get64(c, o + 18u);
    const float nsl = len2_(nsx, nsy);
    if (nsl < 0.25f * nc) return;  // contacts on opposite sides: wedged, left to the solve
    const float nx = nsx / nsl, ny = nsy / nsl;
    const float tx = ny, ty = -nx;
    const float ccx = get64(c, o + 20u) / nc, ccy = get64(c, o + 22u) / nc;
    const float sqx = get64(c, o + 24u) / nc - ccx * ccx;
    const float sqy = get64(c, o + 26u) / nc - ccy * ccy;
    const float sqz = get64(c, o + 28u) / nc - ccx * ccy;
    const float ext = sqrtf(3.0f * fmaxf_(tx * tx * sqx + ty * ty * sqy + 2.0f * tx * ty * sqz, 0.0f)) + 0.5f;
    const float rcx = ccx - cm.x, rcy = ccy - cm.y;
    float ucx = U.x + w * -rcy, ucy = U.y + w * rcx;
    const float vn = ucx * nx + ucy * ny;
    const float s = (cm.x - ccx) * tx + (cm.y - ccy) * ty;
    float vn1 = vn, w1 = w;
    if (fabsf(s) <= ext) {
        vn1 = fmaxf_(vn, 0.0f);
        w1 = clampf_(w, -vn1 / ext, vn1 / ext);
    } else {
        const float edge = signf_(s) * ext;
        vn1 = fmaxf_(vn + w * edge, 0.0f) - w * edge;
    }
    if (vn1 == vn && w1 == w) return;
    const float vt = ucx * tx + ucy * ty;
    const float vt1 = signf_(vt) * fmaxf_(fabsf(vt) - mu * (vn1 - vn), 0.0f);
    ucx += (vn1 - vn) * nx + (vt1 - vt) * tx;
    ucy += (vn1 - vn) * ny + (vt1 - vt) * ty;
    w = w1;
    U.x = ucx - w * -rcy;
    U.y = ucy - w * rcx;
}

// Rigid motion of body slot b (root cell r) at x (cells): (u, v, angular velocity). False if
// the body is too light to have one (body_motion, common.glsl:326-340).
SP_DEV bool body_motion(const Ctx& c, uint32_t b, int r, float px, float py, float3& u) {
    const uint32_t o = 16u + (uint32_t)SP_BODY_WORDS * b;
    const float M = get64(c, o);
    if (M <= 1e-3f) return false;
    const float2 cm = make_float2(get64(c, o + 2u) / M, get64(c, o + 4u) / M);
    float2 U = make_float2(get64(c, o + 6u) / M, get64(c, o + 8u) / M);
    const float L = get64(c, o + 10u) - M * (cm.x * U.y - cm.y * U.x);
    const float I = get64(c, o + 12u) - M * (cm.x * cm.x + cm.y * cm.y);
    float w = I > 1e-3f ? L / I : 0.0f;
    body_contact(c, o, cm, U, w);
    const float dx = px - ((float)(r % c.W) + 0.5f) - cm.x;
    const float dy = py - ((float)(r / c.W) + 0.5f) - cm.y;
    u = make_float3(U.x + w * -dy, U.y + w * dx, w);
    return true;
}
SP_DEV float smoothstep_(float e0, float e1, float x) {
    float t = clampf_((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// Integer hash of the per-particle random numbers (air.glsl:4-11).
SP_DEV uint32_t hash_u(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// Quadratic B-spline weights at fractional offset f in [0.5, 1.5) (common.glsl:638-643).
SP_DEV void bspline3(float f, float w[3]) {
    w[0] = 0.5f * (1.5f - f) * (1.5f - f);
    w[1] = 0.75f - (f - 1.0f) * (f - 1.0f);
    w[2] = 0.5f * (f - 0.5f) * (f - 0.5f);
}

SP_DEV int floori(float x) { return (int)floorf(x); }

// ---- heat (common.glsl:408-453, 485-530, 579-606; particles.comp:67-80) -----------------

// ln of the viscosity factor of row r at T, deg C (temp_law, common.glsl:410-417).
SP_DEV float temp_law(const Ctx& c, int r, float T) {
    const sp_gmat& m = c.mats[r];
    if (m.temp_model == 0) return 0.0f;
    const float Tk = fmaxf_(T + 273.15f, 1.0f);
    const float x = m.temp_model == 1 ? m.temp[0] * (1.0f / Tk - 1.0f / m.temp[2])
                                      : m.temp[0] / fmaxf_(Tk - m.temp[1], 5.0f) - m.temp[0] / fmaxf_(m.temp[2] - m.temp[1], 5.0f);
    return clampf_(x, -30.0f, 30.0f);
}

// Density factor of row r at concentration s and T (expansion, common.glsl:485-494).
SP_DEV float expansion(const Ctx& c, int row, float s, float T) {
    const sp_gmat& m = c.mats[row];
    float b = m.expansion, tr = m.temp[2] - 273.15f, tb = m.th2[2];
    if (m.diluent >= 0) {
        const sp_gmat& d = c.mats[m.diluent];
        b = mixf_(d.expansion, b, s);
        tr = mixf_(d.temp[2] - 273.15f, tr, s);
        tb = mixf_(d.th2[2], tb, s);
    }
    return clampf_(1.0f - b * (fminf_(T, tb) - tr), 0.7f, 1.3f);
}

// Solidus, liquidus, latent heat (K) and heat capacity of row at concentration s
// (thermal_of, common.glsl:500-509).
SP_DEV float4 thermal_of(const Ctx& c, int row, float s) {
    const sp_gmat& m = c.mats[row];
    const float ld = m.th2[0];
    if (m.diluent < 0) return make_float4(m.th[2], m.th[3], ld, m.th[0]);
    const sp_gmat& d = c.mats[m.diluent];
    const float shift = m.freeze_shift * s;
    const float C = mixf_(d.th[0], m.th[0], s);
    return make_float4(d.th[2] + shift, d.th[3] + shift, mixf_(d.th[0] * d.th2[0], m.th[0] * ld, s) / C, C);
}

// Temperature and solid share of enthalpy temperature te (phase_of, common.glsl:515-521).
SP_DEV float2 phase_of(float te, float4 m) {
    if (te >= m.y) return make_float2(te, 0.0f);
    const float lo = m.x - m.z;
    if (te <= lo) return make_float2(te + m.z, 1.0f);
    const float fr = (te - lo) / (m.y - lo);
    return make_float2(m.x + fr * (m.y - m.x), 1.0f - fr);
}
// ln kinematic viscosity (cells^2/s) of material `row` at concentration s and temperature T
// (row_lnu, common.glsl:631-636; an isothermal tank's table has the temperature folded in).
SP_DEV float row_lnu(const Ctx& c, int row, float s, float T, bool thermal) {
    const sp_gmat& m = c.mats[row];
    float l = m.ln_nu + (thermal ? temp_law(c, row, T) : 0.0f);
    const int d = m.diluent;
    if (d >= 0) {
        const float bs = m.visc_exp == 1.0f ? s : powf(s, m.visc_exp);
        l = mixf_(c.mats[d].ln_nu + (thermal ? temp_law(c, d, T) : 0.0f), l, bs);
    }
    return l;
}

// Adds v >= 0 to the ledger's 64-bit fixed-point energy entry e (ledger_add, common.glsl:
// 446-451: the quantum floor(v LEDGER_SCALE + 0.5) with a carry into the high word).
SP_DEV void ledger_add(const Ctx& c, int e, float v) {
    const uint64_t q = (uint64_t)(uint32_t)(v * SP_LEDGER_SCALE + 0.5f);
    atomicAdd(reinterpret_cast<unsigned long long*>(c.ledger + 2 * e), (unsigned long long)q);
}

// Terminal speed of a sphere of diameter d (cells) and density difference drho through a
// liquid of density rho_c and viscosity nu_c, Stokes to Newton drag (common.glsl:370-374).
SP_DEV float terminal(const Ctx& c, float drho, float rho_c, float nu_c, float d) {
    float qa = 0.44f * rho_c, qb = 24.0f * nu_c * rho_c / d, qc = 4.0f / 3.0f * fabsf(drho) * c.P->gravity * d;
    return qc > 0.0f ? 2.0f * qc / (qb + sqrtf(qb * qb + 4.0f * qa * qc)) : 0.0f;
}

// A particle's heat capacity and heat in fixed point (heat_quanta, particles.comp:70-73).
SP_DEV int2 heat_quanta(float C, float T) {
    const int ci = (int)floorf(C * SP_MASS_SCALE + 0.5f);
    return make_int2(ci, (int)floorf((float)ci * (T + 273.15f) / SP_T_SCALE + 0.5f));
}

// The enthalpy temperature that keeps the heat content when the heat capacity changes
// (keep_heat, particles.comp:78-80).
SP_DEV float keep_heat(float t, float c0, float c1) { return c0 == c1 ? t : (t + 273.15f) * c0 / c1 - 273.15f; }

// Boiling point of row at concentration s (particles.comp:368, 957).
SP_DEV float boil_of(const Ctx& c, int row, float s) {
    const sp_gmat& m = c.mats[row];
    return m.diluent >= 0 ? mixf_(c.mats[m.diluent].th2[2], m.th2[2], s) : m.th2[2];
}

// Accumulator channel of phase group g's own temperature sum, -1 none (common.glsl:582-584).
SP_DEV int group_heat_chan(const sp_gparams& P, int g) { return (int)((P.group_heat >> (8u * (uint32_t)g)) & 255u) - 1; }

// Wave64 reductions.
SP_DEV float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf_(v, __shfl_xor(v, o, 64));
    return v;
}
SP_DEV int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}
SP_DEV int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}

// atomicMax on the bits of a non-negative float.
SP_DEV void atomic_max_pos_float(uint32_t* addr, float v) {
    if (v > 0.0f) atomicMax(addr, __float_as_uint(v));
}

// Work-group maximum of a non-negative float folded into *addr (float bits) with one
// atomic per group, and only when the group's maximum exceeds what *addr already holds:
// one atomic per wave on a single word serialises thousands of groups. `lds` holds one
// float per wave64 of the group; every work-item of the group must call it.
template <int NT>
SP_DEV void group_max_to(uint32_t* addr, float v, float* lds) {
    const int t = threadIdx.x + threadIdx.y * blockDim.x;
    v = wave_max(v);
    if ((t & 63) == 0) lds[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        float m = lds[0];
        for (int k = 1; k < NT / 64; k++) m = fmaxf_(m, lds[k]);
        if (m > 0.0f && m > __uint_as_float(__hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
            atomicMax(addr, __float_as_uint(m));
    }
}

// Whether phase group g sorts itself by density below the grid scale (common.glsl:590-602).
SP_DEV bool group_sorts(const sp_gparams& P, const sp_gmat* mats, int g) {
    if (g <= 0) return false;
    int n = 0;
    for (int j = 0; j < P.n_present; j++) {
        int r = P.present[j];
        const sp_gmat& m = mats[r];
        if (m.diluent >= 0 && mats[m.diluent].group == g) return false;
        if (m.group != g) continue;
        if (m.diluent >= 0) return false;
        n++;
    }
    return n >= 2;
}
