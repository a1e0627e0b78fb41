// particles.hip — the per-particle pass of a substep on gfx950, and the particle
// bookkeeping kernels (count, stats); the sort and the claim are in bin.hip.
//
// One work-item per particle slot, 256 per work-group (four wave64s). Each work-item runs
// the gather (grid -> particle, APIC), the advection with collision against the solid SDF,
// and then the scatter (particle -> grid) of mass, momentum and composition as 32-bit fixed
// point (particles.comp:98-1101). The scatter goes to an LDS box that spans the cells the
// group's particles touch when it fits (particles are kept sorted by cell, so it usually
// does), and is flushed to HBM with one atomic per non-zero word; otherwise each node is
// added with a global atomic. Integer sums are order-independent, so the result is the same
// either way and bit-identical to a serial evaluation.
#include "dev.h"

namespace {


struct ScatterIn {
    bool live;
    float px, py, vsx, vsy, Csx, Csy, Csz, Csw, rho, sc, CT;
    int cw, cs, ct, cg;
    int ce;  // first stress channel, -1: no stress to scatter (the stress is then in the slot)
};

// The stress a particle scatters, over el_g in STRESS_FIX units (particles.comp:984). With
// g.ce >= 0 the particle's slot holds it: the gather wrote it, or, on a substep without a
// gather, the particle read it from there.
SP_DEV float3 scatter_stress(const Ctx& c, const ScatterIn& g, uint32_t id) {
    if (g.ce < 0) return make_float3(0.0f, 0.0f, 0.0f);
    const float4 t = c.pstress[id];
    const float e = c.P->el_g;
    return make_float3(t.x / e * SP_STRESS_FIX, t.y / e * SP_STRESS_FIX, t.z / e * SP_STRESS_FIX);
}

SP_DEV int cidx(const Ctx& c, int x, int y) { return y * c.W + x; }

// Relative density of a row at concentration s (common.glsl:475-479).
SP_DEV float row_density(const Ctx& c, int row, float s) {
    int d = c.mats[row].diluent;
    float rho = c.mats[row].rho;
    return d >= 0 ? mixf_(c.mats[d].rho, rho, s) : rho;
}

// Bilinear velocity over live faces, `fb` where no face is live (air.glsl:41-78).
SP_DEV float2 sample_vel(const Ctx& c, float px, float py, float2 fb) {
    float2 o = fb;
    {
        float gx = px, gy = py - 0.5f;
        int bx = floori(gx), by = floori(gy);
        float fx = gx - (float)bx, fy = gy - (float)by;
        float sw = 0.0f, su = 0.0f;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int a = 0; a < 2; a++) {
                float4 q = c.vel[cidx(c, bx + a, by + b)];
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
        int bx = floori(gx), by = floori(gy);
        float fx = gx - (float)bx, fy = gy - (float)by;
        float sw = 0.0f, su = 0.0f;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int a = 0; a < 2; a++) {
                float4 q = c.vel[cidx(c, bx + a, by + b)];
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

// One push out of the solids (particles.comp:51-65).
SP_DEV bool collide(const Ctx& c, float& x, float& y, float& vx, float& vy, bool beyond) {
    float nx, ny;
    float d = sample_sdf(c, x, y, nx, ny);
    float wm = c.P->wall_margin;
    if (d < wm) {
        float l = len2_(nx, ny);
        if (l > 1e-5f) {
            nx = nx / l;
            ny = ny / l;
            float k = wm - d + (beyond ? fmaxf_(-d, 0.0f) : 0.0f);
            x += nx * k;
            y += ny * k;
            float vn = vx * nx + vy * ny;
            if (vn < 0.0f) {
                vx -= vn * nx;
                vy -= vn * ny;
            }
            return true;
        }
    }
    return false;
}

// The upper-convected Maxwell stress of one particle over a substep (particles.comp:249-279):
// dtau/dt = L tau + tau L^T - tau / lambda + G (L + L^T), L the velocity gradient (the APIC
// matrix C = du/dx, du/dy, dv/dx, dv/dy) with at most el_step_max deformation per substep,
// relaxation implicit, the norm capped at el_stress_max G; no stress without a cell's worth
// of liquid in the 3 x 3 cells around the particle. tau = (xx, yy, xy, 0).
SP_DEV float4 stress_update(const Ctx& c, float px, float py, float C0, float C1, float C2, float C3, float4 tau, float G,
                            float lam, float dt) {
    const sp_gparams& P = *c.P;
    float a = C0, b = C1, cc = C2, d = C3;
    const float lnorm = sqrtf(((C0 * C0 + C1 * C1) + C2 * C2) + C3 * C3) * dt;
    if (P.el_step_max > 0.0f && lnorm > P.el_step_max) {
        const float k = P.el_step_max / lnorm;
        a *= k;
        b *= k;
        cc *= k;
        d *= k;
    }
    const float X = tau.x, Y = tau.y, Z = tau.z;
    // L T; T L^T is its transpose (T is symmetric)
    const float lt00 = a * X + b * Z, lt01 = a * Z + b * Y, lt10 = cc * X + d * Z, lt11 = cc * Z + d * Y;
    const float D00 = (lt00 + lt00) + G * (a + a);
    const float D11 = (lt11 + lt11) + G * (d + d);
    const float D01 = (lt01 + lt10) + G * (b + cc);
    const float D10 = (lt10 + lt01) + G * (cc + b);
    const float k = 1.0f + dt / fmaxf_(lam, 1e-4f);
    const float T00 = (X + dt * D00) / k, T11 = (Y + dt * D11) / k;
    const float T01 = (Z + dt * D01) / k, T10 = (Z + dt * D10) / k;
    float4 t = make_float4(T00, T11, 0.5f * (T10 + T01), 0.0f);
    const float tn = sqrtf((t.x * t.x + t.y * t.y) + t.z * t.z);
    if (P.el_stress_max > 0.0f && tn > P.el_stress_max * G) {
        const float s = P.el_stress_max * G / tn;
        t.x *= s;
        t.y *= s;
        t.z *= s;
        t.w *= s;
    }
    const int x1 = clampi_((int)px, 1, c.W - 2), y1 = clampi_((int)py, 1, c.H - 2);
    float around = 0.0f;
#pragma unroll
    for (int bb = -1; bb <= 1; bb++)
#pragma unroll
        for (int aa = -1; aa <= 1; aa++) around += fminf_(c.cell[cidx(c, x1 + aa, y1 + bb)].x, 1.0f);
    if (around < 1.0f) t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return t;
}

// ---- the mouse tool (air.glsl:24-37; particles.comp:594-631) -----------------------------

// The velocity the tool drives a particle towards, (dx, dy) from the particle to the cursor:
// grab (1) carries liquid with the cursor and pulls it in; push / pull (3) drives it radially
// at tool_vel.x with a smooth centre; the nozzle (4) holds it at the launch velocity; whirl
// (7) turns it round the cursor, rigidly inside 0.4 of the radius, with a gentle pull in;
// hold (8) brings it to rest.
SP_DEV float2 tool_target(const sp_gparams& P, float dx, float dy) {
    const float l = len2_(dx, dy);
    if (P.tool_mode == 3) {
        const float m = fmaxf_(l, 0.15f * P.tool_radius);
        return make_float2(-dx / m * P.tool_vel[0], -dy / m * P.tool_vel[0]);
    }
    if (P.tool_mode == 4) return make_float2(P.tool_vel[0], P.tool_vel[1]);
    if (P.tool_mode == 7) {
        const float m = fmaxf_(l, 0.4f * P.tool_radius);
        return make_float2(dy / m * P.tool_vel[0] + dx * 2.0f, -dx / m * P.tool_vel[0] + dy * 2.0f);
    }
    if (P.tool_mode == 8) return make_float2(0.0f, 0.0f);
    return make_float2(P.tool_vel[0] + dx * 4.0f, P.tool_vel[1] + dy * 4.0f);
}

// ---- the air above the liquid (air.glsl:91-175) ----------------------------------------
// Air is not simulated: escaped vapour rides a procedural flow, a buoyant rise (steam.y) plus
// the curl of a stream function of drifting value noise (steam2.z: its speed, steam2.w: its
// size), ramped to zero towards solids so that it flows along them.

constexpr float PUFF_FOLLOW = 0.08f;  // s: how quickly a puff takes up the air's velocity

// Value noise in [-0.5, 0.5) at x and its gradient (air.glsl:108-128): a hash per lattice
// corner, quintic fade.
SP_DEV float4 vnoise3(float x, float y, float z) {
    const float ix = floorf(x), iy = floorf(y), iz = floorf(z);
    const float fx = x - ix, fy = y - iy, fz = z - iz;
    const float ux = fx * fx * fx * (fx * (fx * 6.0f - 15.0f) + 10.0f);
    const float uy = fy * fy * fy * (fy * (fy * 6.0f - 15.0f) + 10.0f);
    const float uz = fz * fz * fz * (fz * (fz * 6.0f - 15.0f) + 10.0f);
    const float dux = 30.0f * fx * fx * (fx * (fx - 2.0f) + 1.0f);
    const float duy = 30.0f * fy * fy * (fy * (fy - 2.0f) + 1.0f);
    const float duz = 30.0f * fz * fz * (fz * (fz - 2.0f) + 1.0f);
    const uint32_t cx = (uint32_t)((int)ix + 4096), cy = (uint32_t)((int)iy + 4096), cz = (uint32_t)((int)iz + 4096);
    float r[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const uint32_t qx = cx + (uint32_t)(k & 1), qy = cy + (uint32_t)((k >> 1) & 1), qz = cz + (uint32_t)((k >> 2) & 1);
        r[k] = (float)(hash_u(qx * 73856093u ^ qy * 19349663u ^ qz * 83492791u) >> 8) / 16777216.0f;
    }
    const float k0 = r[0], k1 = r[1] - r[0], k2 = r[2] - r[0], k3 = r[4] - r[0];
    const float k4 = r[0] - r[1] - r[2] + r[3], k5 = r[0] - r[2] - r[4] + r[6], k6 = r[0] - r[1] - r[4] + r[5];
    const float k7 = -r[0] + r[1] + r[2] - r[3] + r[4] - r[5] - r[6] + r[7];
    const float n = k0 + k1 * ux + k2 * uy + k3 * uz + k4 * ux * uy + k5 * uy * uz + k6 * uz * ux + k7 * ux * uy * uz;
    return make_float4(n - 0.5f, dux * (k1 + k4 * uy + k6 * uz + k7 * uy * uz), duy * (k2 + k4 * ux + k5 * uz + k7 * ux * uz),
                       duz * (k3 + k5 * uy + k6 * ux + k7 * ux * uy));
}

// The eddies' stream function (cells^2 / s) and its gradient at (x, y), time t: two octaves
// that evolve and drift up with the plume (air.glsl:130-141).
SP_DEV float3 air_psi(const sp_gparams& P, float x, float y, float t) {
    const float L = fmaxf_(P.steam2[3], 1.0f);
    const float turn = 0.7f * t * P.steam2[2] / L;
    const float ax = x / L, ay = (y - 0.5f * P.steam[1] * t) / L;
    const float4 n1 = vnoise3(ax, ay, turn);
    const float4 n2 = vnoise3(ax * 2.17f + 17.3f, ay * 2.17f + 17.3f, 2.2f * turn);
    const float amp = 2.2f * P.steam2[2] * L;
    return make_float3(amp * (n1.x + 0.45f * n2.x), amp * ((n1.y + 0.45f * 2.17f * n2.y) / L),
                       amp * ((n1.z + 0.45f * 2.17f * n2.z) / L));
}

// Each puff leans a little to one side as it rises, -0.15 .. 0.15 (air.glsl:143-146).
SP_DEV float puff_lean(uint32_t pid) { return ((float)(hash_u(pid ^ 0x5bd1e995u) >> 8) / 16777216.0f - 0.5f) * 0.3f; }

// Liquid occupancy at (x, y), bilinear over the cell centres (air.glsl:81-88).
SP_DEV float sample_occ(const Ctx& c, float x, float y) {
    const float gx = clampf_(x - 0.5f, 0.0f, (float)c.W - 1.001f), gy = clampf_(y - 0.5f, 0.0f, (float)c.H - 1.001f);
    const int x0 = (int)gx, y0 = (int)gy;
    const float fx = gx - (float)x0, fy = gy - (float)y0;
    const int k = y0 * c.W + x0;
    return mixf_(mixf_(c.cell[k].x, c.cell[k + 1].x, fx), mixf_(c.cell[k + c.W].x, c.cell[k + c.W + 1].x, fx), fy);
}

// Bilinear velocity over live faces, `fb` where no face is live (air.glsl:41-78).
SP_DEV float2 sample_vel(const Ctx& c, float px, float py, float2 fb);

// The air's velocity at (x, y), cells / s, and the liquid occupancy there (air.glsl:148-175).
SP_DEV float2 air_velocity(const Ctx& c, float x, float y, float lean, float& occ) {
    const sp_gparams& P = *c.P;
    const float t = c.S->advanced;
    float nx, ny;
    const float d = sample_sdf(c, x, y, nx, ny);
    const float nl = len2_(nx, ny);
    if (nl > 1e-5f) {
        nx = nx / nl;
        ny = ny / nl;
    } else {
        nx = 0.0f;
        ny = 1.0f;
    }
    const float Lb = fmaxf_(P.steam2[3], 1.0f);
    const float r = clampf_(d / Lb, 0.0f, 1.0f);
    const float ramp = r * (1.875f - r * r * (1.25f - 0.375f * r * r));
    const float dramp = d < Lb && d > 0.0f ? (1.875f - r * r * (3.75f - 1.875f * r * r)) / Lb : 0.0f;
    float vx = 0.0f, vy = 0.0f;
    if (P.steam2[2] > 0.0f) {
        const float3 psi = air_psi(P, x, y, t);
        vx = ramp * psi.z + psi.x * dramp * ny;
        vy = ramp * -psi.y + psi.x * dramp * -nx;
    }
    float rx = lean * P.steam[1], ry = 1.0f * P.steam[1];
    const float into = rx * nx + ry * ny;
    if (into < 0.0f) {
        const float tx = ny, ty = -nx;
        const float side = (rx * tx + ry * ty) + lean * 1e-3f >= 0.0f ? 1.0f : -1.0f;
        rx -= (nx + side * tx * 0.7f) * into * (1.0f - ramp);
        ry -= (ny + side * ty * 0.7f) * into * (1.0f - ramp);
    }
    vx += rx;
    vy += ry;
    // near liquid the air moves with it, and the steam still rises through it
    occ = sample_occ(c, x, y);
    const float2 vl = sample_vel(c, x, y, make_float2(vx, vy - P.steam[1]));
    const float k = smoothstep_(0.0f, 0.4f, occ);
    vx = mixf_(vx, vl.x, k);
    vy = mixf_(vy, vl.y + P.steam[1], k);
    // the tool stirs, pushes and blows the air as it does the liquid (Erase: the caller)
    const int tm = P.tool_mode;
    if (tm == 1 || tm == 3 || tm == 4 || tm == 7) {
        const float tdx = P.tool_pos[0] - x, tdy = P.tool_pos[1] - y;
        const float r2 = (tdx * tdx + tdy * tdy) / (P.tool_radius * P.tool_radius);
        if (r2 < 1.0f) {
            const float2 tt = tool_target(P, tdx, tdy);
            vx = mixf_(vx, tt.x, 1.0f - r2);
            vy = mixf_(vy, tt.y, 1.0f - r2);
        }
    }
    return make_float2(vx, vy);
}

// A puff of vapour that left the tank (particles.comp:103-134): a tracer in the air above the
// liquid, gone after steam.z seconds, at the edge of the grid, or swallowed by the liquid. The
// vapour it lays into the renderer's smoke field is not carried (no smoke field here).
SP_DEV void advance_puff(const Ctx& c, uint32_t id, sp_particle p) {
    const sp_gparams& P = *c.P;
    const float dt = c.S->dt;
    p.s += dt;
    if (p.s >= P.steam[2] || p.pos[0] < 2.01f || p.pos[1] < 2.01f || p.pos[0] > (float)c.W - 2.01f ||
        p.pos[1] > (float)c.H - 2.01f) {
        c.part[id].kind = SP_KIND_DEAD;
        return;
    }
    float occ;
    const float2 target = air_velocity(c, p.pos[0], p.pos[1], puff_lean(p.id), occ);
    if (occ > 0.6f && p.s > 0.1f) {
        c.part[id].kind = SP_KIND_DEAD;
        return;
    }
    if (P.tool_mode == 2 && len2_(P.tool_pos[0] - p.pos[0], P.tool_pos[1] - p.pos[1]) < P.tool_radius) {
        c.part[id].kind = SP_KIND_DEAD;
        return;
    }
    const float k = 1.0f - expf(-dt / PUFF_FOLLOW);
    float vx = mixf_(p.vel[0], target.x, k), vy = mixf_(p.vel[1], target.y, k);
    float x = p.pos[0] + vx * dt, y = p.pos[1] + vy * dt;
    collide(c, x, y, vx, vy, false);
    p.pos[0] = x;
    p.pos[1] = y;
    p.vel[0] = vx;
    p.vel[1] = vy;
    c.part[id] = p;
}

// A particle that leaves the tank as a puff of vapour `fume` from where it is, with velocity
// (vx, vy) (particles.comp:704-712, 796-800).
SP_DEV void become_puff(const Ctx& c, uint32_t id, sp_particle p, float vx, float vy, float fume) {
    p.kind = SP_KIND_PUFF;
    p.s = 0.0f;
    p.vel[0] = vx;
    p.vel[1] = vy;
    p.c[0] = fume;
    p.c[1] = 0.0f;
    p.c[2] = 0.0f;
    p.c[3] = 0.0f;
    c.part[id] = p;
}

// APIC gather of one face component over its 3 x 3 live faces (particles.comp:165-222).
SP_DEV void gather_face(const Ctx& c, float px, float py, int comp, float& u, float& cr, float& Ca, float& Cb) {
    float gx = comp == 0 ? px : px - 0.5f, gy = comp == 0 ? py - 0.5f : py;
    int bx = floori(gx - 0.5f), by = floori(gy - 0.5f);
    float fx = gx - (float)bx, fy = gy - (float)by;
    float wx[3], wy[3];
    bspline3(fx, wx);
    bspline3(fy, wy);
    float sw = 0.0f, su = 0.0f, sc = 0.0f, sdx = 0.0f, sdy = 0.0f, sudx = 0.0f, sudy = 0.0f;
#pragma unroll
    for (int b = 0; b < 3; b++)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            float4 q = c.vel[cidx(c, bx + a, by + b)];
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
        u = uu;
        cr = sc / sw;
        Ca = 4.0f * (sudx - uu * sdx) / sw;
        Cb = 4.0f * (sudy - uu * sdy) / sw;
    }
}

// Gather, drift, exchange, advect and collide one particle (particles.comp:98-985, the
// branches this build carries); fills the scatter inputs. X: the tank may be elastic or
// settle, or hold a volatile liquid (that code is compiled only into the kernels such tanks
// run).
template <bool X>
SP_DEV void advance(const Ctx& c, uint32_t id, int gather, ScatterIn& g, float& umax) {
    const sp_gparams& P = *c.P;
    g.live = false;
    if (id >= P.n_slots) return;
    sp_particle p = c.part[id];
    if (p.kind == SP_KIND_DEAD) return;
    if ((p.kind & 255u) == 0u) {
        // a puff (only tanks that may hold them run the kernel with their code)
        if (X && (P.caps & SP_CAP_PUFFS)) advance_puff(c, id, p);
        return;
    }
    float dt = c.S->dt;
    int row = (int)(p.kind & 255u) - 1;
    int Mgroup = c.mats[row].group, Mcw = c.mats[row].chan_w, Mcs = c.mats[row].chan_s;
    int fam = c.mats[row].fam;
    const bool th_on = (P.caps & SP_CAP_THERMAL) != 0u;
    const bool rig = (P.caps & SP_CAP_RIGID) != 0u;
    // frozen particles, and those too viscous to flow, belong to a rigid body
    // (particles.comp:145-151)
    bool frozen = false;
    if (rig) {
        const float2 ph = th_on ? phase_of(p.t, thermal_of(c, row, p.s)) : make_float2(P.heat_x, 0.0f);
        frozen = ph.y > 0.5f || row_lnu(c, row, p.s, ph.x, th_on) >= P.ln_rigid;
    }
    // viscoelastic stress: a particle that has not run a substep yet starts unstressed (its
    // slot may hold an old stress) (particles.comp:152-156)
    const bool el_on = X && (P.caps & SP_CAP_ELASTIC) != 0u;
    float4 tau = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (el_on && (p.kind & SP_KIND_FRESH) == 0u && c.mats[row].gr[1] > 0.0f) tau = c.pstress[id];
    if (gather) {
        float vx = p.vel[0], vy = p.vel[1];
        float C0 = p.c[0], C1 = p.c[1], C2 = p.c[2], C3 = p.c[3];
        float corrx = 0.0f, corry = 0.0f, slip = 0.0f, crowd = 0.0f;
        gather_face(c, p.pos[0], p.pos[1], 0, vx, corrx, C0, C1);
        gather_face(c, p.pos[0], p.pos[1], 1, vy, corry, C2, C3);
        // a frozen particle, a yield-stress particle on a plug, or any particle in a cell of a
        // body moves exactly with the rigid body around it: the first body in the 3 x 3 cells
        // around its own, own cell first (particles.comp:223-248)
        bool held = false;
        if (rig) {
            const int px1 = clampi_(floori(p.pos[0]), 1, c.W - 2), py1 = clampi_(floori(p.pos[1]), 1, c.H - 2);
            const bool in_body = row != SP_ROW_STEAM && c.rl[cidx(c, px1, py1)].x >= 0;
            // a yield-stress liquid's particle in a cell too sparse to be liquid (the surface
            // of a resting pile), next to a plug
            const int px0 = clampi_(floori(p.pos[0]), 0, c.W - 1), py0 = clampi_(floori(p.pos[1]), 0, c.H - 1);
            const bool plug = (c.mats[row].rheo_p[0] > 0.0f || c.mats[row].temp[3] > 0.0f) && c.cell[cidx(c, px0, py0)].w <= 0.0f;
            if (frozen || plug || in_body) {
                const int ox[9] = {0, -1, 1, 0, 0, -1, 1, -1, 1}, oy[9] = {0, 0, 0, -1, 1, -1, -1, 1, 1};
                for (int k = 0; k < 9 && !held; k++) {
                    const int root = c.rl[cidx(c, px1 + ox[k], py1 + oy[k])].x;
                    if (root < 0 || c.rl[root].y < 0) continue;
                    float3 u;
                    if (body_motion(c, (uint32_t)c.rl[root].y, root, p.pos[0], p.pos[1], u)) {
                        vx = u.x;
                        vy = u.y;
                        C0 = 0.0f;
                        C1 = -u.z;
                        C2 = u.z;
                        C3 = 0.0f;
                        held = true;
                    }
                }
            }
        }
        if (el_on) {
            const float G = c.mats[row].gr[1];
            if (G > 0.0f) tau = stress_update(c, p.pos[0], p.pos[1], C0, C1, C2, C3, tau, G, c.mats[row].gr[2], dt);
            c.pstress[id] = tau;
        }
        if (P.apic_clamp > 0.0f) {
            float lim = P.apic_clamp / dt;
            C0 = clampf_(C0, -lim, lim);
            C1 = clampf_(C1, -lim, lim);
            C2 = clampf_(C2, -lim, lim);
            C3 = clampf_(C3, -lim, lim);
        }
        if (P.corr[0] > 0.0f) {
            float cl = len2_(corrx, corry), cmax = P.corr[0] / dt;
            if (cl > cmax) {
                float k = cmax / cl;
                corrx *= k;
                corry *= k;
            }
        }
        {
            float gx = p.pos[0] - 0.5f, gy = p.pos[1] - 0.5f;
            int bx = floori(gx - 0.5f), by = floori(gy - 0.5f);
            float fx = gx - (float)bx, fy = gy - (float)by;
            float wx[3], wy[3];
            bspline3(fx, wx);
            bspline3(fy, wy);
            float sw = 0.0f, sfw = 0.0f, sfs = 0.0f, sgx = 0.0f, sgy = 0.0f, sgz = 0.0f;
            bool more = (P.group_mask & 12u) != 0u;
            bool groups = (P.group_mask & (P.group_mask - 1u)) != 0u;
            bool sorts = group_sorts(P, c.mats, Mgroup);
            float sbg = 0.0f;
            // the rise of steam through the aqueous liquid around it: the share of steam
            // (particles.comp:303-305)
            const bool steamy = X && (P.caps & SP_CAP_STEAM) && Mgroup == 0;
            float sfv = 0.0f;
            int nmix = P.n_fam - 1;
#pragma unroll
            for (int b = 0; b < 3; b++)
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    int ci = cidx(c, bx + a, by + b);
                    float4 q = c.comp[ci];
                    float wt = wx[a] * wy[b];
                    crowd += wt * q.w;
                    if (q.x > 0.0f) {
                        sw += wt;
                        sgx += wt * q.y;
                        if (more) {
                            float4 ax = c.aux[ci];
                            sgy += wt * ax.z;
                            sgz += wt * ax.w;
                        }
                        if (sorts) sbg += wt * f4c(c.gbase[ci], Mgroup - 1);
                        if (steamy) sfv += wt * c.gbase[ci].w;
                        float sb = fam > 0 ? c.mixb[(size_t)ci * nmix + fam - 1] : q.z;
                        if (sb >= 0.0f) {
                            sfw += wt;
                            sfs += wt * sb;
                        }
                    }
                }
            // Settling (particles.comp:332-351): the cells' gain and loss factors of the
            // particle's family, with its scatter weights; water that is the primary family's
            // diluent takes part at concentration 0 and becomes the concentrate when it gains.
            if (X && (P.caps & SP_CAP_SETTLE)) {
                const bool dil0 = fam < 0 && P.primary >= 0 && P.fams[0][1] == row;
                const int sf = dil0 ? 0 : fam;
                if (sf >= 0 && (p.kind & SP_KIND_FRESH) == 0u && c.mats[dil0 ? P.primary : row].grain > 0.0f) {
                    float dsx = 0.0f, dsy = 0.0f;
                    for (int b = 0; b < 3; b++)
                        for (int a = 0; a < 3; a++) {
                            const float wt = wx[a] * wy[b];
                            const float2 d = c.dsettle[(size_t)cidx(c, bx + a, by + b) * P.n_fam + sf];
                            dsx += wt * d.x;
                            dsy += wt * d.y;
                        }
                    const float s0 = dil0 ? 0.0f : p.s;
                    const float s1 = clampf_(s0 + (1.0f - s0) * dsx - s0 * dsy, 0.0f, 1.0f);
                    const float c0 = th_on ? thermal_of(c, row, p.s).w : 0.0f;
                    if (dil0 && s1 > 0.0f) {
                        row = P.primary;
                        Mgroup = c.mats[row].group;
                        Mcw = c.mats[row].chan_w;
                        Mcs = c.mats[row].chan_s;
                        fam = 0;
                        p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                    }
                    if (!dil0 || s1 > 0.0f) p.s = s1;
                    if (th_on) p.t = keep_heat(p.t, c0, thermal_of(c, row, p.s).w);
                }
            }
            // Heat (particles.comp:353-373): the change of enthalpy temperature the heat
            // pass worked out for the cells this particle scattered to, with the same
            // weights, and the exchange among the particles of its own cell, against its
            // own temperature as it entered the cell's sums.
            if (th_on && (p.kind & SP_KIND_FRESH) == 0u) {
                float dT = 0.0f;
#pragma unroll
                for (int b = 0; b < 3; b++)
#pragma unroll
                    for (int a = 0; a < 3; a++) dT += wx[a] * wy[b] * c.dheat[cidx(c, bx + a, by + b)].x;
                const int ox = clampi_(floori(p.pos[0]), 0, c.W - 1), oy = clampi_(floori(p.pos[1]), 0, c.H - 1);
                const float4 own = c.dheat[cidx(c, ox, oy)];
                if (own.y > 0.0f) {
                    const float4 th = thermal_of(c, row, p.s);
                    const int2 q = heat_quanta(th.w, fminf_(phase_of(p.t, th).x, boil_of(c, row, p.s)));
                    dT += own.y * (own.z - (float)q.y / (float)q.x) * SP_T_SCALE;
                }
                p.t += dT;
            }
            if (sw > 1e-4f) {
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
                            int ci = cidx(c, bx + a, by + b);
                            float4 q = c.comp[ci];
                            if (q.x <= 0.0f) continue;
                            float wt = wx[a] * wy[b];
                            float4 gr = c.grho[ci];
                            sgr[0] += wt * gr.x;
                            sgr[1] += wt * gr.y;
                            sgr[2] += wt * gr.z;
                            sgr[3] += wt * gr.w;
                            float4 ax = c.aux[ci];
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
                    float rho_g = fg > 1e-4f ? sgr[gg] / fg : row_density(c, row, p.s);
                    float rho_o = (sgr[0] + sgr[1] + sgr[2] + sgr[3] - sgr[gg]) / others;
                    float drho = rho_g - rho_o;
                    int pcx = clampi_((int)p.pos[0], 1, c.W - 2), pcy = clampi_((int)p.pos[1], 1, c.H - 2);
                    int k = cidx(c, pcx, pcy);
                    float fsv[4] = {1.0f - sgx - sgy - sgz, sgx, sgy, sgz};
                    int G = fsv[0] >= fmaxf_(fsv[1], fmaxf_(fsv[2], fsv[3])) ? 0 : fsv[1] >= fmaxf_(fsv[2], fsv[3]) ? 1 : fsv[2] >= fsv[3] ? 2 : 3;
                    float rho_c = sgr[G] / fmaxf_(fsv[G], 1e-4f);
                    float nu_c = snw[G] > 1e-6f ? snu[G] / snw[G] : c.aux[k].x;
                    float vt = terminal(c, drho, rho_c, nu_c, P.sep_tau);
                    float sep_max = P.corr[2] > 0.0f ? P.cfl / dt : 1e30f;
                    slip = -fminf_(vt, sep_max) * signf_(drho) * others;
                    float fs[4] = {1.0f - sgx - sgy - sgz, sgx, sgy, sgz};
                    fs[gg] = -1.0f;
                    int h = fs[0] >= fmaxf_(fs[1], fmaxf_(fs[2], fs[3])) ? 0 : fs[1] >= fmaxf_(fs[2], fs[3]) ? 1 : fs[2] >= fs[3] ? 2 : 3;
                    float bond = fabsf(drho) * P.gravity / fmaxf_(P.pair[gg][h], 1e-9f);
                    float hold = clampf_(1.0f - bond, 0.0f, 1.0f);
                    if (slip != 0.0f && hold > 0.0f) {
                        int hi = gg > h ? gg : h;
                        int W = c.W;
                        float grx, gry;
                        if (hi == 1) {
                            grx = 0.5f * (c.surf[k + 1].y - c.surf[k - 1].y);
                            gry = 0.5f * (c.surf[k + W].y - c.surf[k - W].y);
                        } else if (hi == 2) {
                            grx = 0.5f * (c.surf2[k + 1].x - c.surf2[k - 1].x);
                            gry = 0.5f * (c.surf2[k + W].x - c.surf2[k - W].x);
                        } else {
                            grx = 0.5f * (c.surf2[k + 1].y - c.surf2[k - 1].y);
                            gry = 0.5f * (c.surf2[k + W].y - c.surf2[k - W].y);
                        }
                        float resolved = 1.0f - clampf_((0.12f - len2_(grx, gry)) / (0.12f - 0.04f), 0.0f, 1.0f);
                        slip *= 1.0f - hold * resolved;
                    }
                }
                if (sorts && !held) {
                    float fgi = groups ? sgv[Mgroup - 1] : 1.0f;
                    if (fgi > 1e-4f) {
                        float rmean = sbg / sw / fgi;
                        float drho = c.mats[row].rho - rmean;
                        int pcx = clampi_((int)p.pos[0], 0, c.W - 1), pcy = clampi_((int)p.pos[1], 0, c.H - 1);
                        float4 cq = c.cell[cidx(c, pcx, pcy)];
                        float nu_c = cq.y > 0.0f ? cq.z / cq.y : c.aux[cidx(c, pcx, pcy)].x;
                        float vt = terminal(c, drho, rmean, nu_c, P.sep_tau);
                        float sep_max = P.corr[2] > 0.0f ? P.cfl / dt : 1e30f;
                        slip -= fminf_(vt, sep_max) * signf_(drho);
                    }
                }
                // Steam rises through the water around it (particles.comp:505-517): a bubble of
                // the unresolved size at its terminal velocity in water, by the share of water
                // around it; the water drifts down by the share of steam.
                if (steamy && !held) {
                    const float fv = clampf_(sfv / sw, 0.0f, 1.0f);
                    if (fv > 1e-4f) {
                        const float rw = c.mats[SP_ROW_WATER].rho;
                        float vt = terminal(c, c.mats[SP_ROW_STEAM].rho - rw, rw, expf(c.mats[SP_ROW_WATER].ln_nu), P.sep_tau);
                        const float sep_max = P.corr[2] > 0.0f ? P.cfl / dt : 1e30f;
                        vt = fminf_(vt, sep_max);
                        slip += row == SP_ROW_STEAM ? vt * (1.0f - fv) : -vt * fv;
                    }
                }
                bool diluent0 = fam < 0 && P.primary >= 0 && P.fams[0][1] == row;
                if (fam >= 0 || diluent0) {
                    if (sfw > 1e-4f) {
                        float s0 = diluent0 ? 0.0f : p.s;
                        int rr = diluent0 ? P.primary : row;
                        float s1 = clampf_(s0 + (sfs / sfw - s0) * (1.0f - expf(-c.mats[rr].mix_rate * dt)), 0.0f, 1.0f);
                        const float c0 = th_on ? thermal_of(c, row, p.s).w : 0.0f;
                        if (!diluent0) {
                            p.s = s1;
                        } else if (s1 > 0.0f) {
                            row = P.primary;
                            Mgroup = c.mats[row].group;
                            Mcw = c.mats[row].chan_w;
                            Mcs = c.mats[row].chan_s;
                            p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                            p.s = s1;
                        }
                        // the same heat content at the new heat capacity (particles.comp:541)
                        if (th_on) p.t = keep_heat(p.t, c0, thermal_of(c, row, p.s).w);
                    }
                }
                if (fam < 0 && (P.caps & 1u) && (int)(p.kind & 255u) - 1 == row && c.mats[row].fam < 0) {
                    for (int j = P.primary >= 0 ? 1 : 0; j < P.n_fam; j++) {
                        if (P.fams[j][1] != row) continue;
                        float found = 0.0f;
                        if (j == 0) {
                            found = sfw;
                        } else {
                            for (int b = 0; b < 3; b++)
                                for (int a = 0; a < 3; a++)
                                    if (c.mixb[(size_t)cidx(c, bx + a, by + b) * nmix + j - 1] >= 0.0f) found += wx[a] * wy[b];
                        }
                        if (found > 1e-4f) {
                            row = P.fams[j][0];
                            Mgroup = c.mats[row].group;
                            Mcw = c.mats[row].chan_w;
                            Mcs = c.mats[row].chan_s;
                            p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                            p.s = 0.0f;
                            break;
                        }
                    }
                }
            }
        }
        // a held particle takes the density correction only where its body holds a clump
        // (particles.comp:573-576)
        if (held && crowd <= SP_REST_PHI + fmaxf_(P.corr[1], 0.05f)) {
            corrx = 0.0f;
            corry = 0.0f;
        }
        // Drains (particles.comp:578-592): a zone with a negative rate removes the liquid in it,
        // each particle with probability -rate x dt, falling off to a quarter at the rim.
        if (X && (P.caps & SP_CAP_CURRENTS)) {
            const int nz = (int)P.ext[3];
            for (int i = 0; i < nz; i++) {
                const float zw = P.current_zone[i][3];
                if (zw >= 0.0f) continue;
                const float dzx = p.pos[0] - P.current_zone[i][0], dzy = p.pos[1] - P.current_zone[i][1];
                const float rz = P.current_zone[i][2];
                const float r2 = (dzx * dzx + dzy * dzy) / (rz * rz);
                if (r2 < 1.0f) {
                    const float u =
                        (float)(hash_u(p.id * 2246822519u + c.S->substeps * 3266489917u + (uint32_t)i) >> 8) / 16777216.0f;
                    if (u < -zw * (1.0f - 0.75f * r2) * dt) {
                        c.part[id].kind = SP_KIND_DEAD;
                        return;
                    }
                }
            }
        }
        // The mouse tool (particles.comp:594-631), inside its radius: Erase removes the
        // particle; Heat and Cool add or take heat_power x (1 - r^2) x dt kelvin for water
        // (not below -150 deg C), counted in the ledger; the others drive the velocity towards
        // the tool's target by strength x (1 - r^2) x dt, less in a crowd (but the nozzle).
        bool pushed = false;
        if (X && P.tool_mode != 0) {
            const float tdx = P.tool_pos[0] - p.pos[0], tdy = P.tool_pos[1] - p.pos[1];
            const float r2 = (tdx * tdx + tdy * tdy) / (P.tool_radius * P.tool_radius);
            if (r2 < 1.0f) {
                if (P.tool_mode == 2) {
                    c.part[id].kind = SP_KIND_DEAD;
                    return;
                }
                const float fall = 1.0f - r2;
                if (P.tool_mode == 5 || P.tool_mode == 6) {
                    if (th_on) {
                        float e = P.heat_power * fall * dt;
                        const float4 tth = thermal_of(c, row, p.s);
                        const float C = tth.w;
                        if (P.tool_mode == 6) e = fminf_(e, fmaxf_(phase_of(p.t, tth).x + 150.0f, 0.0f) * C);
                        p.t += (P.tool_mode == 5 ? e : -e) / C;
                        ledger_add(c, P.tool_mode == 5 ? 4 : 5, e * P.inv_ppc);
                    }
                } else {
                    float k = clampf_(P.tool_strength * fall * dt, 0.0f, 1.0f);
                    if (P.tool_crowd_max > 0.0f && P.tool_mode != 4)
                        k *= clampf_((P.tool_crowd_max - crowd) / (0.25f * P.tool_crowd_max), 0.0f, 1.0f);
                    const float2 tt = tool_target(P, tdx, tdy);
                    vx = mixf_(vx, tt.x, k);
                    vy = mixf_(vy, tt.y, k);
                    pushed = true;
                }
            }
        }
        // Boiling (particles.comp:633-731). Liquid past its boiling point takes up latent heat
        // as it goes on heating; the share it holds is its vapour quality x. An aqueous liquid
        // in a tank that makes steam turns into steam at x = steam.w, keeping its mass and heat
        // and the row it boiled from (KIND_SRC); steam condenses back below x = steam.x, and at
        // the free surface escapes as a puff with probability x, else gives up its latent heat
        // there and stays, as its liquid at the boiling point. Any other liquid leaves at x = 1.
        // What leaves leaves its residue behind, one particle per residue share on average.
        if (X && th_on && (P.caps & SP_CAP_BOILS)) {
            const float4 th = thermal_of(c, row, p.s);
            const int dil = c.mats[row].diluent;
            float bo_t = c.mats[row].th2[2], bo_l = c.mats[row].th2[3];
            if (dil >= 0) {
                bo_t = mixf_(c.mats[dil].th2[2], bo_t, p.s);
                bo_l = mixf_(c.mats[dil].th2[3], bo_l, p.s);
            }
            const bool vapour = (P.caps & SP_CAP_STEAM) && Mgroup == 0 && (P.steam2[1] > 0.5f || row == SP_ROW_STEAM);
            bool leave = false;
            if (vapour && row == SP_ROW_STEAM) {
                const float x = (p.t - bo_t) / fmaxf_(bo_l, 1e-3f);
                const int pcx = clampi_((int)p.pos[0], 0, c.W - 1), pcy = clampi_((int)p.pos[1], 0, c.H - 1);
                const bool surface = c.cell[cidx(c, pcx, pcy)].w <= 0.0f || c.cell[cidx(c, pcx, clampi_(pcy + 1, 0, c.H - 1))].w <= 0.0f;
                const float u = (float)(hash_u(p.id ^ hash_u(c.S->substeps ^ 0x2545f491u)) >> 8) / 16777216.0f;
                if (surface && u < x) {
                    leave = true;
                } else if (x < P.steam[0] || surface) {
                    // back into the liquid it boiled from: at the surface, its latent heat
                    // given up to the air; below it, condensed
                    if (surface) {
                        ledger_add(c, 6, fmaxf_(th.w * (p.t - bo_t), 0.0f) * P.inv_ppc);
                        p.t = bo_t;
                    }
                    const uint32_t src = (p.kind & SP_KIND_SRC) >> 8;
                    row = src > 0u ? (int)src - 1 : SP_ROW_WATER;
                    Mgroup = c.mats[row].group;
                    Mcw = c.mats[row].chan_w;
                    Mcs = c.mats[row].chan_s;
                    fam = c.mats[row].fam;
                    p.kind = (p.kind & ~(255u | SP_KIND_SRC)) | (uint32_t)(row + 1);
                    p.t = keep_heat(p.t, th.w, thermal_of(c, row, p.s).w);
                }
            } else if (p.t >= bo_t + bo_l * (vapour ? P.steam[3] : 1.0f)) {
                if (vapour) {
                    const float c1 = thermal_of(c, SP_ROW_STEAM, 1.0f).w;
                    atomicAdd(&c.ledger[16 + 2 * SP_MAX_MATERIALS + row], 1u);
                    p.kind = (p.kind & ~(255u | SP_KIND_SRC)) | ((uint32_t)(row + 1) << 8) | (uint32_t)(SP_ROW_STEAM + 1);
                    row = SP_ROW_STEAM;
                    Mgroup = c.mats[row].group;
                    Mcw = c.mats[row].chan_w;
                    Mcs = c.mats[row].chan_s;
                    fam = -1;
                    p.t = keep_heat(p.t, th.w, c1);
                } else {
                    leave = true;
                }
            }
            if (leave) {
                const uint32_t src = row == SP_ROW_STEAM ? (p.kind & SP_KIND_SRC) >> 8 : (uint32_t)(row + 1);
                const int from = src > 0u ? (int)src - 1 : SP_ROW_WATER;
                const int res = c.mats[from].residue;
                const float share = c.mats[from].residue_share * (c.mats[from].diluent >= 0 ? p.s : 1.0f);
                // heat content counted from absolute zero, as the ledger's other entries
                float e = th.w * (p.t + 273.15f);
                atomicAdd(&c.ledger[16 + row], 1u);
                if (res >= 0 && (float)(hash_u(p.id ^ 0x9e3779b9u) >> 8) / 16777216.0f < share) {
                    atomicAdd(&c.ledger[16 + SP_MAX_MATERIALS + res], 1u);
                    row = res;
                    Mgroup = c.mats[row].group;
                    Mcw = c.mats[row].chan_w;
                    Mcs = c.mats[row].chan_s;
                    fam = c.mats[row].fam;
                    p.kind = (uint32_t)(row + 1);
                    p.s = 1.0f;
                    p.t = bo_t - thermal_of(c, row, 1.0f).z;  // the residue is a solid
                    e -= thermal_of(c, row, 1.0f).w * (p.t + 273.15f);
                    ledger_add(c, 6, fmaxf_(e, 0.0f) * P.inv_ppc);
                } else {
                    ledger_add(c, 6, fmaxf_(e, 0.0f) * P.inv_ppc);
                    const float fume = row == SP_ROW_STEAM ? 1.0f : c.mats[row].fume;
                    if (fume > 0.0f) become_puff(c, id, p, vx, vy, fume);
                    else c.part[id].kind = SP_KIND_DEAD;
                    return;
                }
            }
        }
        // Heat made inside (particles.comp:733-739): the liquid's own K/s, counted in the ledger.
        if (X && th_on && (P.caps & SP_CAP_SELF_HEAT) && c.mats[row].rx[2] > 0.0f) {
            const float e = c.mats[row].rx[2] * c.mats[row].th[0] * (c.mats[row].diluent >= 0 ? p.s : 1.0f) * dt * P.heat_rate;
            p.t += e / thermal_of(c, row, p.s).w;
            ledger_add(c, 7, e * P.inv_ppc);
        }
        // A reaction (particles.comp:740-770): where the particle's cell also holds its partner
        // it turns into the product with probability rate x dt, its heat of reaction warming it;
        // with no product it is used up (counted in the ledger).
        if (X && (P.caps & SP_CAP_REACT) && c.mats[row].ic[0] >= 0) {
            const int rx = clampi_((int)p.pos[0], 0, c.W - 1), ry = clampi_((int)p.pos[1], 0, c.H - 1);
            if ((c.rmask[cidx(c, rx, ry)] & (1u << (uint32_t)c.mats[row].ic[2])) != 0u) {
                const float own = c.mats[row].diluent >= 0 ? p.s : 1.0f;
                const float u = (float)(hash_u(p.id * 0x9e3779b1u ^ hash_u(c.S->substeps + 0x51ed270bu)) >> 8) / 16777216.0f;
                if (u < c.mats[row].rx[0] * dt && own > SP_REACT_MIN) {
                    const float c0 = thermal_of(c, row, p.s).w;
                    const float e = c.mats[row].rx[1] * c.mats[row].th[0] * own;
                    if (th_on) {
                        p.t += e / c0;
                        ledger_add(c, 7, e * P.inv_ppc);
                    }
                    const int into = c.mats[row].ic[1];
                    if (into < 0) {
                        if (th_on) ledger_add(c, 6, fmaxf_(c0 * (p.t + 273.15f), 0.0f) * P.inv_ppc);
                        atomicAdd(&c.ledger[16 + row], 1u);
                        c.part[id].kind = SP_KIND_DEAD;
                        return;
                    }
                    row = into;
                    Mgroup = c.mats[row].group;
                    Mcw = c.mats[row].chan_w;
                    Mcs = c.mats[row].chan_s;
                    fam = c.mats[row].fam;
                    p.kind = (p.kind & ~255u) | (uint32_t)(row + 1);
                    if (c.mats[row].diluent < 0) p.s = 1.0f;
                    if (th_on) p.t = keep_heat(p.t, c0, thermal_of(c, row, p.s).w);
                }
            }
        }
        // Evaporation below the boiling point (particles.comp:771-807): a particle of a volatile
        // liquid in a cell open to the air (not one that only touches a solid) leaves with
        // probability rate x dt, the rate doubling every 10 K and a concentrate's scaled by its
        // concentration; it takes its heat with it (ledger entry 6) and is counted.
        if (X && (P.caps & SP_CAP_VOLATILE) && c.mats[row].rx[3] > 0.0f) {
            const int qx = clampi_((int)p.pos[0], 1, c.W - 2), qy = clampi_((int)p.pos[1], 1, c.H - 2);
            const int qi = cidx(c, qx, qy);
            bool open_air = false;
            for (int q = 0; q < 4; q++) {
                const int n = qi + (q == 0 ? -1 : q == 1 ? 1 : q == 2 ? -c.W : c.W);
                if (c.cell[n].w <= 0.0f && c.solid[n].w > 0.5f) open_air = true;
            }
            if (open_air) {
                const float4 th = thermal_of(c, row, p.s);
                const float Tp = th_on ? phase_of(p.t, th).x : P.heat_x;
                const float rate = c.mats[row].rx[3] * (c.mats[row].diluent >= 0 ? p.s : 1.0f) *
                                   exp2f(clampf_((Tp - 20.0f) / 10.0f, -6.0f, 5.0f));
                const float u = (float)(hash_u(p.id * 0x85ebca6bu ^ hash_u(c.S->substeps + 0x27d4eb2fu)) >> 8) / 16777216.0f;
                if (u < rate * dt) {
                    if (th_on) ledger_add(c, 6, fmaxf_(th.w * (p.t + 273.15f), 0.0f) * P.inv_ppc);
                    atomicAdd(&c.ledger[16 + row], 1u);
                    // a liquid that fumes leaves a puff of mist
                    const float fume = c.mats[row].fume;
                    if (fume > 0.0f) become_puff(c, id, p, vx, vy, fume);
                    else c.part[id].kind = SP_KIND_DEAD;
                    return;
                }
            }
        }
        // Currents (particles.comp:809-829): the velocity along each current is raised towards
        // its speed, never braked, and the motion across it left alone; a particle already in a
        // crowd is driven less, and not at all past tool_crowd_max.
        if (X && (P.caps & SP_CAP_CURRENTS)) {
            const float drive =
                P.tool_crowd_max > 0.0f ? clampf_((P.tool_crowd_max - crowd) / (0.25f * P.tool_crowd_max), 0.0f, 1.0f) : 1.0f;
            const int nz = (int)P.ext[3];
            for (int i = 0; i < nz && drive > 0.0f; i++) {
                const float zw = P.current_zone[i][3];
                if (zw <= 0.0f) continue;
                const float dzx = p.pos[0] - P.current_zone[i][0], dzy = p.pos[1] - P.current_zone[i][1];
                const float rz = P.current_zone[i][2];
                const float r2 = (dzx * dzx + dzy * dzy) / (rz * rz);
                if (r2 < 1.0f) {
                    const float tx = P.current_vel[i][0], ty = P.current_vel[i][1];
                    const float speed = len2_(tx, ty);
                    const float m = fmaxf_(speed, 1e-6f);
                    const float dx_ = tx / m, dy_ = ty / m;
                    const float fall = (1.0f - r2) * (1.0f - r2);
                    const float gain = fmaxf_(speed - (vx * dx_ + vy * dy_), 0.0f);
                    const float relax = 1.0f - expf(-zw * fall * dt);
                    vx += dx_ * gain * relax * drive;
                    vy += dy_ * gain * relax * drive;
                }
            }
        }
        float vcap = P.ext[0];
        {
            float sp = len2_(vx, vy);
            if (sp > vcap) {
                float k = vcap / sp;
                vx *= k;
                vy *= k;
            }
        }
        float vax = vx, vay = vy;
        if (P.ext[1] >= 2.0f && !pushed && !held) {
            float h = 0.5f * dt;
            float ex = (vx + corrx) + 0.0f, ey = (vy + corry) + slip;
            float xmx = clampf_(p.pos[0] + h * ex, 2.01f, (float)c.W - 2.01f);
            float xmy = clampf_(p.pos[1] + h * ey, 2.01f, (float)c.H - 2.01f);
            float2 here = sample_vel(c, p.pos[0], p.pos[1], make_float2(vx, vy));
            float2 mid = sample_vel(c, xmx, xmy, here);
            vax = vx + mid.x - here.x;
            vay = vy + mid.y - here.y;
            float sp = len2_(vax, vay);
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
            float mine = len2_(dispx, dispy);
            if (crowd < SP_LIQUID_THR) mine /= P.ext[2];
            umax = fmaxf_(umax, mine);
        }
        float sdx = dt * dispx, sdy = dt * dispy;
        int pieces = clampi_((int)ceilf(len2_(sdx, sdy)), 1, 32);
        sdx /= (float)pieces;
        sdy /= (float)pieces;
        float x = p.pos[0], y = p.pos[1];
        for (int k = 1; k < pieces; k++) {
            x += sdx;
            y += sdy;
            collide(c, x, y, vx, vy, true);
        }
        x += sdx;
        y += sdy;
        float excess = clampf_(crowd - 1.5f, 0.0f, 1.0f);
        if (excess > 0.0f && P.crowd_jitter > 0.0f) {
            uint32_t hh = hash_u(p.id ^ hash_u(__float_as_uint(p.pos[0]) ^ c.S->substeps));
            x += ((float)(hh & 0xffffu) / 65535.0f - 0.5f) * P.crowd_jitter * excess;
            y += ((float)(hh >> 16) / 65535.0f - 0.5f) * P.crowd_jitter * excess;
        }
        collide(c, x, y, vx, vy, true);
        bool again = collide(c, x, y, vx, vy, true);
        if (again) {
            float gx_, gy_;
            if (sample_sdf(c, x, y, gx_, gy_) < -1.0f) collide(c, x, y, vx, vy, false);
        }
        x = clampf_(x, 2.01f, (float)c.W - 2.01f);
        y = clampf_(y, 2.01f, (float)c.H - 2.01f);
        p.pos[0] = x;
        p.pos[1] = y;
        p.vel[0] = vx;
        p.vel[1] = vy;
        p.c[0] = C0;
        p.c[1] = C1;
        p.c[2] = C2;
        p.c[3] = C3;
        p.kind &= ~SP_KIND_FRESH;
        c.part[id] = p;
    }
    g.live = true;
    g.px = p.pos[0];
    g.py = p.pos[1];
    g.vsx = p.vel[0] * dt;
    g.vsy = p.vel[1] * dt;
    g.Csx = p.c[0] * dt;
    g.Csy = p.c[1] * dt;
    g.Csz = p.c[2] * dt;
    g.Csw = p.c[3] * dt;
    float rho_p = row_density(c, row, p.s);
    // Thermal state for the scatter (particles.comp:951-969): the temperature capped at the
    // boiling point, the expansion and the solid's density where frozen, and the
    // particle's quanta into its own cell's sums for the exchange inside the cell.
    float CT = 0.0f;
    if (th_on) {
        const sp_gmat& Mr = c.mats[row];
        const float4 th = thermal_of(c, row, p.s);
        const float2 ph = phase_of(p.t, th);
        const float tb = boil_of(c, row, p.s);
        rho_p *= expansion(c, row, p.s, ph.x);
        if (ph.y > 0.0f) {
            const float rs = Mr.diluent >= 0 ? mixf_(c.mats[Mr.diluent].rho_solid, Mr.rho_solid, p.s) : Mr.rho_solid;
            rho_p = mixf_(rho_p, rs, ph.y);
        }
        CT = th.w * (fminf_(ph.x, tb) + 273.15f) / SP_T_SCALE;
        const int hx = clampi_(floori(p.pos[0]), 0, c.W - 1), hy = clampi_(floori(p.pos[1]), 0, c.H - 1);
        const int2 q = heat_quanta(th.w, fminf_(ph.x, tb));
        int2* hc = &c.hcell[cidx(c, hx, hy)];
        atomicAdd(&hc->x, q.x);
        atomicAdd(&hc->y, q.y);
    }
    g.rho = rho_p;
    g.sc = p.s;
    g.cw = Mcw;
    g.cs = Mcs;
    g.CT = CT;
    g.ct = th_on ? P.chan_t : -1;
    g.cg = th_on ? group_heat_chan(P, c.mats[row].group) : -1;
    // whether there is stress to scatter (particles.comp:983)
    g.ce = el_on && (tau.x != 0.0f || tau.y != 0.0f || tau.z != 0.0f || tau.w != 0.0f) ? P.chan_e : -1;
}

// The scatter of one particle (particles.comp:997-1062) through `add(node x, node y,
// channel, value)`.
template <class Add>
SP_DEV void deposit(const ScatterIn& g, float3 te, Add add) {
    float wx[3], wy[3];
    {
        float gx = g.px, gy = g.py - 0.5f;
        int bx = floori(gx - 0.5f), by = floori(gy - 0.5f);
        float fx = gx - (float)bx, fy = gy - (float)by;
        bspline3(fx, wx);
        bspline3(fy, wy);
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float m = wx[a] * wy[b] * g.rho;
                float dx = (float)a - fx, dy = (float)b - fy;
                add(bx + a, by + b, 0, (int)floorf(m * SP_MASS_SCALE + 0.5f));
                add(bx + a, by + b, 1, (int)floorf(m * (g.vsx + (g.Csx * dx + g.Csy * dy)) * SP_MOM_SCALE + 0.5f));
            }
    }
    {
        float gx = g.px - 0.5f, gy = g.py;
        int bx = floori(gx - 0.5f), by = floori(gy - 0.5f);
        float fx = gx - (float)bx, fy = gy - (float)by;
        bspline3(fx, wx);
        bspline3(fy, wy);
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float m = wx[a] * wy[b] * g.rho;
                float dx = (float)a - fx, dy = (float)b - fy;
                add(bx + a, by + b, 2, (int)floorf(m * SP_MASS_SCALE + 0.5f));
                add(bx + a, by + b, 3, (int)floorf(m * (g.vsy + (g.Csz * dx + g.Csw * dy)) * SP_MOM_SCALE + 0.5f));
            }
    }
    {
        float gx = g.px - 0.5f, gy = g.py - 0.5f;
        int bx = floori(gx - 0.5f), by = floori(gy - 0.5f);
        float fx = gx - (float)bx, fy = gy - (float)by;
        bspline3(fx, wx);
        bspline3(fy, wy);
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                int wi = (int)floorf(wx[a] * wy[b] * SP_MASS_SCALE + 0.5f);
                add(bx + a, by + b, 4, wi);
                if (g.cw >= 0) add(bx + a, by + b, g.cw, wi);
                if (g.cs >= 0 && g.sc > 0.0f) add(bx + a, by + b, g.cs, (int)floorf(wx[a] * wy[b] * g.sc * SP_MASS_SCALE + 0.5f));
                // heat: C (T + 273.15) / T_SCALE into the sum and the group's own (particles.comp:1047-1051)
                if (g.ct >= 0) {
                    const int h = (int)floorf(wx[a] * wy[b] * g.CT * SP_MASS_SCALE + 0.5f);
                    add(bx + a, by + b, g.ct, h);
                    if (g.cg >= 0) add(bx + a, by + b, g.cg, h);
                }
                // viscoelastic stress over el_g, fixed point (particles.comp:1053-1058)
                if (g.ce >= 0) {
                    const float wab = wx[a] * wy[b];
                    add(bx + a, by + b, g.ce, (int)floorf(wab * te.x + 0.5f));
                    add(bx + a, by + b, g.ce + 1, (int)floorf(wab * te.y + 0.5f));
                    add(bx + a, by + b, g.ce + 2, (int)floorf(wab * te.z + 0.5f));
                }
            }
    }
}

}  // namespace

// The particle pass (particles.comp:1064-1101), NT work-items per group, a scatter box of
// BOX ints in LDS, registers for at least MINB groups per compute unit; X for tanks that are
// elastic or settle. Every (NT, BOX) gives the same result (integer sums; the speed maximum is
// a maximum), so the launch shape is free to tune.
// gather < 0: each tank decides from its own step state (the first substep after a reset,
// which k_step has just counted, has no grid velocities to gather yet).
template <int NT, int BOX, int MINB, bool X>
__global__ __launch_bounds__(NT, MINB) void k_particles(const TankBatch tb, int gather) {
    static_assert(NT % 64 == 0 && NT <= 1024 && BOX * 4 <= 64 * 1024, "launch shape");
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (gather < 0) gather = c.S->substeps > 1u ? 1 : 0;
    __shared__ int sacc[BOX];
    __shared__ int sbox[4];
    const int t = threadIdx.x;
    const uint32_t id = blockIdx.x * NT + t;
    ScatterIn g;
    float umax = 0.0f;
    __shared__ float s_max[NT / 64];
    advance<X>(c, id, gather, g, umax);
    group_max_to<NT>(&c.S->umax_part, umax, s_max);

    if (t == 0) {
        sbox[0] = 0x7fffffff;
        sbox[1] = 0x7fffffff;
        sbox[2] = -0x7fffffff;
        sbox[3] = -0x7fffffff;
    }
    __syncthreads();
    int fx = g.live ? floori(g.px) : 0x7fffffff, fy = g.live ? floori(g.py) : 0x7fffffff;
    int lox = wave_min_i(fx), loy = wave_min_i(fy);
    int hix = wave_max_i(g.live ? fx : -0x7fffffff), hiy = wave_max_i(g.live ? fy : -0x7fffffff);
    if ((t & 63) == 0 && lox != 0x7fffffff) {
        atomicMin(&sbox[0], lox);
        atomicMin(&sbox[1], loy);
        atomicMax(&sbox[2], hix);
        atomicMax(&sbox[3], hiy);
    }
    __syncthreads();
    if (sbox[0] == 0x7fffffff) return;  // no live particle in the group
    const int stride = c.P->acc_stride;
    // In LDS a cell's channels take an odd number of words (the accumulator's stride is a
    // multiple of 4): the same channel of neighbouring cells, which a wave's lanes add to
    // together, then falls into different banks. The extra word stays zero.
    const int sstride = stride | 1;
    const int b0x = sbox[0] - 1, b0y = sbox[1] - 1;
    const int bw = sbox[2] + 2 - b0x + 1;
    const int bh = sbox[3] + 2 - b0y + 1;
    const int n = bw * bh * sstride;
    const bool shared = n <= BOX;
    if (shared) {
        for (int k = t; k < n; k += NT) sacc[k] = 0;
        __syncthreads();
        if (g.live)
            deposit(g, scatter_stress(c, g, id),
                    [&](int x, int y, int ch, int v) { atomicAdd(&sacc[((y - b0y) * bw + (x - b0x)) * sstride + ch], v); });
        __syncthreads();
        for (int k = t; k < n; k += NT) {
            int v = sacc[k];
            if (v == 0) continue;
            int cell = k / sstride;
            int x = b0x + cell % bw, y = b0y + cell / bw;
            atomicAdd(&c.acc[(size_t)(y * c.W + x) * stride + (k - cell * sstride)], v);
        }
    } else if (g.live) {
        deposit(g, scatter_stress(c, g, id), [&](int x, int y, int ch, int v) { atomicAdd(&c.acc[(size_t)(y * c.W + x) * stride + ch], v); });
    }
}

// Live particles per cell in a window (count.comp:17-23).
__global__ __launch_bounds__(256) void k_count(const TankBatch tb, int x0, int y0, int w, int h, uint32_t* out) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= c.P->n_slots) return;
    sp_particle p = c.part[id];
    if ((p.kind & 255u) == 0u) return;
    int x = floori(p.pos[0]) - x0, y = floori(p.pos[1]) - y0;
    if (x < 0 || y < 0 || x >= w || y >= h) return;
    atomicAdd(&out[y * w + x], 1u);
}

// Live count and per-row counts / concentration sums (stats.comp:14-42).
__global__ __launch_bounds__(64) void k_stats(const TankBatch tb, uint32_t* counter) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const uint32_t RUN = 64u;
    uint32_t first = (blockIdx.x * 64 + threadIdx.x) * RUN;
    uint32_t n_slots = c.P->n_slots;
    if (first >= n_slots) return;
    uint32_t last = first + RUN < n_slots ? first + RUN : n_slots;
    uint32_t live = 0, cur = 0, n = 0, conc = 0;
    for (uint32_t id = first; id < last; id++) {
        uint32_t kind = c.part[id].kind & 255u;
        if (kind == 0u || kind > (uint32_t)SP_MAX_MATERIALS) continue;
        live++;
        if (kind != cur) {
            if (n > 0u) {
                atomicAdd(&counter[8u + 2u * (cur - 1u)], n);
                atomicAdd(&counter[9u + 2u * (cur - 1u)], conc);
            }
            cur = kind;
            n = 0u;
            conc = 0u;
        }
        n++;
        conc += (uint32_t)(clampf_(c.part[id].s, 0.0f, 1.0f) * 256.0f + 0.5f);
    }
    if (n > 0u) {
        atomicAdd(&counter[8u + 2u * (cur - 1u)], n);
        atomicAdd(&counter[9u + 2u * (cur - 1u)], conc);
    }
    if (live > 0u) atomicAdd(&counter[4], live);
}
