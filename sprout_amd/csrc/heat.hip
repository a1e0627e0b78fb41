// heat.hip — heat conduction on the grid (heat.comp), one work-item per cell on 16 x 16
// tiles, for a thermal tank (CAP_THERMAL) with a heat rate.
//
// grid_prep left each cell's temperature, heat capacity, conductivity and frozen share in
// therm and its emissivity in dheat.x; the particle pass left the heat sums of the particles
// whose position lies in each cell in hcell. Per cell this works out the energy it gains in
// the substep (conduction from liquid neighbours with the harmonic mean of the two
// conductivities, sub-grid eddy mixing within one liquid, the boundary-layer exchange
// between different liquids in a form that cannot overshoot, the air film, grey-body
// radiation and evaporation at a free surface, solids held at a temperature), and stores it
// as the change of enthalpy temperature the particles gather next substep, with the share
// and mean of the exchange among the particles of the cell.
//
// Exchange with air and solids goes into the ledger (64-bit fixed point). The reference adds
// every cell's quanta with 32-bit atomics and a carry; the sum of integer quanta does not
// depend on their order, so here each wave64 sums its quanta in 64 bits first and adds one
// 64-bit atomic per entry: the same ledger, 64 times fewer atomics on eight words.
#include "dev.h"

namespace {

constexpr float HEAT_PHI_MIN = 0.04f;  // heat.comp:57
constexpr float HEAT_PR_T = 0.85f;     // heat.comp:58: turbulent Prandtl number

// Mass fraction of water vapour in saturated air at T, deg C (Tetens, heat.comp:66-70).
SP_DEV float y_sat(float T) {
    T = clampf_(T, -40.0f, 99.0f);
    const float p = 0.61078f * expf(17.27f * T / (T + 237.3f));
    return 0.622f * p / (101.325f - 0.378f * p);
}

// Velocity at a cell's centre from its live faces (heat.comp:73-79).
SP_DEV float2 centre_vel(const Ctx& c, int i) {
    const int last = c.W * c.H - 1;
    const float4 v = c.vel[i], vr = c.vel[i + 1 < last ? i + 1 : last], vt = c.vel[i + c.W < last ? i + c.W : last];
    const float ax = face_live(v.z) ? v.x : 0.0f, ay = face_live(v.w) ? v.y : 0.0f;
    const float bx = face_live(vr.z) ? vr.x : 0.0f, by = face_live(vt.w) ? vt.y : 0.0f;
    return make_float2(0.5f * (ax + bx), 0.5f * (ay + by));
}

// The leading material of cell i, + 1 (0: none), as grid_prep left it in sprop.w.
SP_DEV int lead_of(const Ctx& c, int i) {
    const float w = c.sprop[i].w;
    return (int)((w - floorf(w)) * 128.0f + 0.5f);
}

// The ledger quantum of v >= 0 (ledger_add, common.glsl:446-451).
SP_DEV uint64_t ledger_quantum(float v) { return (uint64_t)(uint32_t)(v * SP_LEDGER_SCALE + 0.5f); }

SP_DEV uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace

// heat.comp:81-198.
__global__ __launch_bounds__(256) void k_heat(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    const bool in = cx < c.W && cy < c.H;
    const sp_gparams& P = *c.P;
    const int W = c.W, H = c.H;
    const int i = cy * W + cx;
    // ledger quanta of this work-item: air gained / lost, solids gained / lost
    uint64_t q0 = 0, q1 = 0, q2 = 0, q3 = 0;
    if (in) {
        const float k = c.S->dt * P.heat_rate;
        const float4 me = c.therm[i];
        const float emis = c.dheat[i].x;
        float phi = c.cell[i].x;
        // exchange inside the cell (heat.comp:88-106)
        const int2 hsum = c.hcell[i];
        float own_x = 0.0f, own_y = 0.0f;
        if (hsum.x > 0 && me.y > 0.0f && phi > 0.0f) {
            float rate = 5.78f * me.z * 3.14159265f / (P.inv_ppc * me.y * P.inv_ppc / phi);
            // a cell of aqueous liquid that holds steam mixes far faster (its bubbles condense
            // on the water around them), unless another liquid shares the cell (heat.comp:95-104)
            if ((P.caps & SP_CAP_STEAM) && c.gbase[i].w > 0.0f) {
                const bool others = (P.group_mask & ~1u) != 0u && c.comp[i].y + c.aux[i].z + c.aux[i].w > 1e-3f;
                if (!others) rate = fmaxf_(rate, P.steam2[0]);
            }
            own_x = 1.0f - expf(-k * rate);
            own_y = (float)hsum.y / (float)hsum.x;
        }
        if (phi <= HEAT_PHI_MIN || me.y <= 0.0f || cx < 1 || cy < 1 || cx >= W - 1 || cy >= H - 1) {
            c.dheat[i] = make_float4(0.0f, own_x, own_y, 0.0f);
        } else {
            phi = fminf_(phi, 1.0f);
            const int nb[4] = {i - 1, i + 1, i - W, i + W};
            const float open[4] = {c.solid[i].x, c.solid[i + 1].x, c.solid[i].y, c.solid[i + W].y};
            const float Tk = me.x + 273.15f, Ta = P.heat_x + 273.15f;
            float cond = 0.0f, air = 0.0f, wall = 0.0f, contact = 0.0f;
            const float cpc = me.y * P.inv_ppc;  // the cell's heat capacity, cell volumes of water
            const float hc = P.contact;
            const int lead = lead_of(c, i);
            // eddy mixing: (C_s dx)^2 / Pr_t times the strain rate at each face, at most what
            // keeps the explicit step stable (heat.comp:120-128)
            const float eddy_l2 = P.eddy * P.eddy / HEAT_PR_T;
            const bool eddy = eddy_l2 > 0.0f && me.w < 0.5f;
            const float2 u_me = eddy ? centre_vel(c, i) : make_float2(0.0f, 0.0f);
            const float eddy_max = 0.2f / fmaxf_(k, 1e-9f);
            const float per_rate = 1.0f / fmaxf_(P.heat_rate, 1e-6f);
            const float solid_ti = c.solid_t[i];
            float evap = 0.0f;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int n = nb[q];
                const float a = open[q];
                if (a > 0.0f) {
                    const float pn = c.cell[n].x;
                    const float4 o = c.therm[n];
                    // liquid against liquid through the part both fill; against air through
                    // the part only this cell fills
                    const float pl = (pn > HEAT_PHI_MIN && o.y > 0.0f) ? fminf_(pn, 1.0f) : 0.0f;
                    if (pl > 0.0f) {
                        float gc = 2.0f * me.z * o.z / fmaxf_(me.z + o.z, 1e-12f);
                        const int ln = lead_of(c, n);
                        if (eddy && o.w < 0.5f && ln == lead) {
                            const float2 un = centre_vel(c, n);
                            const float strain = 1.4142f * len2_(un.x - u_me.x, un.y - u_me.y);
                            const float alpha = fminf_(eddy_l2 * strain * per_rate, eddy_max);
                            const float c1 = cpc / phi, c2 = o.y * P.inv_ppc / pl;
                            gc += alpha * 2.0f * c1 * c2 / fmaxf_(c1 + c2, 1e-12f);
                        }
                        cond += a * fminf_(phi, pl) * gc * (o.x - me.x);
                        if (hc > 0.0f && ln != lead) {
                            // each half cell conducts at least as well as the boundary layer
                            const float g1 = fmaxf_(2.0f * me.z, hc), g2 = fmaxf_(2.0f * o.z, hc);
                            const float gx = fmaxf_(g1 * g2 / (g1 + g2) - gc, 0.0f) * a * fminf_(phi, pl);
                            const float cn = o.y * P.inv_ppc;
                            const float x = gx * k * (1.0f / cpc + 1.0f / cn);
                            contact += 0.25f * (1.0f - expf(-4.0f * x)) * cpc * cn / (cpc + cn) * (o.x - me.x);
                        }
                    }
                    const float exposed = fmaxf_(phi - pl, 0.0f);
                    if (exposed > 0.0f && c.solid[n].w > 0.02f) {
                        // free surface: film + radiation, linearised about the two temperatures
                        const float h = P.film + emis * P.radiation * (Ta * Ta + Tk * Tk) * (Ta + Tk);
                        air += a * exposed * h * (P.heat_x - me.x);
                        // evaporation from an aqueous surface, never below the air's own
                        // saturation
                        if (P.evap > 0.0f && lead > 0 && c.mats[lead - 1].group == 0 && c.mats[lead - 1].th2[2] < 150.0f && me.w < 0.5f)
                            evap += a * exposed * P.film * P.evap * fmaxf_(y_sat(me.x) - y_sat(P.heat_x), 0.0f);
                    }
                }
                if (a < 1.0f) {
                    const float ts = fmaxf_(c.solid_t[n], solid_ti);
                    if (ts > 0.5f * SP_NO_TEMP) wall += (1.0f - a) * phi * 2.0f * me.z * (ts - me.x);
                }
            }
            if (evap > 0.0f) air -= evap;
            air *= k;
            wall *= k;
            const float lim = cpc * fabsf(P.heat_x - me.x);
            air = clampf_(air, -lim, lim);
            const float e = k * cond + contact + air + wall;
            c.dheat[i] = make_float4(e / cpc, own_x, own_y, 0.0f);
            if (air > 0.0f) q0 = ledger_quantum(air);
            else q1 = ledger_quantum(-air);
            if (wall > 0.0f) q2 = ledger_quantum(wall);
            else q3 = ledger_quantum(-wall);
        }
    }
    // the ledger, a wave's quanta at a time (every work-item of the wave takes part)
    q0 = wave_sum_u64(q0);
    q1 = wave_sum_u64(q1);
    q2 = wave_sum_u64(q2);
    q3 = wave_sum_u64(q3);
    if (((threadIdx.y * 16 + threadIdx.x) & 63) == 0) {
        unsigned long long* L = reinterpret_cast<unsigned long long*>(c.ledger);
        if (q0) atomicAdd(&L[0], (unsigned long long)q0);
        if (q1) atomicAdd(&L[1], (unsigned long long)q1);
        if (q2) atomicAdd(&L[2], (unsigned long long)q2);
        if (q3) atomicAdd(&L[3], (unsigned long long)q3);
    }
}
