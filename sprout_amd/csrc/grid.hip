// grid.hip — the per-cell passes of a substep on gfx950: time step, accumulator clear,
// grid preparation, face coefficients, surface tension, the pressure right-hand side and
// the projection (the viscosity passes are in visc.hip).
//
// Every pass is one work-item per cell in 16 x 16 work-groups (four wave64s), the tile the
// pressure hierarchy skips by. They are stencil work with no contraction: HBM-bound, each
// reading a few float4 fields of the cell and its neighbours and writing one or two.
#include "dev.h"

namespace {

SP_DEV bool cell_of_thread(const Ctx& c, int& x, int& y, int& i) {
    x = blockIdx.x * 16 + threadIdx.x;
    y = blockIdx.y * 16 + threadIdx.y;
    i = y * c.W + x;
    return x < c.W && y < c.H;
}

}  // namespace

// step.comp:22-50: the substep length and the Kahan-summed simulated time.
__global__ void k_step(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sp_stepbuf* S = c.S;
    const sp_gparams& P = *c.P;
    float u = fmaxf_(fminf_(__uint_as_float(S->umax_grid), P.ext[0]), __uint_as_float(S->umax_part));
    u = fmaxf_(u, P.ext_speed);
    float dt_cfl = P.cfl / (u + sqrtf(5.0f * fmaxf_(P.gravity, 0.0f)) + 1e-6f);
    float dt;
    if (P.dt_fixed > 0.0f) {
        dt = P.dt_fixed;
    } else {
        uint32_t left = S->batch_steps > S->batch_done ? S->batch_steps - S->batch_done : 1u;
        dt = fminf_(fminf_(dt_cfl, P.dt_max), fmaxf_(S->frame_left, 0.0f) / (float)left);
    }
    dt = fmaxf_(dt, 1e-6f);
    S->dt_prev = S->dt > 0.0f ? S->dt : dt;
    S->dt = dt;
    S->dt_cfl = fminf_(dt_cfl, P.dt_max);
    S->umax_seen = u;
    S->frame_left -= dt;
    float lo = __uint_as_float(S->pad0);
    float yv = dt - lo;
    float t = S->advanced + yv;
    S->pad0 = __float_as_uint((t - S->advanced) - yv);
    S->advanced = t;
    S->batch_done += 1u;
    S->substeps += 1u;
    S->umax_grid = 0u;
    S->umax_part = 0u;
    *c.region_air = 0u;
}

// Clears n 32-bit words, 16 bytes per work-item per iteration.
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, size_t n, uint32_t v) {
    size_t n4 = n / 4;
    uint4* p4 = reinterpret_cast<uint4*>(p);
    uint4 v4 = make_uint4(v, v, v, v);
    for (size_t k = blockIdx.x * (size_t)256 + threadIdx.x; k < n4; k += (size_t)gridDim.x * 256) p4[k] = v4;
    size_t k = n4 * 4 + blockIdx.x * (size_t)256 + threadIdx.x;
    if (blockIdx.x == 0 && k < n) p[k] = v;
}

// Each tank's accumulator cleared (record_substep, mod.rs:2628): W H stride words in a
// grid-stride loop of 16-byte stores, the last few singly.
__global__ __launch_bounds__(256) void k_clear_acc(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const size_t n = (size_t)c.N * (size_t)c.P->acc_stride, n4 = n / 4;
    uint4* p4 = reinterpret_cast<uint4*>(c.acc);
    for (size_t k = blockIdx.x * (size_t)256 + threadIdx.x; k < n4; k += (size_t)gridDim.x * 256) p4[k] = make_uint4(0, 0, 0, 0);
    const size_t k = n4 * 4 + blockIdx.x * (size_t)256 + threadIdx.x;
    if (blockIdx.x == 0 && k < n) c.acc[k] = 0;
    // and the particles' heat sums of each cell while the tank is thermal (mod.rs:2629-2631)
    if (c.P->caps & SP_CAP_THERMAL) {
        const size_t h2 = (size_t)c.N;  // int2 per cell: N / 2 uint4 and a possible tail
        uint4* h4 = reinterpret_cast<uint4*>(c.hcell);
        for (size_t q = blockIdx.x * (size_t)256 + threadIdx.x; q < h2 / 2; q += (size_t)gridDim.x * 256) h4[q] = make_uint4(0, 0, 0, 0);
        if (blockIdx.x == 0 && threadIdx.x == 0 && (h2 & 1)) c.hcell[h2 - 1] = make_int2(0, 0);
    }
}

// Each tank's active viscosity tiles cleared before k_visc_setup marks them.
__global__ __launch_bounds__(256) void k_clear_vtiles(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int n = ((c.W + 15) >> 4) * ((c.H + 15) >> 4);
    for (int k = threadIdx.x; k < n; k += 256) c.vtiles[k] = 0u;
}

namespace {
SP_DEV bool packed(const Ctx& c, int i) {
    float4 sol = c.solid[i];
    float phi_f = (float)c.acc[(size_t)i * c.P->acc_stride + 4] / SP_MASS_SCALE * c.P->inv_ppc;
    return sol.w > 0.02f && phi_f > 0.04f && phi_f + sol.z > SP_LIQUID_THR;
}
}  // namespace

// grid_prep.comp:20-263, with the CAP-gated fields this build carries (heat, rigid bodies,
// rheology, viscoelasticity).
__global__ __launch_bounds__(256) void k_grid_prep(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    const sp_gparams& P = *c.P;
    const int W = c.W, H = c.H;
    const size_t k = (size_t)i * P.acc_stride;
    const float dt = c.S->dt;
    int um = c.acc[k], umom = c.acc[k + 1], vm = c.acc[k + 2], vmom = c.acc[k + 3];
    int cw = c.acc[k + 4];
    float4 sol = c.solid[i];
    float4 o = make_float4(0.0f, 0.0f, SP_FACE_EMPTY, SP_FACE_EMPTY);
    if (sol.x <= 0.0f) {
        o.z = SP_FACE_CLOSED;
    } else if (um >= SP_MIN_FACE_MASS) {
        o.x = ((float)umom / SP_MOM_SCALE) / ((float)um / SP_MASS_SCALE) / dt;
        o.z = 0.0f;
    }
    if (sol.y <= 0.0f) {
        o.w = SP_FACE_CLOSED;
    } else if (vm >= SP_MIN_FACE_MASS) {
        o.y = ((float)vmom / SP_MOM_SCALE) / ((float)vm / SP_MASS_SCALE) / dt - P.gravity * dt;
        o.w = 0.0f;
    }
    c.vel[i] = o;
    c.vel0[i] = make_float2(o.x, o.y);
    const int nmix = P.n_fam - 1;
    float phi_f = (float)cw / SP_MASS_SCALE * P.inv_ppc;
    float f[4] = {0, 0, 0, 0}, rg[4] = {0, 0, 0, 0}, lg[4] = {0, 0, 0, 0};
    float bg0 = 0.0f, bg1 = 0.0f, bg2 = 0.0f;
    float sigma = 0.0f, csx_ = 0.0f, csy_ = 0.0f, fam0 = -1.0f;
    float rho = 0.0f, mu = 0.0f, rho_c = 0.0f, nu_c = 0.0f;
    int major = 0, top_row = -1;
    float top_share = 0.0f, fv = 0.0f;  // fv: the share of steam
    const bool th_on = (P.caps & SP_CAP_THERMAL) != 0u;
    float T = P.heat_x, kappa = 0.0f, emis = 0.0f, frozen = 0.0f, Craw = 0.0f;
    float4 nn = make_float4(0.0f, -1.0f, 1.0f, 0.0f);  // rh: the cell's non-Newtonian material
    float chi = 0.0f;
    for (int q = 0; q < nmix; q++) c.mixb[(size_t)i * nmix + q] = -1.0f;
    const bool react = (P.caps & SP_CAP_REACT) != 0u;
    if (cw <= 0 && react) c.rmask[i] = 0u;  // grid_prep.comp:71
    // each family's occupancy of the cell for settling (grid_prep.comp:69-70, 147)
    const bool settle = (P.caps & SP_CAP_SETTLE) != 0u;
    if (settle)
        for (int q = 0; q < P.n_fam; q++) c.famocc[(size_t)i * P.n_fam + q] = 0.0f;
    if (cw > 0) {
        int others = 0;
        for (int j = 0; j < P.n_present; j++) {
            int ch = c.mats[P.present[j]].chan_w;
            if (ch >= 0) others += c.acc[k + ch];
        }
        float sh[SP_MAX_PRESENT], scn[SP_MAX_PRESENT];
        float Csh = 0.0f, Cg[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int j = 0; j < P.n_present; j++) {
            sh[j] = 0.0f;
            scn[j] = 1.0f;
            int r = P.present[j];
            const sp_gmat& m = c.mats[r];
            if (m.chan_w < 0 && P.primary >= 0 && r != P.primary) continue;
            int w = m.chan_w >= 0 ? c.acc[k + m.chan_w] : (cw - others > 0 ? cw - others : 0);
            if (w <= 0) continue;
            sh[j] = w >= cw ? 1.0f : clampf_((float)w / (float)cw, 0.0f, 1.0f);
            if (m.diluent >= 0 && m.chan_s >= 0) scn[j] = clampf_((float)c.acc[k + m.chan_s] / (float)w, 0.0f, 1.0f);
            // with heat: the cell's heat capacity, conductivity and emissivity (grid_prep.comp:100-108)
            if (th_on) {
                const float cr = sh[j] * thermal_of(c, r, scn[j]).w;
                Csh += cr;
                Cg[m.group] += cr;
                const float kp = m.th[1];
                kappa += sh[j] * (m.diluent >= 0 ? mixf_(c.mats[m.diluent].th[1], kp, scn[j]) : kp);
                const float em = m.th2[1];
                emis += sh[j] * (m.diluent >= 0 ? mixf_(c.mats[m.diluent].th2[1], em, scn[j]) : em);
            }
        }
        // which rows the cell holds, for the reactions (grid_prep.comp:110-118)
        if (react) {
            uint32_t mask = 0u;
            for (int j = 0; j < P.n_present; j++)
                if (sh[j] * (c.mats[P.present[j]].diluent >= 0 ? scn[j] : 1.0f) > SP_REACT_MIN) mask |= 1u << (uint32_t)j;
            c.rmask[i] = mask;
        }
        // the temperature of each phase group from its own sum; T leaves as the cell's mean
        // (group_temperatures, common.glsl:598-622)
        float Tg[4] = {T, T, T, T};
        if (th_on && Csh > 0.0f) {
            const int Htot = c.acc[k + P.chan_t];
            T = (float)Htot * SP_T_SCALE / (Csh * (float)cw) - 273.15f;
            for (int q = 0; q < 4; q++) Tg[q] = T;
            int rest = Htot, first = -1;
            for (int q = 0; q < 4; q++) {
                const int ch = group_heat_chan(P, q);
                if (ch >= 0) {
                    const int h = c.acc[k + ch];
                    rest -= h;
                    if (Cg[q] > 0.0f) Tg[q] = (float)h * SP_T_SCALE / (Cg[q] * (float)cw) - 273.15f;
                } else if (first < 0 && (P.group_mask & (1u << q)) != 0u) {
                    first = q;
                }
            }
            if (first >= 0 && Cg[first] > 0.0f) Tg[first] = (float)rest * SP_T_SCALE / (Cg[first] * (float)cw) - 273.15f;
        }
        for (int j = 0; j < P.n_present; j++) {
            float share = sh[j];
            if (share <= 0.0f) continue;
            int r = P.present[j];
            const sp_gmat& m = c.mats[r];
            float sc = scn[j];
            float rho_m = m.rho, lnu = m.ln_nu, sig = m.sigma;
            const float Tm = Tg[m.group];
            if (th_on) lnu += temp_law(c, r, Tm);
            const float xp = th_on ? expansion(c, r, sc, Tm) : 1.0f;
            float csx = m.cos_c, csy = m.sin_c;
            if (m.diluent >= 0) {
                const sp_gmat& d = c.mats[m.diluent];
                rho_m = mixf_(d.rho, rho_m, sc);
                float bs = m.visc_exp == 1.0f ? sc : powf(sc, m.visc_exp);
                lnu = mixf_(th_on ? d.ln_nu + temp_law(c, m.diluent, Tm) : d.ln_nu, lnu, bs);
                sig = mixf_(d.sigma, sig, sc);
                csx = mixf_(d.cos_c, csx, sc);
                csy = mixf_(d.sin_c, csy, sc);
                if (m.fam == 0) fam0 = sc;
                else if (m.fam > 0) c.mixb[(size_t)i * nmix + m.fam - 1] = sc;
                if (settle && m.fam >= 0) c.famocc[(size_t)i * P.n_fam + m.fam] = share * phi_f;
            }
            rho_m *= xp;
            // frozen liquid: the frozen viscosity and the solid's density (grid_prep.comp:150-159)
            if (th_on) {
                const float4 th = thermal_of(c, r, sc);
                const float fs = th.y > -1e8f ? clampf_((th.y - Tm) / fmaxf_(th.y - th.x, 1e-3f), 0.0f, 1.0f) : 0.0f;
                if (fs > 0.0f) {
                    lnu = mixf_(lnu, fmaxf_(lnu, P.ln_rigid), fs);
                    const float rs = m.rho_solid;
                    rho_m = mixf_(rho_m, m.diluent >= 0 ? mixf_(c.mats[m.diluent].rho_solid, rs, sc) : rs, fs);
                    frozen += share * fs;
                }
            }
            // the cell's non-Newtonian material: the one with the largest share (grid_prep.comp:162)
            if ((m.rheo != 0 || m.temp[3] > 0.0f || m.gr[0] > 0.0f) && share > nn.x) nn = make_float4(share, (float)r, sc, 0.0f);
            // steam's share, which never makes it the cell's top material (grid_prep.comp:163-167)
            if (r == SP_ROW_STEAM) {
                fv = share;
            } else if (share > top_share) {
                top_share = share;
                top_row = r;
            }
            chi += share * m.gr[3];  // the cell's susceptibility (grid_prep.comp:168)
            int g = m.group;
            f[g] += share;
            rg[g] += share * rho_m;
            if (g == 1) bg0 += share * m.rho;
            else if (g == 2) bg1 += share * m.rho;
            else if (g == 3) bg2 += share * m.rho;
            lg[g] += share * lnu;
            sigma += share * sig;
            csx_ += share * csx;
            csy_ += share * csy;
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
    bool has = phi_f > 0.04f && sol.w > 0.02f;
    float level = has ? phi_t - SP_LIQUID_THR : -SP_LIQUID_THR;
    if (has && level <= 0.0f && cx > 0 && cy > 0 && cx < W - 1 && cy < H - 1) {
        if ((packed(c, i - 1) && packed(c, i + 1)) || (packed(c, i - W) && packed(c, i + W))) level = 0.05f;
        else if (P.gravity > 0.0f && nn.x <= 0.0f && phi_t > 0.25f && packed(c, i - W)) level = 0.05f;
        // with a magnet, "lying on" is towards gravity and the last substep's magnetic pull
        // together (grid_prep.comp:211-221)
        else if ((P.caps & SP_CAP_MAGNET) && nn.x <= 0.0f && phi_t > 0.25f && chi > 0.0f) {
            const float km = chi / P.mag[1] / fmaxf_(rho, 0.05f);
            const float px = -km * 0.5f * (c.magc[i + 1].y - c.magc[i - 1].y) - 0.0f;
            const float py = -km * 0.5f * (c.magc[i + W].y - c.magc[i - W].y) - P.gravity;
            const int toward = fabsf(px) > fabsf(py) ? (px > 0.0f ? i + 1 : i - 1) : (py > 0.0f ? i + W : i - W);
            if (packed(c, toward)) level = 0.05f;
        }
    }
    if (sol.w > 0.0f && level <= 0.0f) {
        uint32_t bit = 1u << (c.region[i] & SP_REGION_MASK);
        if ((*c.region_air & bit) == 0u) atomicOr(c.region_air, bit);
    }
    // the viscosity solve stops at rigid_viscosity: a cell that would need more is stiff and
    // moves rigidly (grid_prep.comp:230-237)
    if (P.caps & (SP_CAP_RIGID | SP_CAP_RHEO)) {
        const float cap = rho * expf(P.ln_rigid);
        if (mu >= cap && level > 0.0f) nn.w = 1.0f;
        mu = fminf_(mu, cap);
        c.rh[i] = nn;
    }
    // the viscoelastic stress of the cell: its particles' stress sums, by volume
    // (grid_prep.comp:241-244)
    if (P.caps & SP_CAP_ELASTIC) {
        const float sc = P.el_g * P.inv_ppc / SP_STRESS_FIX;
        c.estr[i] = make_float4((float)c.acc[k + P.chan_e] * sc, (float)c.acc[k + P.chan_e + 1] * sc,
                                (float)c.acc[k + P.chan_e + 2] * sc, 0.0f);
    }
    if (P.caps & SP_CAP_MAGNET) c.magc[i].x = chi;  // grid_prep.comp:240
    c.cell[i] = make_float4(phi_f, rho, mu, level);
    c.comp[i] = make_float4(rho, f[1], fam0, phi_t);
    uint32_t gm = P.group_mask;
    if (group_sorts(P, c.mats, 1) || group_sorts(P, c.mats, 2) || group_sorts(P, c.mats, 3) || (P.caps & SP_CAP_STEAM))
        c.gbase[i] = make_float4(bg0, bg1, bg2, fv);
    if ((gm & (gm - 1u)) != 0u) {
        c.aux[i] = make_float4(nu_c, rho_c, f[2], f[3]);
        c.grho[i] = make_float4(rg[0], rg[1], rg[2], rg[3]);
    }
    c.sprop[i] = make_float4(sigma, csx_, csy_, (float)major + (float)(top_row + 1) / 128.0f);
    if (th_on) {
        c.therm[i] = make_float4(T, Craw, kappa, frozen);
        // the heat pass reads the emissivity here, then stores its result (grid_prep.comp:256-261)
        c.dheat[i] = make_float4(P.heat_rate > 0.0f ? emis : 0.0f, 0.0f, 0.0f, 0.0f);
    }
    if (level <= 0.0f) c.pres[i] = make_float2(0.0f, 0.0f);
}

namespace {
// Face coefficient of the variable-density solve (grid_coef.comp:12-25).
SP_DEV void face_coef(float4 a, float4 b, float open, float& g, float& rho) {
    float wsum = a.x + b.x;
    rho = wsum > 1e-6f ? (a.x * a.y + b.x * b.y) / wsum : 1.0f;
    g = 0.0f;
    bool fa = a.w > 0.0f, fb = b.w > 0.0f;
    if (open <= 0.0f || !(fa || fb)) return;
    float theta = 1.0f;
    if (fa != fb) {
        float lf = fa ? a.w : b.w;
        float la = fa ? b.w : a.w;
        theta = clampf_(lf / (lf - la), SP_THETA_MIN, 1.0f);
    }
    g = 1.0f / (fmaxf_(rho, 0.05f) * theta);
}
}  // namespace

// grid_coef.comp:27-38 fused with surf_smooth.comp:16-44 (both read only cell, comp,
// aux and solid, which grid_prep has finished).
__global__ __launch_bounds__(256) void k_coef_smooth(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    // the tile and a ring of two cells, each cell's smoothing inputs worked out once:
    // its open-weighted fill phi (-1 where it takes no part: outside the grid or closed),
    // phi times its group-1 share, and its group-2 and group-3 shares
    constexpr int S = 20;
    __shared__ float s_phi[S * S], s_py[S * S], s_az[S * S], s_aw[S * S];
    const bool more = (c.P->group_mask & 12u) != 0u;
    const int x0 = (int)blockIdx.x * 16 - 2, y0 = (int)blockIdx.y * 16 - 2;
    for (int k = (int)(threadIdx.y * 16 + threadIdx.x); k < S * S; k += 256) {
        const int nx = x0 + k % S, ny = y0 + k / S;
        float phi = -1.0f, py = 0.0f, az = 0.0f, aw = 0.0f;
        if (nx >= 0 && ny >= 0 && nx < c.W && ny < c.H) {
            const int j = ny * c.W + nx;
            const float open = c.solid[j].w;
            if (open > 0.02f) {
                phi = clampf_(c.cell[j].x / (SP_REST_PHI * open), 0.0f, 1.0f);
                py = phi * c.comp[j].y;
                if (more) {
                    const float4 ax = c.aux[j];
                    az = ax.z;
                    aw = ax.w;
                }
            }
        }
        s_phi[k] = phi;
        s_py[k] = py;
        s_az[k] = az;
        s_aw[k] = aw;
    }
    __syncthreads();
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    float4 o = make_float4(0.0f, 0.0f, 1.0f, 1.0f);
    if (cx > 0 && cy > 0) {
        float4 me = c.cell[i];
        float4 sol = c.solid[i];
        face_coef(me, c.cell[i - 1], sol.x, o.x, o.z);
        face_coef(me, c.cell[i - c.W], sol.y, o.y, o.w);
    }
    c.coef[i] = o;
    const float B5[5] = {1.0f, 4.0f, 6.0f, 4.0f, 1.0f};
    float sx = 0, sy = 0, s2x = 0, s2y = 0, wsum = 0;
    const int lx = (int)threadIdx.x + 2, ly = (int)threadIdx.y + 2;
#pragma unroll
    for (int b = -2; b <= 2; b++)
#pragma unroll
        for (int a = -2; a <= 2; a++) {
            const int k = (ly + b) * S + lx + a;
            const float phi = s_phi[k];
            if (phi < 0.0f) continue;
            float w = B5[a + 2] * B5[b + 2];
            sx += w * phi;
            sy += w * s_py[k];
            if (more) {
                float wp = w * phi;
                s2x += wp * s_az[k];
                s2y += wp * s_aw[k];
            }
            wsum += w;
        }
    float4 s = c.surf[i];
    s.x = wsum > 0.0f ? sx / wsum : 0.0f;
    s.y = wsum > 0.0f ? sy / wsum : 0.0f;
    c.surf[i] = s;
    if (more) {
        float4 s2 = c.surf2[i];
        s2.x = wsum > 0.0f ? s2x / wsum : 0.0f;
        s2.y = wsum > 0.0f ? s2y / wsum : 0.0f;
        c.surf2[i] = s2;
    }
}

namespace {
SP_DEV float sfield(const Ctx& c, int x, int y, int k) {
    x = clampi_(x, 0, c.W - 1);
    y = clampi_(y, 0, c.H - 1);
    int i = y * c.W + x;
    if (k < 2) return k == 0 ? c.surf[i].x : c.surf[i].y;
    return k == 2 ? c.surf2[i].x : c.surf2[i].y;
}

SP_DEV float node_sdf(const Ctx& c, int x, int y) {
    x = clampi_(x, 0, c.W);
    y = clampi_(y, 0, c.H);
    return c.sdf[y * (c.W + 1) + x];
}

// Contact angle of the liquid around node n (surf_curv.comp:49-58).
SP_DEV float2 contact_at(const Ctx& c, int nx, int ny) {
    float4 sp = c.sprop[clampi_(ny, 0, c.H - 1) * c.W + clampi_(nx, 0, c.W - 1)];
    float cx = sp.y, cy = sp.z;
    if (cx * cx + cy * cy > 1e-12f) {
        float l = len2_(cx, cy);
        return make_float2(cx / l, cy / l);
    }
    for (int b = -3; b <= 2; b++)
        for (int a = -3; a <= 2; a++) {
            int j = clampi_(ny + b, 0, c.H - 1) * c.W + clampi_(nx + a, 0, c.W - 1);
            float occ = c.cell[j].x;
            float4 spj = c.sprop[j];
            cx += spj.y * occ;
            cy += spj.z * occ;
        }
    if (cx * cx + cy * cy > 1e-12f) {
        float l = len2_(cx, cy);
        return make_float2(cx / l, cy / l);
    }
    return make_float2(0.0f, 1.0f);
}

// Outward normal of field k at node n, bent to the contact angle at walls
// (surf_curv.comp:60-91).
SP_DEV float2 normal_node(const Ctx& c, int nx, int ny, int k, float& mag) {
    const float G_MIN = 0.05f;
    float a = sfield(c, nx, ny, k), b = sfield(c, nx - 1, ny, k), d = sfield(c, nx, ny - 1, k), e = sfield(c, nx - 1, ny - 1, k);
    float gx = 0.5f * (a + d - b - e), gy = 0.5f * (a + b - d - e);
    mag = len2_(gx, gy);
    if (mag < G_MIN) return make_float2(0.0f, 0.0f);
    float2 nn = make_float2(-gx / mag, -gy / mag);
    if (k == 0) {
        float dd = node_sdf(c, nx, ny);
        if (dd < 0.5f) {
            float gwx = node_sdf(c, nx + 1, ny) - node_sdf(c, nx - 1, ny);
            float gwy = node_sdf(c, nx, ny + 1) - node_sdf(c, nx, ny - 1);
            float lw = len2_(gwx, gwy);
            if (lw > 1e-4f) {
                float nwx = gwx / lw, nwy = gwy / lw;
                float gdn = gx * nwx + gy * nwy;
                float gtx = gx - gdn * nwx, gty = gy - gdn * nwy;
                float along = len2_(gtx, gty);
                if (along > 0.5f * G_MIN) {
                    float tx = -gtx / along, ty = -gty / along;
                    float2 cs = contact_at(c, nx, ny);
                    float ncx = cs.x * nwx + cs.y * tx, ncy = cs.x * nwy + cs.y * ty;
                    float sm = smoothstep_(0.5f * G_MIN, G_MIN, along);
                    float mx = mixf_(nn.x, ncx, sm), my = mixf_(nn.y, ncy, sm);
                    float lm = len2_(mx, my);
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

SP_DEV float curvature(const Ctx& c, int cx, int cy, int k, float kmax) {
    float m00, m10, m01, m11;
    float2 n00 = normal_node(c, cx, cy, k, m00);
    float2 n10 = normal_node(c, cx + 1, cy, k, m10);
    float2 n01 = normal_node(c, cx, cy + 1, k, m01);
    float2 n11 = normal_node(c, cx + 1, cy + 1, k, m11);
    if (0.25f * (m00 + m10 + m01 + m11) < 0.05f) return 0.0f;
    float kap = 0.5f * (n10.x + n11.x - n00.x - n01.x) + 0.5f * (n01.y + n11.y - n00.y - n10.y);
    return kmax > 0.0f ? clampf_(kap, -kmax, kmax) : kap;
}
}  // namespace

// surf_curv.comp:104-121.
__global__ __launch_bounds__(256) void k_surf_curv(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    float kmax = c.P->kappa_max;
    uint32_t gm = c.P->group_mask;
    float k0 = curvature(c, cx, cy, 0, kmax);
    float k1 = curvature(c, cx, cy, 1, kmax);
    float4 s = c.surf[i];
    s.z = k0;
    s.w = k1;
    c.surf[i] = s;
    if (gm & 12u) {
        float4 s2 = c.surf2[i];
        s2.z = (gm & 4u) ? curvature(c, cx, cy, 2, kmax) : 0.0f;
        s2.w = (gm & 8u) ? curvature(c, cx, cy, 3, kmax) : 0.0f;
        c.surf2[i] = s2;
    }
}

namespace {
SP_DEV float air_jump(const Ctx& c, int a, int b) {
    bool la = c.cell[a].w > 0.0f;
    int liq = la ? a : b, air = la ? b : a;
    float lf = c.cell[liq].w, lgv = c.cell[air].w;
    float theta = clampf_(lf / (lf - lgv), SP_THETA_MIN, 1.0f);
    float kl = c.surf[liq].z, kg = c.surf[air].z;
    float kap = kg != 0.0f ? mixf_(kl, kg, theta) : kl;
    return c.S->dt * c.sprop[liq].x * kap;
}

SP_DEV float share_of(const Ctx& c, int i, int g) {
    if (g == 1) return c.comp[i].y;
    if (g == 2) return c.aux[i].z;
    if (g == 3) return c.aux[i].w;
    return 1.0f - c.comp[i].y - c.aux[i].z - c.aux[i].w;
}

SP_DEV float curv_of(const Ctx& c, int i, int g) { return g == 1 ? c.surf[i].w : g == 2 ? c.surf2[i].z : c.surf2[i].w; }

SP_DEV float pair_jump(const Ctx& c, int ci, int n) {
    int gc = (int)c.sprop[ci].w, gn = (int)c.sprop[n].w;
    if (gc == gn) return 0.0f;
    int hi = gc > gn ? gc : gn, lo = gc < gn ? gc : gn;
    float dc = share_of(c, ci, hi) - share_of(c, ci, lo), dn = share_of(c, n, hi) - share_of(c, n, lo);
    if (dc * dn >= 0.0f) return 0.0f;
    float theta = -dn / (dc - dn);
    float kc = curv_of(c, ci, hi), kn = curv_of(c, n, hi);
    float kap = kc == 0.0f ? kn : kn == 0.0f ? kc : mixf_(kn, kc, theta);
    float j = c.S->dt * c.P->pair[hi][lo] * kap;
    return gc == hi ? j : -j;
}

SP_DEV float face_jump(const Ctx& c, int ci, int n, float open) {
    if (open <= 0.0f) return 0.0f;
    bool lc = c.cell[ci].w > 0.0f, ln = c.cell[n].w > 0.0f;
    if (lc != ln) return air_jump(c, ci, n);
    if (lc && ln) return pair_jump(c, ci, n);
    return 0.0f;
}
}  // namespace

// surf_apply.comp:71-82.
__global__ __launch_bounds__(256) void k_surf_apply(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    float2 pj = make_float2(0.0f, 0.0f);
    if (cx > 0 && cy > 0) {
        float4 sol = c.solid[i];
        pj.x = face_jump(c, i, i - 1, sol.x);
        pj.y = face_jump(c, i, i - c.W, sol.y);
    }
    c.pjump[i] = pj;
}

// Face states of the cell's u and v faces (vel.zw: live, or the FACE_EMPTY / FACE_CLOSED
// sentinels), used by the viscosity kernels.
namespace {
SP_DEV bool live_u(const Ctx& c, int i) { return face_live(c.vel[i].z); }
SP_DEV bool live_v(const Ctx& c, int i) { return face_live(c.vel[i].w); }
SP_DEV bool closed_u(const Ctx& c, int i) { return face_closed(c.vel[i].z); }
SP_DEV bool closed_v(const Ctx& c, int i) { return face_closed(c.vel[i].w); }
SP_DEV float vel_u(const Ctx& c, int i) { return live_u(c, i) ? c.vel[i].x : 0.0f; }
SP_DEV float vel_v(const Ctx& c, int i) { return live_v(c, i) ? c.vel[i].y : 0.0f; }
}  // namespace

// rhs.comp:19-93: the level-0 matrix, both right-hand sides, and the tile flags.
__global__ __launch_bounds__(256) void k_rhs(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    const bool in = cell_of_thread(c, cx, cy, i);
    __shared__ int s_any;
    if (threadIdx.x == 0 && threadIdx.y == 0) s_any = 0;
    __syncthreads();
    if (in) {
        const sp_gparams& P = *c.P;
        const int W = c.W, H = c.H;
        float4 me = c.cell[i];
        if (me.w <= 0.0f || cx == 0 || cy == 0 || cx == W - 1 || cy == H - 1) {
            c.mga[i] = make_float4(0, 0, 0, 0);
            c.mgd[i] = 0.0f;
            c.prhs[i] = make_float4(0, 0, 0, 0);
        } else {
            const float dt = c.S->dt;
            float4 sol = c.solid[i];
            float4 cf = c.coef[i];
            float aL = sol.x, aB = sol.y, aR = c.solid[i + 1].x, aT = c.solid[i + W].y;
            float4 w = make_float4(aL * cf.x, aR * c.coef[i + 1].x, aB * cf.y, aT * c.coef[i + W].y);
            float diag = w.x + w.y + w.z + w.w;
            uint32_t rg = c.region[i];
            bool sealed = (*c.region_air & (1u << (rg & SP_REGION_MASK))) == 0u;
            if ((rg & SP_REGION_VENT) != 0u && sealed && P.sealed_vent > 0.0f) diag += 1.0f / fmaxf_(me.y, 0.05f);
            float4 vme = c.vel[i];
            float div = aR * c.vel[i + 1].x - aL * vme.x + aT * c.vel[i + W].y - aB * vme.y;
            bool lR = c.cell[i + 1].w > 0.0f, lT = c.cell[i + W].w > 0.0f;
            float2 pji = c.pjump[i];
            float pjr = c.pjump[i + 1].x, pjt = c.pjump[i + W].y;
            float jump = w.x * pji.x + w.y * (lR ? -pjr : pjr) + w.z * pji.y + w.w * (lT ? -pjt : pjt);
            float err = clampf_(c.comp[i].w - SP_REST_PHI, -0.5f, 0.5f);
            {
                bool interior = (aL <= 0.0f || c.cell[i - 1].w > 0.0f) && (aR <= 0.0f || c.cell[i + 1].w > 0.0f)
                    && (aB <= 0.0f || c.cell[i - W].w > 0.0f) && (aT <= 0.0f || c.cell[i + W].w > 0.0f);
                bool pressed = c.pres[i].x * (dt / c.S->dt_prev) > 0.5f * me.y * P.gravity * dt;
                if (interior && !pressed) err = fmaxf_(err - P.corr[1], 0.0f);
                else if (!interior && err < 0.0f) err = 0.0f;
            }
            float gain = fminf_(1.0f - expf(-dt / P.tol[3]), P.density_k);
            float sv = gain * err;
            if (sealed && P.sealed_vent > 0.0f) sv = 0.0f;
            if (diag > 1e-9f) s_any = 1;
            c.mga[i] = diag > 1e-9f ? w : make_float4(0, 0, 0, 0);
            c.mgd[i] = diag > 1e-9f ? diag : 0.0f;
            c.prhs[i] = make_float4(-div + jump, sv / dt, 0.0f, 0.0f);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && threadIdx.y == 0) c.ptile[blockIdx.y * gridDim.x + blockIdx.x] = s_any ? 1u : 0u;
}

namespace {
SP_DEV float difference(bool lc, bool ln, float pc, float pn, float j) {
    if (lc && ln) return pc - pn - j;
    return (lc ? pc : j) - (ln ? pn : j);
}
}  // namespace

// project.comp:26-58.
__global__ __launch_bounds__(256) void k_project(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    const bool in = cell_of_thread(c, cx, cy, i);
    float m = 0.0f;
    if (in && cx >= 1 && cy >= 1) {
        float4 cf = c.coef[i];
        float4 me = c.vel[i];
        if (c.mgd[i] > 0.0f) c.prhs[i].z = c.pres[i].x;
        if (cf.x != 0.0f || cf.y != 0.0f) {
            float2 p = c.pres[i];
            bool liq = c.cell[i].w > 0.0f;
            if (face_live(me.z) && cf.x != 0.0f) {
                int n = i - 1;
                me.x -= cf.x * difference(liq, c.cell[n].w > 0.0f, p.x, c.pres[n].x, c.pjump[i].x);
                me.z = -cf.x * (p.y - c.pres[n].y);
            }
            if (face_live(me.w) && cf.y != 0.0f) {
                int n = i - c.W;
                me.y -= cf.y * difference(liq, c.cell[n].w > 0.0f, p.x, c.pres[n].x, c.pjump[i].y);
                me.w = -cf.y * (p.y - c.pres[n].y);
            }
            c.vel[i] = me;
        }
        if (face_live(me.z) && cf.x != 0.0f) m = fabsf(me.x);
        if (face_live(me.w) && cf.y != 0.0f) m = fmaxf_(m, fabsf(me.y));
    }
    __shared__ float s_max[4];
    group_max_to<256>(&c.S->umax_grid, m, s_max);
}
