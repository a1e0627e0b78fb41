// visc.hip — implicit viscosity on gfx950: backward Euler on the full stress tensor
// (visc.glsl:1-18), A = M + G^T W G on the live u and v faces of the tiles whose diffusion
// number reaches implicit_viscosity_min, solved by conjugate gradients preconditioned by an
// aggregation multigrid V-cycle over both components at once (vmg.glsl) or by its diagonal;
// then the second pressure solve's right-hand side and projection (rhs_visc.comp,
// project_visc.comp). Record order: record_viscosity_split / record_viscosity_solve,
// mod.rs:2805-2927.
//
// Every kernel works on one 16 x 16 tile per 256-work-item group (the tiles are the
// viscosity's unit of activity). The multigrid level kernels stage the tile and 2 NU (+1)
// rings of cells in LDS, matrix, right-hand side and iterate, and smooth there, each
// half-sweep on a ring one cell narrower than the last, so that a V-cycle level is one
// launch whatever the number of sweeps; the cells of the rings are recomputed by each tile
// that needs them. The solve's scalars live in device memory (sp_solve, FP64) and are
// advanced by one-group reduction kernels over per-tile FP64 partial sums in a fixed order;
// every kernel of an iteration reads `v_done` at entry and returns when the solve has
// converged, so a batch records a fixed number of iterations.
//
// Face helpers (live_u, closed_u, vel_u ...) come from grid.hip, level helpers (Lv, level,
// inside, at) and reductions (group_sum4, total) from solve.hip: the unity build includes
// this file after both.
#include "dev.h"

namespace {

SP_DEV bool v_interior(const Ctx& c, int x, int y) { return x >= 1 && y >= 1 && x < c.W - 1 && y < c.H - 1; }
SP_DEV int vtile_of(const Ctx& c, int x, int y) { return (y >> 4) * ((c.W + 15) >> 4) + (x >> 4); }
SP_DEV bool v_on(const Ctx& c, int x, int y) { return c.vtiles[vtile_of(c, x, y)] != 0u; }
SP_DEV bool v_done(const Ctx& c) { return c.K->v_done != 0u || c.K->v_halt != 0u; }
// GLSL max(x, y): y if x < y, otherwise x.
SP_DEV float gmax_(float x, float y) { return x < y ? y : x; }
SP_DEV double dot2d(float2 a, float2 b) { return (double)a.x * (double)b.x + (double)a.y * (double)b.y; }
SP_DEV float2 live2(const Ctx& c, int i) {
    return make_float2(live_u(c, i) ? 1.0f : 0.0f, live_v(c, i) ? 1.0f : 0.0f);
}

// Strain of the current velocities with weights (wx, wy) (visc.glsl:47-63).
SP_DEV float4 strain_of_vel(const Ctx& c, int i, float wx, float wy) {
    const int W = c.W;
    float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (wx > 0.0f) {
        o.x = wx * (vel_u(c, i + 1) - vel_u(c, i));
        o.y = wx * (vel_v(c, i + W) - vel_v(c, i));
    }
    if (wy > 0.0f) {
        float ua = vel_u(c, i), ub = vel_u(c, i - W), va = vel_v(c, i), vb = vel_v(c, i - 1);
        if (closed_u(c, i) && live_u(c, i - W)) ua = -ub;
        if (closed_u(c, i - W) && live_u(c, i)) ub = -ua;
        if (closed_v(c, i) && live_v(c, i - 1)) va = -vb;
        if (closed_v(c, i - 1) && live_v(c, i)) vb = -va;
        o.z = wy * 0.5f * ((ua - ub) + (va - vb));
    }
    return o;
}

// G^T of the weighted strain field st at the faces of cell i (visc_apply.comp:16-28).
SP_DEV float2 gts_of(const Ctx& c, const float4* __restrict__ st, int i) {
    const int W = c.W;
    float2 g = make_float2(0.0f, 0.0f);
    if (live_u(c, i)) {
        float c1 = closed_u(c, i - W) ? 1.0f : 0.5f, c2 = closed_u(c, i + W) ? -1.0f : -0.5f;
        g.x = st[i - 1].x - st[i].x + st[i].z * c1 + st[i + W].z * c2;
    }
    if (live_v(c, i)) {
        float d1 = closed_v(c, i - 1) ? 1.0f : 0.5f, d2 = closed_v(c, i + 1) ? -1.0f : -0.5f;
        g.y = st[i - W].y - st[i].y + st[i].z * d1 + st[i + 1].z * d2;
    }
    return g;
}

}  // namespace

// visc_setup.comp:14-58: strain weights at the cell centre (2 dt mu) and its lower-left node
// (4 dt mu V), and the strain of the current velocities with them; with `mark`, the tiles
// within one cell of a cell whose diffusion number reaches implicit_viscosity_min (the tile
// flags are cleared before; each group collects the marks for its tile and the eight around
// it in LDS and stores them once).
__global__ __launch_bounds__(256) void k_visc_setup(const TankBatch tb, int mark) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ uint32_t s_mark[9];
    const int t = tid2();
    if (mark && t < 9) s_mark[t] = 0u;
    if (mark) __syncthreads();
    int cx, cy, i;
    if (cell_of_thread(c, cx, cy, i)) {
        const int W = c.W, H = c.H;
        const float dt = c.S->dt;
        float wc = 0.0f, wn = 0.0f;
        float4 me = c.cell[i];
        if (cx >= 1 && cy >= 1 && cx < W - 1 && cy < H - 1) {
            auto ok_u = [&](int j) { return live_u(c, j) || closed_u(c, j); };
            auto ok_v = [&](int j) { return live_v(c, j) || closed_v(c, j); };
            if (me.w > 0.0f && ok_u(i) && ok_u(i + 1) && ok_v(i) && ok_v(i + W)) wc = 2.0f * dt * me.z;
            int cs4[4] = {i - W - 1, i - W, i - 1, i};
            int nliq = 0;
            bool any_air = false;
            float mu = 0.0f;
            for (int k = 0; k < 4; k++) {
                float4 q = c.cell[cs4[k]];
                if (q.w > 0.0f) {
                    nliq++;
                    mu += q.z;
                } else if (c.solid[cs4[k]].w > 0.02f) {
                    any_air = true;
                }
            }
            if (nliq > 0 && !any_air && ok_u(i) && ok_u(i - W) && ok_v(i) && ok_v(i - 1))
                wn = 4.0f * dt * (mu / (float)nliq) * ((float)nliq / 4.0f);
        }
        float4 vw = c.vw[i];
        vw.x = wc;
        vw.y = wn;
        c.vw[i] = vw;
        c.vs[i] = strain_of_vel(c, i, wc, wn);
        if (mark && me.w > 0.0f && dt * me.z / fmaxf_(me.y, 0.05f) >= c.P->corr[3]) {
            const int bx = blockIdx.x, by = blockIdx.y;
            for (int b = -1; b <= 1; b++)
                for (int a = -1; a <= 1; a++) {
                    int nx = clampi_(cx + a, 0, W - 1), ny = clampi_(cy + b, 0, H - 1);
                    s_mark[((ny >> 4) - by + 1) * 3 + ((nx >> 4) - bx + 1)] = 1u;
                }
        }
    }
    if (!mark) return;
    __syncthreads();
    if (t < 9 && s_mark[t]) {
        int tx = (int)blockIdx.x + t % 3 - 1, ty = (int)blockIdx.y + t / 3 - 1;
        c.vtiles[ty * gridDim.x + tx] = 1u;
    }
}

// visc_apply.comp mode 2 (visc_apply.comp:16-48): the explicit viscous step everywhere.
__global__ __launch_bounds__(256) void k_visc_explicit(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    if (!v_interior(c, cx, cy)) return;
    float2 g = gts_of(c, c.vs, i);
    float4 v = c.vel[i];
    float4 cf = c.coef[i];
    if (face_live(v.z)) v.x -= g.x / cf.z;
    if (face_live(v.w)) v.y -= g.y / cf.w;
    c.vel[i] = v;
}

// visc_split.comp:24-38: the explicit step in the tiles the implicit solve leaves out.
__global__ __launch_bounds__(256) void k_visc_split(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (c.vtiles[tile_slot()] != 0u) return;
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    if (!v_interior(c, cx, cy)) return;
    float2 g = gts_of(c, c.vs, i);
    float4 v = c.vel[i];
    float4 cf = c.coef[i];
    if (face_live(v.z)) v.x -= g.x / cf.z;
    if (face_live(v.w)) v.y -= g.y / cf.w;
    c.vel[i] = v;
}

namespace {

// Strain weights as the implicit system sees them, and its unknowns (visc_diag.comp:75-81).
SP_DEV float2 v_wts(const Ctx& c, int x, int y) {
    if (!v_on(c, x, y)) return make_float2(0.0f, 0.0f);
    float4 w = c.vw[y * c.W + x];
    return make_float2(w.x, w.y);
}
SP_DEV bool unk_u(const Ctx& c, int x, int y) { return v_interior(c, x, y) && v_on(c, x, y) && live_u(c, y * c.W + x); }
SP_DEV bool unk_v(const Ctx& c, int x, int y) { return v_interior(c, x, y) && v_on(c, x, y) && live_v(c, y * c.W + x); }

// The couplings of A at cell (x, y) (visc_diag.comp:83-117).
SP_DEV void couplings(const Ctx& c, int x, int y, float4& a, float4& b) {
    const int W = c.W, i = y * W + x;
    const bool u = unk_u(c, x, y), v = unk_v(c, x, y);
    a = make_float4(0, 0, 0, 0);
    b = make_float4(0, 0, 0, 0);
    if (u && unk_u(c, x - 1, y)) a.x = -v_wts(c, x - 1, y).x;
    if (v && unk_v(c, x, y - 1)) a.w = -v_wts(c, x, y - 1).x;
    float wn = v_wts(c, x, y).y;
    if (wn > 0.0f) {
        float ua = u ? (closed_u(c, i - W) ? 1.0f : 0.5f) : 0.0f;
        float ub = unk_u(c, x, y - 1) ? -(closed_u(c, i) ? 1.0f : 0.5f) : 0.0f;
        float va = v ? (closed_v(c, i - 1) ? 1.0f : 0.5f) : 0.0f;
        float vb = unk_v(c, x - 1, y) ? -(closed_v(c, i) ? 1.0f : 0.5f) : 0.0f;
        a.y = wn * ua * ub;
        a.z = wn * va * vb;
        b.x = wn * ua * va;
        b.y = wn * ua * vb;
    }
    float wu = v_wts(c, x, y + 1).y;
    if (wu > 0.0f && u) {
        float ub = -(closed_u(c, i + W) ? 1.0f : 0.5f);
        float va = unk_v(c, x, y + 1) ? (closed_v(c, i + W - 1) ? 1.0f : 0.5f) : 0.0f;
        float vb = unk_v(c, x - 1, y + 1) ? -(closed_v(c, i + W) ? 1.0f : 0.5f) : 0.0f;
        b.z = wu * ub * va;
        b.w = wu * ub * vb;
    }
}

}  // namespace

// visc_diag.comp:121-184: the Jacobi diagonal (vw.zw), last substep's change rescaled as
// the warm start (vd), the inactive tiles taken out of the system, and (multigrid) the matrix
// of level 0 in the form the V-cycle keeps it. A tile that has just gone inactive has its
// matrix and multigrid vectors cleared once.
__global__ __launch_bounds__(256) void k_visc_diag(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ uint32_t s_was;
    const int slot = tile_slot();
    const bool mg = c.P->v_precond != 0;
    if (mg && tid2() == 0) {
        uint32_t w = c.vtile[slot];
        c.vtile[slot] = c.vtiles[slot];
        s_was = w;
    }
    __syncthreads();
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    const int W = c.W;
    const bool on = c.vtiles[slot] != 0u;
    const bool lu = live_u(c, i), lv = live_v(c, i);
    const bool inner = v_interior(c, cx, cy);
    float2 d = make_float2(1.0f, 1.0f);
    if (inner) {
        float4 cf = c.coef[i];
        float2 me = v_wts(c, cx, cy);
        if (lu) {
            float c1 = closed_u(c, i - W) ? 1.0f : 0.5f, c2 = closed_u(c, i + W) ? 1.0f : 0.5f;
            d.x = cf.z + v_wts(c, cx - 1, cy).x + me.x + me.y * c1 * c1 + v_wts(c, cx, cy + 1).y * c2 * c2;
        }
        if (lv) {
            float d1 = closed_v(c, i - 1) ? 1.0f : 0.5f, d2 = closed_v(c, i + 1) ? 1.0f : 0.5f;
            d.y = cf.w + v_wts(c, cx, cy - 1).x + me.x + me.y * d1 * d1 + v_wts(c, cx + 1, cy).y * d2 * d2;
        }
    }
    float2 dv = make_float2(0.0f, 0.0f);
    if (c.P->v_warm && on && inner) {
        const float ratio = c.S->dt / c.S->dt_prev;
        float2 vd = c.vd[i];
        dv = make_float2(vd.x * ratio * (lu ? 1.0f : 0.0f), vd.y * ratio * (lv ? 1.0f : 0.0f));
    }
    c.vd[i] = dv;
    float4 a = make_float4(0, 0, 0, 0), b = make_float4(0, 0, 0, 0);
    float2 dm = make_float2(0.0f, 0.0f);
    const bool write_mg = mg && (on || s_was != 0u);
    if (write_mg && on && inner) {
        couplings(c, cx, cy, a, b);
        dm = make_float2(d.x * (lu ? 1.0f : 0.0f), d.y * (lv ? 1.0f : 0.0f));
    }
    // Neighbours read only the weights of active tiles (unchanged here), so the writes below
    // cannot race with their reads.
    float4 o = c.vw[i];
    o.z = d.x;
    o.w = d.y;
    if (!on) {
        o.x = o.y = 0.0f;
        const float2 z2 = make_float2(0.0f, 0.0f);
        c.vs[i] = make_float4(0, 0, 0, 0);
        c.vr[i] = z2;
        c.vp[i] = z2;
        c.vp2[i] = z2;
        c.vq[i] = z2;
    }
    c.vw[i] = o;
    if (write_mg) {
        if (!on) {
            const float2 z2 = make_float2(0.0f, 0.0f);
            c.vmz[i] = z2;
            c.vaz[i] = z2;
            c.vr2[i] = z2;
            c.vq2[i] = z2;
        }
        c.vma[i] = a;
        c.vmb[i] = b;
        c.vmd[i] = dm;
    }
}

namespace {
SP_DEV bool vflag(const Ctx& c, int l, int slot) {
    return l == 0 ? c.vtiles[slot] != 0u : c.vtile[c.lev[l].w + slot] != 0u;
}
}  // namespace

// vmg_build.comp:187-240: level l of the viscosity multigrid from level l - 1 (P^T A P with
// 2 x 2 aggregates per component), over coarse 16 x 16 tiles, with the tile flags of
// mg_build.comp.
__global__ __launch_bounds__(256) void k_vmg_build(const TankBatch tb, int l) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ uint32_t s_state;
    const Lv F = level(c, l - 1), C = level(c, l);
    if (tid2() == 0) {
        int tfx = (F.w + 15) >> 4, tfy = (F.h + 15) >> 4;
        uint32_t on = 0;
        for (int k = 0; k < 4; k++) {
            int tx = 2 * (int)blockIdx.x + (k & 1), ty = 2 * (int)blockIdx.y + (k >> 1);
            if (tx < tfx && ty < tfy) on |= vflag(c, l - 1, ty * tfx + tx) ? 1u : 0u;
        }
        uint32_t s = C.toff + tile_slot();
        uint32_t was = c.vtile[s];
        c.vtile[s] = on ? 1u : 0u;
        s_state = (on ? 1u : 0u) | (was ? 2u : 0u);
    }
    __syncthreads();
    const uint32_t state = s_state;
    if (state == 0u) return;
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (!inside(C, cx, cy)) return;
    const int o = at(C, cx, cy);
    if ((state & 1u) == 0u) {
        c.vma[o] = make_float4(0, 0, 0, 0);
        c.vmb[o] = make_float4(0, 0, 0, 0);
        c.vmd[o] = make_float2(0.0f, 0.0f);
        return;
    }
    float2 D = make_float2(0.0f, 0.0f);
    float4 A = make_float4(0, 0, 0, 0), B = make_float4(0, 0, 0, 0);
    const int UVX[4] = {0, -1, 0, -1}, UVY[4] = {0, 0, 1, 1};
    for (int k = 0; k < 4; k++) {
        const int ax = k & 1, ay = k >> 1;
        const int x = 2 * cx + ax, y = 2 * cy + ay;
        if (!inside(F, x, y)) continue;
        const int i = at(F, x, y);
        const float4 a = c.vma[i], b = c.vmb[i];
        const float2 dd = c.vmd[i];
        D.x += dd.x;
        D.y += dd.y;
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
            float bm = f4c(b, m);
            if (sl == 0) B.x += bm;
            else if (sl == 1) B.y += bm;
            else if (sl == 2) B.z += bm;
            else B.w += bm;
        }
    }
    c.vma[o] = A;
    c.vmb[o] = B;
    c.vmd[o] = make_float2(gmax_(D.x, 0.0f), gmax_(D.y, 0.0f));
}

// visc_strain.comp:11-17: the strain of the current velocities in the active tiles (the
// first pressure solve changed them since the split).
__global__ __launch_bounds__(256) void k_visc_strain(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (c.vtiles[tile_slot()] == 0u) return;
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    float4 w = c.vw[i];
    c.vs[i] = strain_of_vel(c, i, w.x, w.y);
}

// visc_apply.comp mode 0: r = b - A u* = -G^T W G u*, p = z = r / diag, and the partial
// sums of r.z and r.r.
__global__ __launch_bounds__(256) void k_visc_start(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[4];
    const int slot = tile_slot();
    if (c.vtiles[slot] == 0u) {
        if (tid2() == 0) c.vpart[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    int cx, cy, i;
    double4 acc = make_double4(0, 0, 0, 0);
    if (cell_of_thread(c, cx, cy, i)) {
        if (v_interior(c, cx, cy)) {
            float2 g = gts_of(c, c.vs, i);
            float2 lv = live2(c, i);
            float4 w = c.vw[i];
            float2 r = make_float2(-g.x * lv.x, -g.y * lv.y);
            float2 z = make_float2(r.x / w.z, r.y / w.w);
            c.vr[i] = r;
            c.vp[i] = z;
            acc.x = dot2d(r, z);
            acc.y = dot2d(r, r);
        } else {
            c.vr[i] = make_float2(0.0f, 0.0f);
            c.vp[i] = make_float2(0.0f, 0.0f);
        }
    }
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) c.vpart[slot] = s;
}

namespace {

constexpr int PW = 20;  // a direction on the tile and two rings
constexpr int SW = 18;  // its strain on the tile and one ring

// The direction p (ring 2) and its weighted strain (ring 1) of the tile in LDS
// (visc_iter.comp:55-112). dir(x, y, i) gives p of an interior cell.
struct DirStage {
    float2 sp[PW * PW];
    float4 ss[SW * SW];
};

template <class Dir>
SP_DEV void stage_direction(const Ctx& c, int x0, int y0, DirStage& S, Dir dir) {
    const int t = tid2();
    for (int k = t; k < PW * PW; k += 256) {
        int gx = x0 - 2 + k % PW, gy = y0 - 2 + k / PW;
        float2 p = make_float2(0.0f, 0.0f);
        if (gx >= 0 && gy >= 0 && gx < c.W && gy < c.H && v_interior(c, gx, gy)) p = dir(gx, gy, gy * c.W + gx);
        S.sp[k] = p;
    }
    __syncthreads();
    const int W = c.W;
    for (int k = t; k < SW * SW; k += 256) {
        int gx = x0 - 1 + k % SW, gy = y0 - 1 + k / SW;
        float4 o = make_float4(0, 0, 0, 0);
        if (gx >= 0 && gy >= 0 && gx < c.W && gy < c.H) {
            const int i = gy * W + gx;
            const float4 w = c.vw[i];
            auto P = [&](int ox, int oy) { return S.sp[(gy + oy - y0 + 2) * PW + (gx + ox - x0 + 2)]; };
            if (w.x > 0.0f) {
                float u1 = live_u(c, i + 1) ? P(1, 0).x : 0.0f, u0 = live_u(c, i) ? P(0, 0).x : 0.0f;
                float v1 = live_v(c, i + W) ? P(0, 1).y : 0.0f, v0 = live_v(c, i) ? P(0, 0).y : 0.0f;
                o.x = w.x * (u1 - u0);
                o.y = w.x * (v1 - v0);
            }
            if (w.y > 0.0f) {
                float ua = live_u(c, i) ? P(0, 0).x : 0.0f, ub = live_u(c, i - W) ? P(0, -1).x : 0.0f;
                float va = live_v(c, i) ? P(0, 0).y : 0.0f, vb = live_v(c, i - 1) ? P(-1, 0).y : 0.0f;
                if (closed_u(c, i) && live_u(c, i - W)) ua = -ub;
                if (closed_u(c, i - W) && live_u(c, i)) ub = -ua;
                if (closed_v(c, i) && live_v(c, i - 1)) va = -vb;
                if (closed_v(c, i - 1) && live_v(c, i)) vb = -va;
                o.z = w.y * 0.5f * ((ua - ub) + (va - vb));
            }
        }
        S.ss[k] = o;
    }
    __syncthreads();
}

// G^T s at the faces of tile cell (x, y) from the staged strain (visc_iter.comp:79-92).
SP_DEV float2 gts_staged(const Ctx& c, const DirStage& S, int x0, int y0, int x, int y, int i) {
    const int W = c.W;
    auto s = [&](int ox, int oy) { return S.ss[(y + oy - y0 + 1) * SW + (x + ox - x0 + 1)]; };
    float2 g = make_float2(0.0f, 0.0f);
    if (live_u(c, i)) {
        float c1 = closed_u(c, i - W) ? 1.0f : 0.5f, c2 = closed_u(c, i + W) ? -1.0f : -0.5f;
        g.x = s(-1, 0).x - s(0, 0).x + s(0, 0).z * c1 + s(0, 1).z * c2;
    }
    if (live_v(c, i)) {
        float d1 = closed_v(c, i - 1) ? 1.0f : 0.5f, d2 = closed_v(c, i + 1) ? -1.0f : -0.5f;
        g.y = s(0, -1).y - s(0, 0).y + s(0, 0).z * d1 + s(1, 0).z * d2;
    }
    return g;
}

}  // namespace

// visc_iter.comp with WARM (viscosity_warm): the warm start of the solve. It steps from
// x0 = u* by d = vd, last substep's change, rescaled and masked by k_visc_diag: x = u* + d,
// r = r0 - A d, with the partial sums of r.z and r.r.
__global__ __launch_bounds__(256) void k_visc_warm(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ DirStage S;
    __shared__ double4 lds[4];
    if (v_done(c)) return;
    const int slot = tile_slot();
    if (c.vtiles[slot] == 0u) {
        if (tid2() == 0) c.vpart[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    stage_direction(c, x0, y0, S, [&](int, int, int i) { return c.vd[i]; });
    int cx, cy, i;
    double4 acc = make_double4(0, 0, 0, 0);
    if (cell_of_thread(c, cx, cy, i) && v_interior(c, cx, cy)) {
        float2 g = gts_staged(c, S, x0, y0, cx, cy, i);
        float2 lv = live2(c, i);
        float4 cf = c.coef[i];
        float2 pd = S.sp[(cy - y0 + 2) * PW + (cx - x0 + 2)];
        float2 d = make_float2(pd.x * lv.x, pd.y * lv.y);
        float2 q = make_float2((cf.z * d.x + g.x) * lv.x, (cf.w * d.y + g.y) * lv.y);
        float4 v = c.vel[i];
        v.x += d.x;
        v.y += d.y;
        c.vel[i] = v;
        float2 r0 = c.vr[i];
        float2 r = make_float2(r0.x - q.x, r0.y - q.y);
        c.vr[i] = r;
        float4 w = c.vw[i];
        float2 z = make_float2(r.x / w.z, r.y / w.w);
        acc.x = dot2d(r, z);
        acc.y = dot2d(r, r);
    }
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) c.vpart[slot] = s;
}

// visc_iter.comp: one Jacobi-preconditioned iteration up to its step length: p = z + beta
// p_old (z = r / diag), s = W G p, q = M p + G^T s, and the partial sums of p.q.
__global__ __launch_bounds__(256) void k_visc_iter(const TankBatch tb, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    // the direction alternates between vp and vp2 with the iteration's parity
    const float2* __restrict__ p_old = par ? c.vp2 : c.vp;
    float2* __restrict__ p_new = par ? c.vp : c.vp2;
    __shared__ DirStage S;
    __shared__ double4 lds[4];
    if (v_done(c)) return;
    const int slot = tile_slot();
    if (c.vtiles[slot] == 0u) {
        if (tid2() == 0) c.vpart[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const float beta = (float)c.K->v_beta;
    stage_direction(c, x0, y0, S, [&](int, int, int i) {
        float4 w = c.vw[i];
        float2 r = c.vr[i];
        float2 z = make_float2(r.x / w.z, r.y / w.w);
        if (beta == 0.0f) return z;
        float2 po = p_old[i];
        return make_float2(z.x + beta * po.x, z.y + beta * po.y);
    });
    int cx, cy, i;
    double4 acc = make_double4(0, 0, 0, 0);
    if (cell_of_thread(c, cx, cy, i)) {
        float2 pn = S.sp[(cy - y0 + 2) * PW + (cx - x0 + 2)];
        p_new[i] = pn;
        if (v_interior(c, cx, cy)) {
            float2 g = gts_staged(c, S, x0, y0, cx, cy, i);
            float2 lv = live2(c, i);
            float4 cf = c.coef[i];
            float2 p = make_float2(pn.x * lv.x, pn.y * lv.y);
            float2 q = make_float2((cf.z * p.x + g.x) * lv.x, (cf.w * p.y + g.y) * lv.y);
            c.vq[i] = q;
            acc.z = dot2d(p, q);
        } else {
            c.vq[i] = make_float2(0.0f, 0.0f);
        }
    }
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) c.vpart[slot] = s;
}

// visc_update.comp:11-36: u += alpha p, vd += alpha p, r -= alpha q, and the partial sums of
// r.z and r.r.
__global__ __launch_bounds__(256) void k_visc_update(const TankBatch tb, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ p_dir = par ? c.vp : c.vp2;
    __shared__ double4 lds[4];
    if (v_done(c)) return;
    const int slot = tile_slot();
    if (c.vtiles[slot] == 0u) {
        if (tid2() == 0) c.vpart[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    int cx, cy, i;
    double4 acc = make_double4(0, 0, 0, 0);
    if (cell_of_thread(c, cx, cy, i) && v_interior(c, cx, cy)) {
        const float a = (float)c.K->v_alpha;
        float4 v = c.vel[i];
        float2 p = p_dir[i];
        const bool lu = face_live(v.z), lv = face_live(v.w);
        if (lu) v.x += a * p.x;
        if (lv) v.y += a * p.y;
        c.vel[i] = v;
        float2 vd = c.vd[i];
        vd.x += a * (lu ? p.x : 0.0f);
        vd.y += a * (lv ? p.y : 0.0f);
        c.vd[i] = vd;
        float2 r0 = c.vr[i], q = c.vq[i];
        float2 r = make_float2(r0.x - a * q.x, r0.y - a * q.y);
        c.vr[i] = r;
        float4 w = c.vw[i];
        float2 z = make_float2(r.x / w.z, r.y / w.w);
        acc.x = dot2d(r, z);
        acc.y = dot2d(r, r);
    }
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) c.vpart[slot] = s;
}

// ---- the viscosity multigrid (vmg.glsl) -----------------------------------------------------

namespace {

#define SP_MEM __device__ __forceinline__

// The staged region of a tile: the tile and RH rings of its level (zero outside the level).
template <int NU, int EXTRA>
struct VRegion {
    static constexpr int HALO = 2 * NU;
    static constexpr int RH = HALO + EXTRA;
    static constexpr int RW = 16 + 2 * RH;
    static constexpr int RR = RW * RW;
    static constexpr int PER = (RR + 255) / 256;
    float4 sa[RR];
    float4 sb[RR];
    float2 sd[RR];
    float2 sx[RR];
    float2 sr[RR];

    SP_DEV static int cell_x(int x0, int k) { return x0 - RH + k % RW; }
    SP_DEV static int cell_y(int y0, int k) { return y0 - RH + k / RW; }
    SP_DEV static int tile_slot_of(int tx, int ty) { return (ty + RH) * RW + tx + RH; }
    SP_DEV static bool in_ring(int k, int lo) {
        int lx = k % RW, ly = k / RW;
        return lx >= lo && ly >= lo && lx < RW - lo && ly < RW - lo;
    }

    // Off-diagonal part of A x at slot k (vmg.glsl:69-78); k is never on the outer ring.
    SP_MEM float2 offd(int k) const {
        const float4 a = sa[k], b = sb[k];
        const float2 xl = sx[k - 1], xr = sx[k + 1], xd = sx[k - RW], xu = sx[k + RW];
        float u = a.x * xl.x + sa[k + 1].x * xr.x + a.y * xd.x + sa[k + RW].y * xu.x + b.x * sx[k].y + b.y * xl.y +
                  b.z * xu.y + b.w * sx[k + RW - 1].y;
        float v = a.z * xl.y + sa[k + 1].z * xr.y + a.w * xd.y + sa[k + RW].w * xu.y + b.x * sx[k].x +
                  sb[k + 1].y * xr.x + sb[k - RW].z * xd.x + sb[k + 1 - RW].w * sx[k + 1 - RW].x;
        return make_float2(u, v);
    }

    SP_MEM float2 relax(int k) const {
        const float2 d = sd[k];
        const float2 o = offd(k);
        const float2 s = make_float2(sr[k].x - o.x, sr[k].y - o.y);
        return make_float2(d.x > 0.0f ? s.x / d.x : 0.0f, d.y > 0.0f ? s.y / d.y : 0.0f);
    }

    // One half-sweep of colour `red` over slots [lo, RW - lo): every update reads the
    // values from before it (vmg.glsl:119-134).
    SP_MEM void half_sweep(int x0, int y0, bool red, int lo) {
        const int t = tid2();
        float2 nv[PER];
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int k = t + 256 * j;
            nv[j] = make_float2(0.0f, 0.0f);
            if (k < RR && in_ring(k, lo) && (((cell_x(x0, k) + cell_y(y0, k)) & 1) == 0) == red) nv[j] = relax(k);
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const int k = t + 256 * j;
            if (k < RR && in_ring(k, lo) && (((cell_x(x0, k) + cell_y(y0, k)) & 1) == 0) == red) sx[k] = nv[j];
        }
        __syncthreads();
    }

    // Matrix and right-hand side of the region (vmg.glsl:88-111); rhs(i) of a cell of the
    // level.
    template <class Rhs>
    SP_MEM void load(const Ctx& c, const Lv& L, int x0, int y0, Rhs rhs) {
        for (int k = tid2(); k < RR; k += 256) {
            const int gx = cell_x(x0, k), gy = cell_y(y0, k);
            float4 a = make_float4(0, 0, 0, 0), b = make_float4(0, 0, 0, 0);
            float2 d = make_float2(0.0f, 0.0f), r = make_float2(0.0f, 0.0f);
            if (inside(L, gx, gy)) {
                const int i = at(L, gx, gy);
                a = c.vma[i];
                b = c.vmb[i];
                d = c.vmd[i];
                r = rhs(i);
            }
            sa[k] = a;
            sb[k] = b;
            sd[k] = d;
            sr[k] = r;
        }
    }

    // The way down (vmg.glsl:140-168): pre-smoothing from zero, red then black, the tile's
    // values to vmx and its residual, summed over each 2 x 2 aggregate, to the right-hand
    // side of the next level unless this is the coarsest level of the cycle.
    SP_MEM void way_down(const Ctx& c, const Lv& L, int l, bool coarsest, int x0, int y0, float2* sres) {
        const int t = tid2();
        for (int k = t; k < RR; k += 256) {
            const float2 d = sd[k], r = sr[k];
            const float2 z = make_float2(d.x > 0.0f ? r.x / d.x : 0.0f, d.y > 0.0f ? r.y / d.y : 0.0f);
            sx[k] = (((cell_x(x0, k) + cell_y(y0, k)) & 1) == 0) ? z : make_float2(0.0f, 0.0f);
        }
        __syncthreads();
        for (int h = 1; h < HALO; h++) half_sweep(x0, y0, (h & 1) == 0, h);
        const int tx = threadIdx.x, ty = threadIdx.y;
        const int gx = x0 + tx, gy = y0 + ty;
        const int k = tile_slot_of(tx, ty);
        const float2 x = sx[k], d = sd[k];
        const float2 o = offd(k);
        float2 res;
        res.x = (sr[k].x - d.x * x.x - o.x) * (d.x > 0.0f ? 1.0f : 0.0f);
        res.y = (sr[k].y - d.y * x.y - o.y) * (d.y > 0.0f ? 1.0f : 0.0f);
        if (inside(L, gx, gy)) c.vmx[at(L, gx, gy)] = x;
        if (coarsest) return;
        sres[t] = res;
        __syncthreads();
        const Lv C = level(c, l + 1);
        const int cx = gx >> 1, cy = gy >> 1;
        if (((tx | ty) & 1) == 0 && inside(C, cx, cy)) {
            const float2 r00 = sres[t], r10 = sres[t + 1], r01 = sres[t + 16], r11 = sres[t + 17];
            c.vmr[at(C, cx, cy)] = make_float2(r00.x + r10.x + r01.x + r11.x, r00.y + r10.y + r01.y + r11.y);
        }
    }

    // The start of a way up (vmg_up.comp:277-287): x = vmx plus the coarse-grid correction
    // (none on the coarsest level of the cycle), masked to the unknowns, then NU sweeps
    // black then red.
    SP_MEM void way_up(const Ctx& c, const Lv& L, const Lv& C, bool coarsest, int x0, int y0) {
        for (int k = tid2(); k < RR; k += 256) {
            const int gx = cell_x(x0, k), gy = cell_y(y0, k);
            float2 x = make_float2(0.0f, 0.0f);
            if (inside(L, gx, gy)) {
                x = c.vmx[at(L, gx, gy)];
                if (!coarsest) {
                    float2 zc = c.vmz[at(C, gx >> 1, gy >> 1)];
                    x.x += zc.x;
                    x.y += zc.y;
                }
            }
            const float2 d = sd[k];
            sx[k] = make_float2(x.x * (d.x > 0.0f ? 1.0f : 0.0f), x.y * (d.y > 0.0f ? 1.0f : 0.0f));
        }
        __syncthreads();
        for (int h = 1; h <= HALO; h++) half_sweep(x0, y0, (h & 1) == 0, h);
    }
};

}  // namespace

// vmg_down.comp: the way down of a viscosity V-cycle on level l (level 0 only at the start of
// a solve, from the residual vr; the iterations take theirs inside k_visc_mgu).
template <int NU>
__global__ __launch_bounds__(256) void k_vmg_down(const TankBatch tb, int l, int coarsest) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ VRegion<NU, 0> R;
    __shared__ float2 sres[256];
    if (v_done(c)) return;
    if (!vflag(c, l, tile_slot())) return;
    const Lv L = level(c, l);
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    const float2* rhs = l == 0 ? c.vr : c.vmr;
    R.load(c, L, x0, y0, [&](int i) { return rhs[i]; });
    __syncthreads();
    R.way_down(c, L, l, coarsest != 0, x0, y0, sres);
}

// vmg_up.comp: the way up on a coarse level l, the finished values to vmz.
template <int NU>
__global__ __launch_bounds__(256) void k_vmg_up(const TankBatch tb, int l, int coarsest) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ VRegion<NU, 0> R;
    if (v_done(c)) return;
    if (!vflag(c, l, tile_slot())) return;
    const Lv L = level(c, l);
    const Lv C = level(c, l + 1 < c.n_levels ? l + 1 : c.n_levels - 1);
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    R.load(c, L, x0, y0, [&](int i) { return c.vmr[i]; });
    __syncthreads();
    R.way_up(c, L, C, coarsest != 0, x0, y0);
    const int gx = x0 + threadIdx.x, gy = y0 + threadIdx.y;
    if (inside(L, gx, gy)) c.vmz[at(L, gx, gy)] = R.sx[VRegion<NU, 0>::tile_slot_of(threadIdx.x, threadIdx.y)];
}

// visc_mgu.comp:28-65: p = z + beta p, q = A z + beta q, u += alpha p, r -= alpha q, the
// partial sums of r.r and p.q, then the way down on level 0 from the new residual (which the
// group recomputes for its rings from r_in and q_in, so it need not wait for its neighbours).
template <int NU>
__global__ __launch_bounds__(256) void k_visc_mgu(const TankBatch tb, int par) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    // residual and A p alternate between (vr, vq) and (vr2, vq2) with the iteration's parity
    const float2* __restrict__ r_in = par ? c.vr2 : c.vr;
    const float2* __restrict__ q_in = par ? c.vq2 : c.vq;
    float2* __restrict__ r_out = par ? c.vr : c.vr2;
    float2* __restrict__ q_out = par ? c.vq : c.vq2;
    __shared__ VRegion<NU, 0> R;
    __shared__ double4 lds[4];
    __shared__ float2 sres[256];
    if (v_done(c)) return;
    const int slot = tile_slot();
    if (c.vtiles[slot] == 0u) {
        if (tid2() == 0) c.vpart[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    const float a = (float)c.K->v_alpha, b = (float)c.K->v_beta;
    auto new_q = [&](int i) {
        float2 s = c.vaz[i];
        if (b == 0.0f) return s;
        float2 q = q_in[i];
        return make_float2(s.x + b * q.x, s.y + b * q.y);
    };
    int cx, cy, i;
    double4 acc = make_double4(0, 0, 0, 0);
    if (cell_of_thread(c, cx, cy, i)) {
        float2 z = c.vmz[i];
        float2 p = z;
        if (b != 0.0f) {
            float2 po = c.vp[i];
            p = make_float2(z.x + b * po.x, z.y + b * po.y);
        }
        float2 q = new_q(i);
        float2 r0 = r_in[i];
        float2 r = make_float2(r0.x - a * q.x, r0.y - a * q.y);
        if (v_interior(c, cx, cy)) {
            float4 v = c.vel[i];
            if (face_live(v.z)) v.x += a * p.x;
            if (face_live(v.w)) v.y += a * p.y;
            c.vel[i] = v;
            float2 vd = c.vd[i];
            vd.x += a * p.x;
            vd.y += a * p.y;
            c.vd[i] = vd;
        }
        c.vp[i] = p;
        q_out[i] = q;
        r_out[i] = r;
        acc.y = dot2d(r, r);
        acc.z = dot2d(p, q);
    }
    const Lv L = level(c, 0);
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    R.load(c, L, x0, y0, [&](int j) {
        float2 q = new_q(j);
        float2 r0 = r_in[j];
        return make_float2(r0.x - a * q.x, r0.y - a * q.y);
    });
    __syncthreads();
    R.way_down(c, L, 0, false, x0, y0, sres);
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) c.vpart[slot] = s;
}

// visc_mgz.comp:85-120: the way up on level 0 on one ring more than the tile, so that A z of
// the tile follows in LDS; writes z (vmz) and A z (vaz) with the partial sums of r.z, z.A z
// and z.q (q = A p of this iteration).
template <int NU>
__global__ __launch_bounds__(256) void k_visc_mgz(const TankBatch tb, int sel) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const float2* __restrict__ r = sel ? c.vr2 : c.vr;
    const float2* __restrict__ q = sel ? c.vq2 : c.vq;
    __shared__ VRegion<NU, 1> R;
    __shared__ double4 lds[4];
    if (v_done(c)) return;
    const int slot = tile_slot();
    if (c.vtiles[slot] == 0u) {
        if (tid2() == 0) c.vpart[slot] = make_double4(0, 0, 0, 0);
        return;
    }
    const Lv L = level(c, 0), C = level(c, 1);
    const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
    R.load(c, L, x0, y0, [&](int i) { return r[i]; });
    __syncthreads();
    R.way_up(c, L, C, false, x0, y0);
    const int gx = x0 + threadIdx.x, gy = y0 + threadIdx.y;
    double4 acc = make_double4(0, 0, 0, 0);
    if (inside(L, gx, gy)) {
        const int i = at(L, gx, gy);
        const int k = VRegion<NU, 1>::tile_slot_of(threadIdx.x, threadIdx.y);
        const float2 z = R.sx[k], d = R.sd[k];
        const float2 o = R.offd(k);
        const float2 az = make_float2(d.x * z.x + o.x, d.y * z.y + o.y);
        c.vmz[i] = z;
        c.vaz[i] = az;
        acc.x = dot2d(R.sr[k], z);
        acc.y = dot2d(z, az);
        acc.z = dot2d(z, q[i]);
    }
    double4 s = group_sum4(acc, lds);
    if (tid2() == 0) c.vpart[slot] = s;
}

// ---- the scalar steps (reduce_tail.glsl modes 4-7, 10-13), FP64 -----------------------------

enum { VRED_START = 4, VRED_ALPHA = 5, VRED_BETA = 6, VRED_WARM = 10, VRED_MG_START = 11, VRED_MG_UPDATE = 12,
       VRED_MG_ITER = 13 };

// After the last kernel of an iteration: stop here if this tank records no more.
SP_DEV void v_count_sweep(sp_solve* K) {
    K->v_sweeps += 1u;
    if (K->cap_v != 0u && K->v_sweeps >= K->cap_v && K->v_done == 0u) K->v_halt = 1u;
}

__global__ __launch_bounds__(REDUCE_NT) void k_visc_reduce(const TankBatch tb, int mode, int nblocks) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    __shared__ double4 lds[REDUCE_NT / 64];
    sp_solve* K = c.K;
    if (mode != VRED_START && v_done(c)) return;
    const double tol2 = (double)c.P->tol[2] * (double)c.P->tol[2];
    const double4 a = total_of<REDUCE_NT>(c.vpart, 0, nblocks, lds);
    if (tid2() != 0) return;
    switch (mode) {
    case VRED_START:
        K->v_rz = a.x;
        K->v_r0 = a.y;
        K->v_rr = a.y;
        K->v_beta = 0.0;
        K->v_alpha = 0.0;
        K->v_iters = 0u;
        K->v_sweeps = 0u;
        K->v_halt = 0u;
        K->v_done = (a.y <= 1e-30 || a.x <= 0.0) ? 1u : 0u;
        break;
    case VRED_WARM:
        K->v_rz = a.x;
        K->v_rr = a.y;
        K->v_beta = 0.0;
        K->v_done = (a.y <= tol2 * K->v_r0 || a.x <= 0.0) ? 1u : 0u;
        break;
    case VRED_ALPHA:
        K->v_alpha = (K->v_done == 0u && a.z > 0.0) ? K->v_rz / a.z : 0.0;
        break;
    case VRED_BETA: {
        const uint32_t was = K->v_done;
        K->v_iters += 1u - was;
        const uint32_t done = was | (a.y <= tol2 * K->v_r0 ? 1u : 0u);
        K->v_beta = (done == 0u && K->v_rz > 0.0) ? a.x / K->v_rz : 0.0;
        K->v_rz = a.x;
        if (!was) K->v_rr = a.y;
        K->v_done = done;
        v_count_sweep(K);
        break;
    }
    case VRED_MG_START:
        K->v_rz = a.x;
        K->v_beta = 0.0;
        K->v_pq = a.y;
        K->v_alpha = a.y > 0.0 ? a.x / a.y : 0.0;
        if (a.x <= 0.0 || a.y <= 0.0) K->v_done = 1u;
        break;
    case VRED_MG_UPDATE: {
        const uint32_t was = K->v_done;
        K->v_iters += 1u - was;
        if (!was) K->v_rr = a.y;
        K->v_pq = a.z;
        K->v_done = was | (a.y <= tol2 * K->v_r0 ? 1u : 0u);
        break;
    }
    case VRED_MG_ITER: {
        const double beta = (K->v_done == 0u && K->v_rz > 0.0) ? a.x / K->v_rz : 0.0;
        const double pq = a.y + 2.0 * beta * a.z + beta * beta * K->v_pq;
        K->v_beta = beta;
        K->v_alpha = (K->v_done == 0u && pq > 0.0) ? a.x / pq : 0.0;
        K->v_rz = a.x;
        v_count_sweep(K);
        break;
    }
    default: break;
    }
}

// Statistics at the end of a viscosity solve (reduce_tail.glsl:186-192, mode 7).
__global__ void k_visc_end(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    sp_solve* K = c.K;
    if (K->v_iters > K->v_it_max) K->v_it_max = K->v_iters;
    if (K->v_done == 0u) K->v_unconv += 1u;
    float rel = K->v_r0 > 0.0 ? (float)sqrt(K->v_rr / K->v_r0) : 0.0f;
    if (rel > K->v_res_max) K->v_res_max = rel;
    K->v_iters_last = K->v_iters;
}

// ---- the second pressure solve's right-hand side and projection -----------------------------

// rhs_visc.comp:143-168: b = -div(u) + A P1 (P1 in prhs.z, kept by the first projection),
// and the start P1 + last substep's P2 rescaled (viscous_warm). Neighbours rewrite prhs with
// the same P1 in z, so reading their z while they write is safe.
__global__ __launch_bounds__(256) void k_rhs_visc(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    if (!cell_of_thread(c, cx, cy, i)) return;
    const float D = c.mgd[i];
    if (D <= 0.0f) {
        c.prhs[i].z = 0.0f;
        return;
    }
    const int W = c.W;
    const float4 w = c.mga[i];
    const float p1 = c.prhs[i].z;
    float ap = D * p1;
    if (c.mgd[i - 1] > 0.0f) ap -= w.x * c.prhs[i - 1].z;
    if (c.mgd[i + 1] > 0.0f) ap -= w.y * c.prhs[i + 1].z;
    if (c.mgd[i - W] > 0.0f) ap -= w.z * c.prhs[i - W].z;
    if (c.mgd[i + W] > 0.0f) ap -= w.w * c.prhs[i + W].z;
    const float aL = c.solid[i].x, aB = c.solid[i].y, aR = c.solid[i + 1].x, aT = c.solid[i + W].y;
    const float4 vme = c.vel[i];
    const float div = aR * c.vel[i + 1].x - aL * vme.x + aT * c.vel[i + W].y - aB * vme.y;
    const float ry = c.prhs[i].y;
    c.prhs[i] = make_float4(ap - div, ry, p1, 0.0f);
    if (c.P->v_p2warm) c.pres[i].x = p1 + c.p2keep[i] * (c.S->dt / c.S->dt_prev);
}

// project_visc.comp:187-210: subtract the gradient of P2 = P - P1 (zero in air), keep P2 for
// the next substep, and report the fastest liquid face again.
__global__ __launch_bounds__(256) void k_project_visc(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    int cx, cy, i;
    const bool in = cell_of_thread(c, cx, cy, i);
    float m = 0.0f;
    if (in && cx >= 1 && cy >= 1) {
        auto p2 = [&](int j) { return c.mgd[j] > 0.0f ? c.pres[j].x - c.prhs[j].z : 0.0f; };
        const float pc = p2(i);
        c.p2keep[i] = pc;
        const float4 cf = c.coef[i];
        float4 me = c.vel[i];
        const bool lu = face_live(me.z) && cf.x != 0.0f, lv = face_live(me.w) && cf.y != 0.0f;
        if (lu || lv) {
            if (lu) me.x -= cf.x * (pc - p2(i - 1));
            if (lv) me.y -= cf.y * (pc - p2(i - c.W));
            c.vel[i] = me;
        }
        if (lu) m = fabsf(me.x);
        if (lv) m = fmaxf_(m, fabsf(me.y));
    }
    __shared__ float s_max[4];
    group_max_to<256>(&c.S->umax_grid, m, s_max);
}
