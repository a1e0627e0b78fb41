// rigid.hip — rigid bodies of frozen or too viscous liquid (rigid.comp): after the
// projection, the velocity inside each connected rigid region is replaced by the rigid motion
// with its linear and angular momentum, with contact against the solids it rests on. One
// work-item per cell on 16 x 16 tiles, five passes (modes):
//   0  label: a rigid cell is its own body;
//   1  union: join 8-connected rigid neighbours (union-find by atomicMin; a body's root is its
//      smallest cell index, whatever the order of the joins);
//   2  compress: every cell points at its root; each root claims a slot of the body table;
//   3  sum: mass, moments, momentum, angular momentum, inertia and contacts per body, as
//      64-bit fixed-point integer sums (the same total in any order);
//   4  apply: faces between two cells of one body take the body's motion.
//
// The reference adds each 64-bit quantum as two 32-bit atomics with a carry; here each is one
// 64-bit atomic of the same integer, so the sums are the same words.
#include "dev.h"

namespace {

// Rigid: liquid that is mostly frozen, too viscous to flow, or, after the solve, a
// non-Newtonian cell that does not flow: its stress mu g stays below its yield stress (an
// unyielded plug), or its viscosity at its shear rate reaches the cap (jammed shear-thickening
// liquid) (rigid.comp:34-47).
SP_DEV bool rigid_cell(const Ctx& c, int cx, int cy, int i) {
    const float4 me = c.cell[i];
    if (me.w <= 0.0f) return false;
    if ((c.P->caps & SP_CAP_THERMAL) && c.therm[i].w > SP_RIGID_SHARE) return true;
    const float4 q = c.rh[i];
    if (q.w > 0.5f) return true;
    if (!(c.P->caps & SP_CAP_RHEO) || q.x <= 0.0f || cx < 1 || cy < 1 || cx >= c.W - 1 || cy >= c.H - 1) return false;
    const float g = shear_rate(c, i, false);
    float F, tau;
    nn_eval(c, i, q, g, F, tau);
    if (tau > 0.0f && me.z * g < tau) return true;
    return c.mats[(int)q.y].rheo == 3 && me.z * F >= me.y * expf(c.P->ln_rigid);
}

// The label of cell i as memory holds it now (read past the non-coherent cache: other groups
// lower labels during the pass).
SP_DEV int label_of(const Ctx& c, int i) { return __hip_atomic_load(&c.rl[i].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// Labels only ever fall (to a smaller cell index), so the walk ends at a root; the guard
// bounds it as the reference's does (rigid.comp:63-70).
SP_DEV int find_root(const Ctx& c, int i) {
    int l = label_of(c, i);
    for (int guard = 0; guard < (1 << 20) && l != i; guard++) {
        i = l;
        l = label_of(c, i);
    }
    return i;
}

// rigid.comp:72-88.
SP_DEV void unite(const Ctx& c, int a, int b) {
    bool done = false;
    for (int guard = 0; guard < 4096 && !done; guard++) {
        a = find_root(c, a);
        b = find_root(c, b);
        if (a < b) {
            const int old = atomicMin(&c.rl[b].x, a);
            done = old == b;
            b = old;
        } else if (b < a) {
            const int old = atomicMin(&c.rl[a].x, b);
            done = old == a;
            a = old;
        } else {
            done = true;
        }
    }
}

// Adds round(x) to the 64-bit sum at words w, w + 1 (add64, rigid.comp:92-99).
SP_DEV void add64(const Ctx& c, uint32_t w, float x) {
    const float r = floorf(x + 0.5f);
    atomicAdd(reinterpret_cast<unsigned long long*>(c.rbody + w), (unsigned long long)(long long)r);
}

}  // namespace

// rheo.comp: the non-Newtonian viscosity of each liquid cell from the shear rate of the
// velocity before the solve (lagged by one solve), mu_eff = mu F + tau / g up to the rigid
// cap; a skin cell holds (stiff) until the motion across it breaks it (rheo.comp:31-55).
__global__ __launch_bounds__(256) void k_rheo(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (cx < 1 || cy < 1 || cx >= c.W - 1 || cy >= c.H - 1) return;
    const int W = c.W;
    const int i = cy * W + cx;
    float4 me = c.cell[i];
    float4 q = c.rh[i];
    if (me.w <= 0.0f || q.x <= 0.0f) return;
    const float g = shear_rate(c, i, true);
    float F, tau;
    nn_eval(c, i, q, g, F, tau);
    const float eff = me.z * F + tau / fmaxf_(g, 1e-6f);
    const float cap = me.y * expf(c.P->ln_rigid);
    const float skin = q.x * c.mats[(int)q.y].temp[3];
    if (skin > 0.0f && me.y * g * g < skin &&
        (c.cell[i - 1].w <= 0.0f || c.cell[i + 1].w <= 0.0f || c.cell[i - W].w <= 0.0f || c.cell[i + W].w <= 0.0f)) {
        q.w = 1.0f;
        c.rh[i] = q;
    }
    me.z = fminf_(eff, cap);
    c.cell[i] = me;
}

// rigid.comp:101-196.
__global__ __launch_bounds__(256) void k_rigid(const TankBatch tb, int mode) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (cx >= c.W || cy >= c.H) return;
    const int W = c.W;
    const int i = cy * W + cx;
    if (mode == 0) {
        c.rl[i] = make_int2(rigid_cell(c, cx, cy, i) ? i : -1, -1);
        if (i == 0) c.rbody[0] = 0u;
        return;
    }
    const int l = c.rl[i].x;
    if (mode == 1) {
        if (l < 0) return;
        // four of the eight neighbours, each pair once
        if (cx > 0 && label_of(c, i - 1) >= 0) unite(c, i, i - 1);
        if (cy > 0 && label_of(c, i - W) >= 0) unite(c, i, i - W);
        if (cy > 0 && cx > 0 && label_of(c, i - W - 1) >= 0) unite(c, i, i - W - 1);
        if (cy > 0 && cx < W - 1 && label_of(c, i - W + 1) >= 0) unite(c, i, i - W + 1);
        return;
    }
    if (mode == 2) {
        if (l < 0) return;
        const int r = find_root(c, i);
        c.rl[i].x = r;
        if (r == i) {
            const uint32_t b = atomicAdd(&c.rbody[0], 1u);
            if (b < (uint32_t)SP_RIGID_BODIES) {
                for (uint32_t k = 0; k < (uint32_t)SP_BODY_WORDS; k++) c.rbody[16u + (uint32_t)SP_BODY_WORDS * b + k] = 0u;
                c.rl[i].y = (int)b;
            }
        }
        return;
    }
    if (mode == 3) {
        if (l < 0) return;
        const int b = c.rl[l].y;
        if (b < 0) return;
        const float4 me = c.cell[i];
        const float m = me.y * fminf_(me.x, 1.0f) * SP_RIGID_FIX;
        const float dx = (float)(cx - l % W), dy = (float)(cy - l / W);
        const int last = c.N - 1;
        const float4 q = c.vel[i];
        const float ux = 0.5f * (q.x + c.vel[i + 1 < last ? i + 1 : last].x);
        const float uy = 0.5f * (q.y + c.vel[i + W < last ? i + W : last].y);
        const uint32_t o = 16u + (uint32_t)SP_BODY_WORDS * (uint32_t)b;
        add64(c, o, m);
        add64(c, o + 2u, m * dx);
        add64(c, o + 4u, m * dy);
        add64(c, o + 6u, m * ux);
        add64(c, o + 8u, m * uy);
        add64(c, o + 10u, m * (dx * uy - dy * ux));
        add64(c, o + 12u, m * (dx * dx + dy * dy + 1.0f / 6.0f));
        // a contact: the cell's centre within a cell of a solid surface
        if (c.P->rigid_friction >= 0.0f) {
            float gx, gy;
            const float dist = sample_sdf(c, (float)cx + 0.5f, (float)cy + 0.5f, gx, gy);
            if (dist < 1.0f && gx * gx + gy * gy > 1e-8f) {
                const float gl = len2_(gx, gy);
                const float nx = gx / gl, ny = gy / gl;
                const float F = SP_RIGID_FIX;
                add64(c, o + 14u, F);
                add64(c, o + 16u, F * nx);
                add64(c, o + 18u, F * ny);
                add64(c, o + 20u, F * dx);
                add64(c, o + 22u, F * dy);
                add64(c, o + 24u, F * dx * dx);
                add64(c, o + 26u, F * dy * dy);
                add64(c, o + 28u, F * dx * dy);
            }
        }
        return;
    }
    // mode 4: faces between two cells of one body move with it and take no density correction,
    // except next to a clump (a cell over rest occupancy by more than the dead band)
    float4 q = c.vel[i];
    bool changed = false;
    float3 u;
    const float clump = SP_REST_PHI + fmaxf_(c.P->corr[1], 0.05f);
    const bool cl = c.cell[i].x > clump;
    if (l >= 0 && cx > 0 && c.rl[i - 1].x == l && face_live(q.z) && c.rl[l].y >= 0 &&
        body_motion(c, (uint32_t)c.rl[l].y, l, (float)cx, (float)cy + 0.5f, u)) {
        q.x = u.x;
        if (!cl && c.cell[i - 1].x <= clump) q.z = 0.0f;
        changed = true;
    }
    if (l >= 0 && cy > 0 && c.rl[i - W].x == l && face_live(q.w) && c.rl[l].y >= 0 &&
        body_motion(c, (uint32_t)c.rl[l].y, l, (float)cx + 0.5f, (float)cy, u)) {
        q.y = u.y;
        if (!cl && c.cell[i - W].x <= clump) q.w = 0.0f;
        changed = true;
    }
    if (changed) c.vel[i] = q;
}
