// magnet.hip — magnetism (magnet.comp): the magnetic energy density of a magnetisable liquid
// in the analytic field of the placed magnets (point dipoles in the plane, no field solve),
// and its force on the faces. Mode 0 stores per cell
//     psi = -mu0 Ms^2 / chi_e (x - ln(1 + x)),  x = chi_e H / Ms,  chi_e = chi / (1 + N chi),
// the energy density of a linear liquid that saturates, demagnetised as a sphere (N = 1/3,
// blended towards the surface's own N by `demag`); mode 1 adds -dt s (psi_c - psi_n) / rho
// to each live face before viscosity and pressure, s the face's magnetisable share, rho the
// face density of the projection: inside a body of liquid that is a discrete gradient the
// pressure takes up, so only surfaces and interfaces feel a net force. One work-item per
// cell on 16 x 16 tiles; mode 1 reads its own and its left and lower neighbours' psi.
#include "dev.h"

namespace {

// The field at x (cells) of the placed magnets, B0 on the axis at the rim and constant
// inside it (common.glsl:390-401).
SP_DEV float2 magnet_field(const Ctx& c, float x, float y) {
    const sp_gparams& P = *c.P;
    float Bx = 0.0f, By = 0.0f;
    const int n = (int)P.mag[0];
    for (int k = 0; k < n; k++) {
        const float mx = P.magnet_dir[k][0], my = P.magnet_dir[k][1];
        const float dx = x - P.magnet[k][0], dy = y - P.magnet[k][1];
        const float dl = len2_(dx, dy);
        const float r = fmaxf_(dl, P.magnet[k][2]);
        float ex = mx, ey = my;
        if (dl > 1e-4f) {
            ex = dx / dl;
            ey = dy / dl;
        }
        const float k3 = P.magnet[k][3] * powf(P.magnet[k][2] / r, 3.0f) * 0.5f;
        const float me = 3.0f * (mx * ex + my * ey);
        Bx += k3 * (me * ex - mx);
        By += k3 * (me * ey - my);
    }
    return make_float2(Bx, By);
}

// magnet.comp:63-67.
SP_DEV float mag_face_force(const Ctx& c, int i, int n, float rho) {
    const bool li = c.cell[i].w > 0.0f, ln = c.cell[n].w > 0.0f;
    const float2 a = c.magc[i], b = c.magc[n];
    const float s = (li == ln ? 0.5f * (a.x + b.x) : li ? a.x : b.x) / c.P->mag[1];
    return -c.S->dt * s * (a.y - b.y) / fmaxf_(rho, 0.05f);
}

}  // namespace

// magnet.comp:71-105.
__global__ __launch_bounds__(256) void k_magnet(const TankBatch tb, int mode) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (cx >= c.W || cy >= c.H) return;
    const sp_gparams& P = *c.P;
    const int W = c.W;
    const int i = cy * W + cx;
    if (mode == 1) {
        if (cx < 1 || cy < 1) return;
        float4 me = c.vel[i];
        const float4 cf = c.coef[i];
        if (face_live(me.z) && cf.x != 0.0f) me.x += mag_face_force(c, i, i - 1, cf.z);
        if (face_live(me.w) && cf.y != 0.0f) me.y += mag_face_force(c, i, i - W, cf.w);
        c.vel[i] = me;
        return;
    }
    const float2 B = magnet_field(c, (float)cx + 0.5f, (float)cy + 0.5f);
    const float Bl = len2_(B.x, B.y);
    const float H = Bl / SP_MU0;
    float psi = 0.0f;
    const float chi = P.mag[1];
    if (H > 0.0f && chi > 0.0f) {
        float N = 1.0f / 3.0f;
        if (P.demag > 0.0f) {
            float gx = 0.0f, gy = 0.0f;
            if (cx > 0 && cy > 0 && cx < W - 1 && cy < c.H - 1) {
                gx = 0.5f * (c.surf[i + 1].x - c.surf[i - 1].x);
                gy = 0.5f * (c.surf[i + W].x - c.surf[i - W].x);
            }
            const float gl = len2_(gx, gy);
            const float nh = gl > 1e-6f ? (gx / gl) * (B.x / Bl) + (gy / gl) * (B.y / Bl) : 0.0f;
            N = mixf_(1.0f / 3.0f, nh * nh, clampf_(gl / 0.25f, 0.0f, 1.0f) * P.demag);
        }
        const float ce = chi / (1.0f + N * chi);
        const float Ms = P.mag[2] / SP_MU0;
        const float xs = ce * H / Ms;
        // x - ln(1 + x), by its series where float cancellation would eat it
        const float f = xs < 0.02f ? xs * xs * (0.5f - xs * (1.0f / 3.0f - 0.25f * xs)) : xs - logf(1.0f + xs);
        psi = -SP_MU0 * Ms * Ms / ce * f * P.mag[3];
    }
    c.magc[i].y = psi;
}
