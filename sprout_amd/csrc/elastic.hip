// elastic.hip — the force of the viscoelastic stress on the liquid (elastic.comp): the
// particles' stress, scattered to cell centres by volume in grid_prep (so it falls to zero
// across a free surface), and its divergence added to the live faces before viscosity and
// pressure, explicitly:
//   u += dt / rho_f (d tau_xx / dx + d tau_xy / dy),  v += dt / rho_f (d tau_xy / dx + d tau_yy / dy)
// with tau_xy averaged to the cell corners. One work-item per cell on 16 x 16 tiles; a cell
// reads its own and its left, lower and lower-left neighbours' stress (the corner values),
// and the row above / the column to the right for the far corners.
#include "dev.h"

namespace {

// tau_xy at the lower-left corner of cell i (elastic.comp:15-18).
SP_DEV float txy_node(const Ctx& c, int i) {
    const int W = c.W;
    return 0.25f * (((c.estr[i].z + c.estr[i - 1].z) + c.estr[i - W].z) + c.estr[i - W - 1].z);
}

}  // namespace

// elastic.comp:20-38: cells 2 .. grid - 2 in each direction.
__global__ __launch_bounds__(256) void k_elastic(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (cx >= c.W - 1 || cy >= c.H - 1 || cx < 2 || cy < 2) return;
    const int W = c.W;
    const int i = cy * W + cx;
    float4 q = c.vel[i];
    const float4 a = c.cell[i], l = c.cell[i - 1], b = c.cell[i - W];
    const float dt = c.S->dt;
    if (face_live(q.z)) {
        const float rho = (a.x * a.y + l.x * l.y) / fmaxf_(a.x + l.x, 1e-6f);
        const float f = (c.estr[i].x - c.estr[i - 1].x) + (txy_node(c, i + W) - txy_node(c, i));
        q.x += dt * f / fmaxf_(rho, 0.05f);
    }
    if (face_live(q.w)) {
        const float rho = (a.x * a.y + b.x * b.y) / fmaxf_(a.x + b.x, 1e-6f);
        const float f = (c.estr[i].y - c.estr[i - W].y) + (txy_node(c, i + 1) - txy_node(c, i));
        q.y += dt * f / fmaxf_(rho, 0.05f);
    }
    c.vel[i] = q;
}
