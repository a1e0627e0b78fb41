// settle.hip — settling of suspensions (settle.comp): the grains of a concentrate (mud,
// clay) fall through its diluent at the terminal velocity of a grain, as a flux of
// concentration between vertically adjacent cells of the family,
//     F = v_t s_above (1 - s_below) min(occupancy above, occupancy below),
// the hindered (Kynch) flux of a uniform suspension, limited so that a cell never gives more
// concentrate than it holds nor takes more than it has room for. The pass stores per cell
// and family the gain over the room (1 - s) and the loss over the holding (s); the particle
// pass shares them out with its scatter weights, so the concentrate is conserved and every
// concentration stays in [0, 1]. One work-item per cell on 16 x 16 tiles.
#include "dev.h"

namespace {

// The family's concentration in cell i (family 0 in comp.z, the others in mixb).
SP_DEV float fam_conc(const Ctx& c, int i, int f) {
    return f == 0 ? c.comp[i].z : c.mixb[(size_t)i * (c.P->n_fam - 1) + f - 1];
}

// Flux (cell volumes / s) of family f's concentrate from cell `up` into cell `dn` below it
// (settle.comp:21-30).
SP_DEV float settle_flux(const Ctx& c, int up, int dn, int f, float vt, float dt) {
    const int nf = c.P->n_fam;
    const float ou = c.famocc[(size_t)up * nf + f], od = c.famocc[(size_t)dn * nf + f];
    if (ou <= 1e-3f || od <= 1e-3f) return 0.0f;
    const float su = fam_conc(c, up, f), sd = fam_conc(c, dn, f);
    if (su <= 0.0f || sd < 0.0f) return 0.0f;
    const float F = vt * su * (1.0f - sd) * fminf_(ou, od);
    return fminf_(F, fminf_(su * ou, (1.0f - sd) * od) / dt);
}

}  // namespace

// settle.comp:32-53.
__global__ __launch_bounds__(256) void k_settle(const TankBatch tb) {
    const Ctx c = ctx_of(tb.t[blockIdx.z]);
    const int cx = blockIdx.x * 16 + threadIdx.x, cy = blockIdx.y * 16 + threadIdx.y;
    if (cx >= c.W || cy >= c.H) return;
    const sp_gparams& P = *c.P;
    const int W = c.W;
    const int i = cy * W + cx;
    const int nf = P.n_fam;
    const float dt = c.S->dt;
    const bool th_on = (P.caps & SP_CAP_THERMAL) != 0u;
    for (int f = 0; f < nf; f++) {
        const int row = P.fams[f][0], dl = P.fams[f][1];
        const float d = c.mats[row].grain;
        const float occ = c.famocc[(size_t)i * nf + f];
        float2 ds = make_float2(0.0f, 0.0f);
        if (d > 0.0f && occ > 1e-3f && cy > 0 && cy < c.H - 1) {
            const float drho = c.mats[row].rho - c.mats[dl].rho;
            const float T = th_on ? c.therm[i].x : P.heat_x;
            const float vt = terminal(c, drho, c.mats[dl].rho, expf(row_lnu(c, dl, 1.0f, T, th_on)), d);
            const float gain = drho > 0.0f ? settle_flux(c, i + W, i, f, vt, dt) : settle_flux(c, i - W, i, f, vt, dt);
            const float loss = drho > 0.0f ? settle_flux(c, i, i - W, f, vt, dt) : settle_flux(c, i, i + W, f, vt, dt);
            const float s = fam_conc(c, i, f);
            ds = make_float2(dt * (gain / fmaxf_(occ * (1.0f - s), 1e-6f)), dt * (loss / fmaxf_(occ * s, 1e-6f)));
        }
        c.dsettle[(size_t)i * nf + f] = ds;
    }
}
