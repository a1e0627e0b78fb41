// params.cpp — host-side derivations of the fluid step: default parameters and the
// built-in materials this build carries, the tank's accumulator layout, and the grid-unit
// parameter block and material table the kernels read. No device code; these run on the
// CPU in the C-ABI calls that change them.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "host.h"

namespace sp {

static const float WATER_DENSITY = 998.0f;  // params.rs:808
static const float T_REF = 20.0f;           // material.rs:27
static const float WATER_HEAT = WATER_DENSITY * 4182.0f;  // params.rs:1315
static const float STEFAN_BOLTZMANN = 5.670374e-8f;       // params.rs:1317
static const float NO_MELT = -1.0e9f;                      // mod.rs:76

// MatProps::default (material.rs:344-420), the fields this build reads.
static sp_material default_material() {
    sp_material m;
    memset(&m, 0, sizeof(m));
    m.group = 0;
    m.diluent = -1;
    m.density = 998.0f;
    m.density_solid = 0.0f;
    m.viscosity = 1.0e-6f;
    m.temp_model = 0;
    m.temp_b = 0.0f;
    m.temp_c = 0.0f;
    m.temp_ref = T_REF;
    m.tension = 0.072f;
    m.contact_angle = 90.0f;
    m.mix_rate = 0.0f;
    m.expansion = 4.0e-4f;
    m.melt_lo = -1.0e9f;
    m.melt_hi = -1.0e9f;
    m.visc_exponent = 1.0f;
    m.load = 1.0f;
    m.spawn_temp = T_REF;
    m.latent = 0.0f;
    m.heat_capacity = 4182.0f;
    m.conductivity = 0.6f;  // material.rs:365-378
    m.emissivity = 0.9f;
    m.boil = 1.0e9f;
    m.latent_vap = 0.0f;
    m.freeze_shift = 0.0f;
    m.rheology = 0;  // material.rs:351-355, 384, 405
    m.yield_stress = 0.0f;
    m.flow_index = 1.0f;
    m.shear_time = 1.0f;
    m.visc_ratio = 1.0f;
    m.skin = 0.0f;
    m.friction = 0.0f;
    m.shear_modulus = 0.0f;  // material.rs:380-381
    m.relax_time = 0.0f;
    m.grain = 0.0f;  // material.rs:363
    m.susceptibility = 0.0f;  // material.rs:382-383
    m.saturation = 0.05f;
    m.volatility = 0.0f;  // material.rs:404, 412
    m.fume = 0.0f;
    m.react_with = -1;  // material.rs:403-409
    m.react_into = -1;
    m.react_rate = 0.0f;
    m.react_heat = 0.0f;
    m.self_heat = 0.0f;
    m.residue = -1;  // material.rs:376-377
    m.residue_share = 0.0f;
    return m;
}

void default_params(sp_sim_params* p) {
    memset(p, 0, sizeof(*p));
    // SimParams::default (params.rs:386-460).
    p->cell_size = 0.01f;
    p->gravity = 9.81f;
    p->ambient_temp = T_REF;
    p->rigid_viscosity = 1.0f;
    // default_pair_tension (material.rs:489-505).
    float t[4][4] = {{0.0f, 0.025f, 0.375f, 0.0f}, {0.025f, 0.0f, 0.36f, 0.0f}, {0.375f, 0.36f, 0.0f, 0.0f}, {0.0f, 0.0f, 0.0f, 0.0f}};
    memcpy(p->pair_tension, t, sizeof(t));
    p->droplet_cells = 1.0f;
    p->cfl = 4.0f;
    p->advection_order = 2;
    p->dt_max = 1.0f / 80.0f;
    p->fixed_hz = 0.0f;
    p->pressure_tol = 1e-4f;
    p->pressure_max_iters = 60;
    p->correction_tol = 1e-2f;
    p->viscosity_tol = 1e-3f;
    p->viscosity_max_iters = 200;
    p->implicit_viscosity_min = 1e-3f;
    p->spray_cfl = 4.0f;
    p->speed_cap = 1.0f;
    p->correction_time = (1.0f / 240.0f) / 0.69314718f;
    p->correction_max = 0.5f;
    p->correction_deadband = 0.25f;
    p->correction_step_max = 0.5f;
    p->crowd_jitter = 0.5f;
    p->tool_crowd_max = 2.0f;
    p->apic_clamp = 1.0f;
    p->curvature_max = 0.5f;
    p->slip_cap = 1;
    p->sealed_vent = 1;
    p->min_substep_share = 0.02f;
    // params.rs:433-438
    p->viscous_projection = 1;
    p->viscous_warm = 1;
    p->viscosity_warm = 1;
    p->viscosity_precond = 1;
    p->viscosity_mg_levels = 1;
    p->viscosity_mg_sweeps = 2;
    p->iters_margin = 0.25f;
    p->mg_precision = 0;
    p->heat_rate = 1.0f;  // params.rs:395-399
    p->air_cooling = 10.0f;
    p->contact_heat = 3000.0f;
    p->eddy_heat = 0.17f;
    p->evaporation = 1;
    p->rigid_contact = 1;  // params.rs:400-402
    p->rigid_friction = 0.6f;
    p->elastic_stress_max = 10.0f;  // params.rs:447-448
    p->elastic_step_max = 0.25f;
    p->magnet_demag = 0.0f;  // params.rs:404
    p->steam = 1;  // params.rs:405-412
    p->steam_boil = 0.02f;
    p->steam_condense = 0.005f;
    p->steam_mixing = 1.0f;
    p->steam_rise = 0.4f;
    p->steam_life = 2.0f;
    p->steam_swirl = 0.1f;
    p->steam_eddy = 0.2f;
    p->heat_power = 300.0f;  // params.rs:403
    p->mat_count = SP_BUILTIN_COUNT;
    for (int r = 0; r < SP_MAX_MATERIALS; r++) p->mat[r] = default_material();
    // presets.rs:37-57 water
    {
        sp_material& m = p->mat[SP_ROW_WATER];
        m.temp_model = 2;
        m.temp_b = 578.9f;
        m.temp_c = 137.5f;
        m.melt_lo = -0.5f;
        m.melt_hi = 0.0f;
        m.latent = 334.0e3f;
        m.density_solid = 917.0f;
        m.boil = 100.0f;
        m.latent_vap = 2257.0e3f;
        m.emissivity = 0.96f;
    }
    // presets.rs:61-89 honey
    {
        sp_material& m = p->mat[SP_ROW_HONEY];
        m.diluent = SP_ROW_WATER;
        m.density = 1.4f * 998.0f;
        m.viscosity = 0.06f;
        m.temp_model = 1;
        m.temp_b = 10500.0f;
        m.tension = 0.076f;
        m.mix_rate = 1.5f;
        m.heat_capacity = 2300.0f;
        m.conductivity = 0.35f;
        m.expansion = 5.0e-4f;
        m.boil = 110.0f;
        m.latent_vap = 2257.0e3f;
    }
    // presets.rs:92-111 olive oil
    {
        sp_material& m = p->mat[SP_ROW_OIL];
        m.group = 1;
        m.density = 0.9f * 998.0f;
        m.viscosity = 6.0e-5f;
        m.temp_model = 1;
        m.temp_b = 3900.0f;
        m.tension = 0.032f;
        m.heat_capacity = 1970.0f;
        m.conductivity = 0.17f;
        m.expansion = 7.0e-4f;
        m.melt_lo = -8.0f;
        m.melt_hi = -4.0f;
        m.latent = 150.0e3f;
        m.density_solid = 925.0f;
    }
    // presets.rs:298-327 mercury
    {
        sp_material& m = p->mat[SP_ROW_MERCURY];
        m.group = 2;
        m.density = 13534.0f;
        m.density_solid = 14184.0f;
        m.viscosity = 1.13e-7f;
        m.tension = 0.485f;
        m.contact_angle = 140.0f;
        m.heat_capacity = 139.5f;
        m.conductivity = 8.3f;
        m.emissivity = 0.1f;
        m.boil = 356.7f;
        m.latent_vap = 295.0e3f;
        m.expansion = 1.8e-4f;
        m.melt_lo = -39.0f;
        m.melt_hi = -38.8f;
        m.latent = 11.4e3f;
    }
    // presets.rs:491-519 brine: boils at 108.8 deg C and leaves salt, 14.4 % of its volume
    {
        sp_material& m = p->mat[SP_ROW_BRINE];
        m.diluent = SP_ROW_WATER;
        m.density = 1200.0f;
        m.viscosity = 1.6e-6f;
        m.mix_rate = 1.0f;
        m.tension = 0.082f;
        m.heat_capacity = 3300.0f;
        m.conductivity = 0.57f;
        m.latent = 250.0e3f;
        m.freeze_shift = -21.0f;
        m.boil = 108.8f;
        m.latent_vap = 2257.0e3f;
        m.residue = SP_ROW_SALT;
        m.residue_share = 0.144f;
    }
    // presets.rs:1064-1087 salt: solid at every temperature the tank reaches
    {
        sp_material& m = p->mat[SP_ROW_SALT];
        m.density = 2165.0f;
        m.viscosity = 1.0e-3f;
        m.heat_capacity = 880.0f;
        m.conductivity = 6.5f;
        m.expansion = 1.0e-4f;
        m.melt_lo = 800.0f;
        m.melt_hi = 801.0f;
        m.latent = 482.0e3f;
        m.tension = 0.07f;
    }
    // presets.rs:1089-1107 steam: the vapour's density and viscosity; it carries the mass and
    // heat of the water it boiled from (particle_heat_vol, params.rs:889-892)
    {
        sp_material& m = p->mat[SP_ROW_STEAM];
        m.density = 0.6f;
        m.viscosity = 2.0e-5f;
        m.expansion = 0.0f;
        m.boil = 100.0f;
        m.latent_vap = 2257.0e3f;
        m.emissivity = 0.0f;
    }
}

static bool freezes(const sp_material& m) { return m.melt_hi > -273.0f; }

// MatProps::temp_factor_ln (material.rs:432-442).
static float temp_factor_ln(const sp_material& m, float t) {
    float tk = std::max(t + 273.15f, 1.0f);
    float tref = m.temp_ref + 273.15f;
    float x = 0.0f;
    if (m.temp_model == 1) x = m.temp_b * (1.0f / tk - 1.0f / tref);
    else if (m.temp_model == 2) x = m.temp_b / std::max(tk - m.temp_c, 5.0f) - m.temp_b / std::max(tref - m.temp_c, 5.0f);
    return std::min(std::max(x, -30.0f), 30.0f);
}

static float viscosity_at(const sp_material& m, float t) { return m.viscosity * expf(temp_factor_ln(m, t)); }

static float tension_cells(const sp_sim_params& p, float sigma) {
    float dx = p.cell_size;
    return sigma / (WATER_DENSITY * dx * dx * dx);
}

// SimParams::dt_capillary (params.rs:962-984).
static float dt_capillary(const sp_sim_params& p, const std::vector<int>& rows_in) {
    // steam's surface is the water around the bubble (params.rs:964-965)
    std::vector<int> rows;
    for (int i : rows_in)
        if (i != SP_ROW_STEAM) rows.push_back(i);
    float worst = 0.0f;
    for (int i : rows) {
        const sp_material& m = p.mat[i];
        worst = std::max(worst, tension_cells(p, m.tension) / (m.density / WATER_DENSITY));
        for (int j : rows) {
            const sp_material& n = p.mat[j];
            if (n.group != m.group) {
                float sigma = tension_cells(p, p.pair_tension[m.group][n.group]);
                worst = std::max(worst, sigma / (std::min(m.density, n.density) / WATER_DENSITY));
            }
        }
    }
    if (worst <= 0.0f) return INFINITY;
    return 0.5f * sqrtf(1.0f / (2.0f * 3.14159265f * worst));
}

// Volumetric heat capacity a particle of the row carries, J/(m^3 K): its own, except
// steam, which carries the water it boiled from (particle_heat_vol, params.rs:888-893).
static float heat_vol(const sp_sim_params& p, int row) {
    const sp_material& m = p.mat[row == SP_ROW_STEAM ? SP_ROW_WATER : row];
    return m.density * m.heat_capacity;
}

// SimParams::dt_thermal (params.rs:914-923): a fifth of the fastest cell diffusion time.
static float dt_thermal(const sp_sim_params& p, const std::vector<int>& rows) {
    if (p.heat_rate <= 0.0f) return INFINITY;
    float dx2 = p.cell_size * p.cell_size, dt = INFINITY;
    for (int i : rows) dt = std::min(dt, 0.2f * dx2 * heat_vol(p, i) / std::max(p.mat[i].conductivity, 1e-6f) / p.heat_rate);
    return dt;
}

// SimParams::dt_elastic (params.rs:951-960): elastic waves, up to sqrt(1 + cap) times the
// shear-wave speed sqrt(G / rho), cross at most half a cell per substep (with no cap, as if
// it were 50 G).
static float dt_elastic(const sp_sim_params& p, const std::vector<int>& rows) {
    const float stiff = 1.0f + (p.elastic_stress_max > 0.0f ? p.elastic_stress_max : 50.0f);
    float dt = INFINITY;
    for (int i : rows) {
        const sp_material& m = p.mat[i];
        if (m.shear_modulus > 0.0f) dt = std::min(dt, 0.5f * p.cell_size / sqrtf(stiff * m.shear_modulus / m.density));
    }
    return dt;
}

// FluidSim::dt_limit (mod.rs:2331-2338).
float dt_limit(const sp_sim_params& p, const std::vector<int>& rows, bool thermal) {
    float dt = std::min(p.dt_max, dt_capillary(p, rows));
    if (thermal) dt = std::min(dt, dt_thermal(p, rows));
    return std::min(dt, dt_elastic(p, rows));
}


// FluidSim::viscosity_mode (mod.rs:2768-2792). A thermal tank counts a liquid that can
// freeze or has a temperature law as arbitrarily viscous.
int viscosity_mode(const sp_sim_params& p, const std::vector<int>& rows, bool thermal) {
    float dt = dt_limit(p, rows, thermal);
    float d = 0.0f;
    for (int r : rows) {
        const sp_material& m = p.mat[r];
        float cells = 1.0f / (p.cell_size * p.cell_size);
        float nu;
        if (!thermal) nu = viscosity_at(m, p.ambient_temp) * cells;
        else if (freezes(m) || m.temp_model != 0) nu = INFINITY;
        else nu = m.viscosity * cells;
        d = std::max(d, nu * dt);
    }
    if (d <= 0.0f) return 0;
    if (d <= 0.05f) return 1;
    return 2;
}

// The rows a tank holding `present` needs: those, their diluents, what they leave behind
// when they boil, and the products of reactions whose partner is present
// (SimParams::closure, params.rs:927-941).
std::vector<int> closure(const sp_sim_params& p, const std::vector<int>& present) {
    std::vector<int> rows = present;
    for (size_t k = 0; k < rows.size(); k++) {
        const sp_material& m = p.mat[rows[k]];
        const bool partner = m.react_with >= 0 && std::find(present.begin(), present.end(), m.react_with) != present.end();
        for (int extra : {m.diluent, m.residue, partner ? m.react_into : -1})
            if (extra >= 0 && extra < SP_MAX_MATERIALS && std::find(rows.begin(), rows.end(), extra) == rows.end())
                rows.push_back(extra);
    }
    std::sort(rows.begin(), rows.end());
    return rows;
}

// Tank::layout (mod.rs:271-331): the materials' channels, then the temperature sums of a
// thermal tank, then the three stress channels while an elastic material is present.
bool tank_layout(const sp_sim_params& p, const std::vector<int>& rows, Tank& t, bool thermal) {
    if ((int)rows.size() > SP_MAX_PRESENT) return false;
    t = Tank();
    t.rows = rows;
    int next = SP_ACC_FIRST;
    uint32_t group_mask = 0;
    int primary = -1;
    for (size_t k = 0; k < rows.size(); k++)
        if (p.mat[rows[k]].diluent == rows[0]) {
            primary = (int)k;
            break;
        }
    int families = primary >= 0 ? 1 : 0;
    for (size_t k = 0; k < rows.size(); k++) {
        const sp_material& m = p.mat[rows[k]];
        group_mask |= 1u << std::min(m.group, 3);
        int c[3] = {-1, -1, -1};
        if (k > 0 && (int)k != primary) c[0] = next++;
        if (m.diluent >= 0) {
            c[1] = next++;
            if ((int)k == primary) c[2] = 0;
            else c[2] = families++;
        }
        t.chan.push_back({c[0], c[1], c[2]});
    }
    // the temperature sum, and one per phase group after the first (mod.rs:303-318)
    t.thermal = thermal;
    if (thermal) {
        t.chan_t = next++;
        bool first = true;
        for (int g = 0; g < 4; g++)
            if (group_mask & (1u << g)) {
                if (!first) t.chan_g[g] = next++;
                first = false;
            }
    }
    bool elastic = false;
    for (int r : rows) elastic |= p.mat[r].shear_modulus > 0.0f;
    if (elastic) {
        t.chan_e = next;
        next += 3;
    }
    if (next > SP_ACC_MAX) return false;
    int stride = std::max(next, SP_ACC_MIN);
    stride = (stride + 3) / 4 * 4;
    t.stride = stride;
    t.families = families;
    t.group_mask = group_mask;
    t.primary = primary >= 0 ? rows[primary] : -1;
    return true;
}

// Tank::layout_steam (mod.rs:247-264): with steam and the water it condenses into added to
// the rows while the tank is thermal, boiling makes steam and an aqueous liquid in it can
// boil; without them if they do not fit (boiled liquid then leaves at once).
bool tank_layout_steam(const sp_sim_params& p, const std::vector<int>& rows, Tank& t, bool thermal) {
    bool boils = false;
    for (int r : rows) boils |= r != SP_ROW_STEAM && p.mat[r].group == 0 && p.mat[r].boil < 1.0e8f;
    if (thermal && p.steam != 0 && boils && std::find(rows.begin(), rows.end(), SP_ROW_STEAM) == rows.end()) {
        std::vector<int> more = rows;
        for (int r : {SP_ROW_WATER, SP_ROW_STEAM})
            if (std::find(more.begin(), more.end(), r) == more.end()) more.push_back(r);
        std::sort(more.begin(), more.end());
        if (tank_layout(p, more, t, thermal)) return true;
    }
    return tank_layout(p, rows, t, thermal);
}

// FluidSim::caps (mod.rs:2717-2766) for the CAP-gated passes this build carries: families,
// heat, rigid bodies, rheology, viscoelasticity (stress channels in the layout) and settling. A non-Newtonian liquid, or one with a skin or internal
// friction, turns rheology on, and rigid bodies for a yield stress, shear thickening, a skin
// or friction. A thermal tank turns rigid bodies on for any liquid that can freeze or has a
// temperature law; any tank for a liquid already as viscous as rigid_viscosity, or frozen at
// the ambient temperature.
static uint32_t tank_caps(const sp_sim_params& p, const Tank& t, size_t n_magnets) {
    uint32_t caps = 0;
    if (t.families > 0) caps |= SP_CAP_FAMILIES;
    if (t.thermal) caps |= SP_CAP_THERMAL;
    bool rigid = false;
    for (int r : t.rows) {
        const sp_material& m = p.mat[r];
        const sp_material* dil = m.diluent >= 0 ? &p.mat[m.diluent] : nullptr;
        if (m.rheology != 0 || m.skin > 0.0f || m.friction > 0.0f) {
            caps |= SP_CAP_RHEO;
            rigid |= m.rheology == 2 || m.rheology == 3 || m.skin > 0.0f || m.friction > 0.0f;
        }
        if (t.thermal)
            rigid |= freezes(m) || (dil && freezes(*dil)) || m.temp_model != 0 || (dil && dil->temp_model != 0);
        rigid |= viscosity_at(m, p.ambient_temp) >= p.rigid_viscosity || (!t.thermal && freezes(m) && p.ambient_temp < m.melt_hi);
    }
    if (rigid) caps |= SP_CAP_RIGID;
    if (t.chan_e >= 0) caps |= SP_CAP_ELASTIC;
    // FluidSim::settles (mod.rs:2139-2143): a family whose concentrate has grains, under gravity
    bool settles = false;
    for (int r : t.rows) settles |= p.mat[r].diluent >= 0 && p.mat[r].grain > 0.0f;
    if (t.families > 0 && p.gravity > 0.0f && settles) caps |= SP_CAP_SETTLE;
    // a magnet and a liquid with a susceptibility (mod.rs:2749-2751)
    bool magnetic = false;
    for (int r : t.rows) magnetic |= p.mat[r].susceptibility > 0.0f;
    if (n_magnets > 0 && magnetic) caps |= SP_CAP_MAGNET;
    // steam carried in a thermal tank (mod.rs:2752-2756); and, for this build's kernel choice,
    // a liquid that can boil there, and puffs of vapour that escaped steam or a fuming liquid
    // leaves in the air
    const bool steam = t.thermal && std::find(t.rows.begin(), t.rows.end(), SP_ROW_STEAM) != t.rows.end();
    if (steam) caps |= SP_CAP_STEAM | SP_CAP_PUFFS;
    for (int r : t.rows) {
        const sp_material& m = p.mat[r];
        const bool boils = t.thermal && m.boil < 1.0e8f;
        if (boils) caps |= SP_CAP_BOILS;
        if (m.fume > 0.0f && (m.volatility > 0.0f || boils)) caps |= SP_CAP_PUFFS;
        if (m.volatility > 0.0f) caps |= SP_CAP_VOLATILE;
        if (t.thermal && m.self_heat > 0.0f) caps |= SP_CAP_SELF_HEAT;
        // two liquids that react with each other (mod.rs:2756-2758)
        if (m.react_with >= 0 && m.react_rate > 0.0f && std::find(t.rows.begin(), t.rows.end(), m.react_with) != t.rows.end())
            caps |= SP_CAP_REACT;
    }
    return caps;
}

// ToolState::speed (params.rs:1202-1210): the fastest velocity the tool imposes, cells / s.
static float tool_speed(const sp_tool_state& t) {
    switch (t.mode) {
    case 1: return hypotf(t.vel[0], t.vel[1]) + 4.0f * t.radius;
    case 3:
    case 4: return hypotf(t.vel[0], t.vel[1]);
    case 7: return fabsf(t.vel[0]) + 2.0f * t.radius;
    default: return 0.0f;
    }
}

int derive(const sp_sim_params& p, int W, int H, const Tank& t, uint32_t n_slots, sp_gparams& g, sp_gmat* mats, std::string* why,
           const std::vector<sp_magnet>& magnets, float ext_speed, const std::vector<sp_current>& currents,
           const sp_tool_state* tool) {
    memset(&g, 0, sizeof(g));
    memset(mats, 0, sizeof(sp_gmat) * SP_MAX_MATERIALS);
    const float dx = p.cell_size;
    // gpu_materials (mod.rs:2179-2236), isothermal.
    for (int r = 0; r < SP_MAX_MATERIALS; r++) {
        const sp_material& m = p.mat[r];
        int chan[3] = {-1, -1, -1};
        for (size_t k = 0; k < t.rows.size(); k++)
            if (t.rows[k] == r) {
                chan[0] = t.chan[k][0];
                chan[1] = t.chan[k][1];
                chan[2] = t.chan[k][2];
            }
        float ang = m.contact_angle * 0.017453292f;
        float sn = sinf(ang), cs = cosf(ang);
        float solid = m.density_solid > 0.0f ? m.density_solid : m.density;
        float ln_nu = logf(std::max(m.viscosity / (dx * dx), 1e-8f));
        float rho = m.density;
        // An isothermal tank: the temperature law and the frozen state are constants,
        // folded in here (mod.rs:2194-2206); a thermal tank evaluates them per cell.
        if (!t.thermal) {
            ln_nu += temp_factor_ln(m, p.ambient_temp);
            rho *= std::min(std::max(1.0f - m.expansion * (p.ambient_temp - m.temp_ref), 0.7f), 1.3f);
            if (freezes(m) && m.diluent < 0 && p.ambient_temp < m.melt_hi) {
                float fs = std::min(std::max((m.melt_hi - p.ambient_temp) / std::max(m.melt_hi - m.melt_lo, 1e-3f), 0.0f), 1.0f);
                float frozen = std::max(logf(p.rigid_viscosity / (dx * dx)), ln_nu);
                ln_nu += (frozen - ln_nu) * fs;
                rho += (solid - rho) * fs;
            }
        }
        sp_gmat& o = mats[r];
        o.group = std::min(m.group, 3);
        o.diluent = m.diluent;
        o.chan_w = chan[0];
        o.chan_s = chan[1];
        o.fam = chan[2];
        o.rheo = (int32_t)m.rheology;
        o.temp_model = m.temp_model;
        o.residue = m.residue;
        o.residue_share = m.residue_share;  // mod.rs:2215 el.w
        o.fume = m.fume;                    // mod.rs:2232 gr.y
        o.rho = rho / WATER_DENSITY;
        o.ln_nu = ln_nu;
        o.sigma = tension_cells(p, m.tension);
        o.mix_rate = m.mix_rate;
        o.cos_c = cs;
        o.sin_c = sn;
        o.grain = m.grain / dx;  // mod.rs:2210
        o.rho_solid = solid / WATER_DENSITY;
        o.visc_exp = m.visc_exponent;
        o.expansion = m.expansion;
        o.load = m.load;
        o.freeze_shift = m.freeze_shift;
        // thermal rows (mod.rs:2213-2214)
        o.th[0] = heat_vol(p, r) / WATER_HEAT;
        o.th[1] = m.conductivity / (WATER_HEAT * dx * dx);
        o.th[2] = std::max(m.melt_lo, NO_MELT);
        o.th[3] = std::max(std::max(m.melt_hi, m.melt_lo + 0.25f), NO_MELT);
        o.th2[0] = (freezes(m) || m.diluent >= 0) ? m.latent / m.heat_capacity : 0.0f;
        o.th2[1] = m.emissivity;
        o.th2[2] = m.boil;
        o.th2[3] = m.latent_vap / m.heat_capacity;
        o.temp[0] = m.temp_b;
        o.temp[1] = m.temp_c;
        o.temp[2] = m.temp_ref + 273.15f;
        o.temp[3] = tension_cells(p, m.skin);
        // rheology (mod.rs:2184, 2211): the yield stress in cells^2/s^2 per unit density
        o.rheo_p[0] = m.yield_stress / (WATER_DENSITY * dx * dx);
        o.rheo_p[1] = m.flow_index;
        o.rheo_p[2] = m.shear_time;
        o.rheo_p[3] = m.visc_ratio;
        o.gr[0] = m.friction;
        // viscoelasticity (mod.rs:2215): the shear modulus as a stress per unit density
        o.gr[1] = m.shear_modulus / (WATER_DENSITY * dx * dx);
        o.gr[2] = m.relax_time;
        o.gr[3] = m.susceptibility;  // mod.rs:2215 el.z
        o.rx[3] = m.volatility;      // mod.rs:2231
        o.rx[0] = m.react_rate;
        o.rx[1] = m.react_heat / m.heat_capacity;
        o.rx[2] = m.self_heat / m.heat_capacity;
        // the reaction while the partner is in the tank (mod.rs:2217-2230)
        int at = -1;
        for (size_t k = 0; k < t.rows.size(); k++)
            if (m.react_with >= 0 && t.rows[k] == m.react_with) at = (int)k;
        o.ic[0] = at >= 0 && m.react_rate > 0.0f ? m.react_with : -1;
        o.ic[1] = m.react_into;
        o.ic[2] = at >= 0 ? at : 0;
        o.ic[3] = 0;
    }
    const uint32_t caps = tank_caps(p, t, magnets.size());
    // gpu_params (mod.rs:2238-2330).
    g.W = W;
    g.H = H;
    g.n_slots = n_slots;
    g.spawn_count = 0;
    g.dt_fixed = p.fixed_hz > 0.0f ? 1.0f / p.fixed_hz : 0.0f;
    g.gravity = p.gravity / p.cell_size;
    g.inv_ppc = 1.0f / (float)SP_PPC;
    g.apic_clamp = p.apic_clamp;
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) g.pair[a][b] = tension_cells(p, p.pair_tension[a][b]);
    g.kappa_max = p.curvature_max;
    g.sealed_vent = (float)p.sealed_vent;
    g.cell_size = p.cell_size;
    g.rigid_friction = p.rigid_contact != 0 ? p.rigid_friction : -1.0f;  // misc.w (mod.rs:2280)
    g.n_present = (int)t.rows.size();
    g.acc_stride = std::max(t.stride, SP_ACC_MIN);
    g.n_fam = t.families;
    g.group_mask = t.group_mask;
    g.caps = caps;
    g.primary = t.rows.empty() ? -1 : t.primary;
    g.group_heat = 0;
    for (int q = 0; q < 4; q++) g.group_heat += ((uint32_t)(t.chan_g[q] + 1) & 255u) << (8 * q);
    g.visc_mode = viscosity_mode(p, t.rows, t.thermal);
    for (int k = 0; k < SP_MAX_PRESENT; k++) {
        g.present[k] = 0;
        g.fams[k][0] = -1;
        g.fams[k][1] = -1;
    }
    for (size_t k = 0; k < t.rows.size(); k++) {
        g.present[k] = t.rows[k];
        if (t.chan[k][2] >= 0) {
            g.fams[t.chan[k][2]][0] = t.rows[k];
            g.fams[t.chan[k][2]][1] = p.mat[t.rows[k]].diluent;
        }
    }
    g.sep_tau = p.droplet_cells;
    g.cfl = p.cfl;
    g.tool_crowd_max = p.tool_crowd_max;
    g.density_k = p.correction_max;
    g.dt_max = dt_limit(p, t.rows, t.thermal);
    g.crowd_jitter = p.crowd_jitter;
    g.wall_margin = 0.1f;
    g.tol[0] = p.pressure_tol;
    g.tol[1] = p.correction_tol;
    g.tol[2] = p.viscosity_tol;
    g.tol[3] = p.correction_time;
    {
        float gc = p.gravity / p.cell_size;
        g.ext[0] = (p.speed_cap > 0.0f && gc > 0.0f) ? p.speed_cap * sqrtf(2.0f * gc * (float)H) : 1e30f;
    }
    g.ext[1] = (float)p.advection_order;
    g.ext[2] = p.spray_cfl;
    // currents and drains (mod.rs:2265-2270, 2309): zone (position, radius at least 0.01, rate)
    // and velocity; their number in ext.w
    const size_t nc = std::min<size_t>(currents.size(), SP_MAX_CURRENTS);
    for (size_t k = 0; k < nc; k++) {
        const sp_current& cu = currents[k];
        g.current_zone[k][0] = cu.pos[0];
        g.current_zone[k][1] = cu.pos[1];
        g.current_zone[k][2] = std::max(cu.radius, 0.01f);
        g.current_zone[k][3] = cu.rate;
        g.current_vel[k][0] = cu.vel[0];
        g.current_vel[k][1] = cu.vel[1];
    }
    g.ext[3] = (float)nc;
    if (nc > 0) g.caps |= SP_CAP_CURRENTS;
    // the tool (mod.rs:2300-2305); the speed it imposes or the fastest queued spawn's,
    // whichever is larger
    const sp_tool_state idle{};
    const sp_tool_state& tl = tool ? *tool : idle;
    g.ext_speed = std::max(tool_speed(tl), std::max(ext_speed, 0.0f));
    g.tool_pos[0] = tl.pos[0];
    g.tool_pos[1] = tl.pos[1];
    g.tool_vel[0] = tl.vel[0];
    g.tool_vel[1] = tl.vel[1];
    g.tool_radius = std::max(tl.radius, 0.01f);
    g.tool_mode = tl.mode;
    g.tool_strength = tl.strength;
    g.heat_power = p.heat_power;
    // steam (mod.rs:2322-2323)
    g.steam[0] = std::min(std::max(p.steam_condense, 0.0f), 1.0f);
    g.steam[1] = p.steam_rise / dx;
    g.steam[2] = p.steam_life;
    g.steam[3] = std::min(std::max(p.steam_boil, g.steam[0]), 1.0f);
    g.steam2[0] = p.steam_mixing;
    g.steam2[1] = (float)p.steam;
    g.steam2[2] = p.steam_swirl / dx;
    g.steam2[3] = p.steam_eddy / dx;
    g.corr[0] = p.correction_step_max;
    g.corr[1] = p.correction_deadband;
    g.corr[2] = (float)p.slip_cap;
    g.corr[3] = p.implicit_viscosity_min;
    g.p_max_iters = (int)p.pressure_max_iters;
    g.v_max_iters = (int)p.viscosity_max_iters;
    g.heat_x = p.ambient_temp;
    g.ln_rigid = logf(p.rigid_viscosity / (dx * dx));
    // The implicit viscosity's options: twice_projected (mod.rs:2712-2715), the cg_scale warm
    // mode (mod.rs:2978-2979), visc_mg_levels / visc_mg_sweeps (mod.rs:2353-2366).
    const bool implicit = g.visc_mode == 2;
    g.v_twice = implicit && p.viscous_projection != 0 ? 1 : 0;
    g.v_p2warm = g.v_twice && p.viscous_warm != 0 ? 1 : 0;
    g.v_warm = p.viscosity_warm != 0 ? 1 : 0;
    g.v_precond = p.viscosity_precond != 0 ? 1 : 0;
    {
        sp_levels_info L;
        levels(W, H, L);
        int want = std::max((int)p.viscosity_mg_levels, 1);
        g.v_mg_levels = std::min(want, L.tail);
    }
    g.v_mg_sweeps = (int)std::min(std::max(p.viscosity_mg_sweeps, 1u), 2u);
    g.mg_half = p.mg_precision == 1u ? 1 : 0;
    // heat (mod.rs:2299-2300, 2321)
    g.heat_rate = p.heat_rate;
    g.film = p.air_cooling / (WATER_HEAT * dx);
    g.radiation = STEFAN_BOLTZMANN / (WATER_HEAT * dx);
    g.contact = p.contact_heat / (WATER_HEAT * dx);
    g.eddy = std::max(p.eddy_heat, 0.0f);
    g.evap = p.evaporation != 0 ? 2.257e6f / 1005.0f : 0.0f;
    g.chan_t = t.chan_t;
    // viscoelasticity (mod.rs:2316-2321)
    float gmax = 0.0f;
    for (int r : t.rows) gmax = std::max(gmax, p.mat[r].shear_modulus);
    g.chan_e = t.chan_e;
    g.el_g = std::max(gmax, 1e-3f) / (WATER_DENSITY * dx * dx);
    g.el_stress_max = p.elastic_stress_max;
    g.el_step_max = p.elastic_step_max;
    // magnetism (mod.rs:2258-2264, 2313, 2321): the most susceptible liquid's susceptibility
    // and saturation (the first of equals; 0.05 T with none)
    float chi_max = 0.0f, sat = 0.05f;
    for (int r : t.rows)
        if (p.mat[r].susceptibility > chi_max) {
            chi_max = p.mat[r].susceptibility;
            sat = p.mat[r].saturation;
        }
    const size_t nm = std::min<size_t>(magnets.size(), SP_MAX_MAGNETS);
    g.mag[0] = (float)nm;
    g.mag[1] = std::max(chi_max, 1e-6f);
    g.mag[2] = sat;
    g.mag[3] = 1.0f / (WATER_DENSITY * dx * dx);
    for (size_t k = 0; k < nm; k++) {
        const sp_magnet& mg = magnets[k];
        g.magnet[k][0] = mg.pos[0];
        g.magnet[k][1] = mg.pos[1];
        g.magnet[k][2] = std::max(mg.radius, 0.5f);
        g.magnet[k][3] = mg.strength;
        g.magnet_dir[k][0] = mg.dir[0];
        g.magnet_dir[k][1] = mg.dir[1];
    }
    g.demag = std::min(std::max(p.magnet_demag, 0.0f), 1.0f);
    return SP_OK;
}

// Levels::new (mod.rs:727-753).
void levels(int W, int H, sp_levels_info& L) {
    memset(&L, 0, sizeof(L));
    L.dims[0][0] = W;
    L.dims[0][1] = H;
    int n = 1;
    while (L.dims[n - 1][0] * L.dims[n - 1][1] > SP_COARSEST_CELLS && n < SP_MAX_LEVELS) {
        L.dims[n][0] = (L.dims[n - 1][0] + 1) / 2;
        L.dims[n][1] = (L.dims[n - 1][1] + 1) / 2;
        n++;
    }
    L.n = n;
    int total = 0;
    for (int l = 0; l < n; l++) {
        L.offs[l] = total;
        total += L.dims[l][0] * L.dims[l][1];
    }
    L.total = total;
    int tail = n - 1;
    for (int l = 1; l < n; l++)
        if (L.dims[l][0] * L.dims[l][1] <= SP_TAIL_CELLS) {
            tail = l;
            break;
        }
    tail = std::max(1, std::min(tail, n - 1));
    L.tail = tail;
    int tiles = 0;
    for (int l = 0; l <= tail; l++) {
        L.tile_offs[l] = tiles;
        tiles += ((L.dims[l][0] + 15) / 16) * ((L.dims[l][1] + 15) / 16);
    }
    L.tiles = tiles;
}

// Particle::settle_spawn (params.rs:70-88), isothermal tanks.
void settle_spawn(const sp_sim_params& p, sp_particle& q) {
    uint32_t low = q.kind & 255u;
    if (q.kind == SP_KIND_DEAD || low == 0u) return;
    const sp_material& def = p.mat[low - 1];
    if (q.t == -1.0e4f) {
        q.t = def.spawn_temp;
        q.kind |= SP_KIND_REAL_T;
    }
    if (q.s == -1.0f) q.s = def.diluent >= 0 ? std::min(std::max(def.load, 0.0f), 1.0f) : 1.0f;
    if (q.kind & SP_KIND_REAL_T) {
        q.kind &= ~SP_KIND_REAL_T;
        // SimParams::enthalpy_temp (params.rs:844-855): the temperature itself above the
        // melting range, which is all an isothermal tank carries.
        const sp_material& own = def;
        float lo, hi, l;
        if (own.diluent < 0) {
            lo = own.melt_lo;
            hi = std::max(own.melt_hi, own.melt_lo + 0.25f);
            l = freezes(own) ? own.latent / own.heat_capacity : 0.0f;
        } else {
            const sp_material& d = p.mat[own.diluent];
            lo = d.melt_lo;
            hi = std::max(d.melt_hi, d.melt_lo + 0.25f);
            l = 0.0f;
        }
        float t = q.t;
        if (lo < -273.0f || t >= hi) q.t = t;
        else if (t <= lo) q.t = t - l;
        else {
            float fr = (t - lo) / std::max(hi - lo, 1e-3f);
            q.t = (lo - l) + fr * (hi - lo + l);
        }
    }
}

}  // namespace sp
