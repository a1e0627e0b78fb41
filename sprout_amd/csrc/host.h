// host.h — host-side declarations of the MI355X fluid step library.
#pragma once

#include <array>
#include <string>
#include <vector>

#include "../../include/sprout_hip.h"

namespace sp {

// What the tank holds and the accumulator channel of each row (Tank, mod.rs:225-245).
struct Tank {
    std::vector<int> rows;
    std::vector<std::array<int, 3>> chan;  // weight channel, concentration channel, family
    int stride = SP_ACC_MIN;
    int families = 0;
    uint32_t group_mask = 0;
    int primary = -1;
    bool thermal = false;                   // temperatures carried and conducted
    int chan_t = -1;                        // channel of the temperature sum
    std::array<int, 4> chan_g{-1, -1, -1, -1};  // per phase group after the first: its own sum
    int chan_e = -1;                        // first of the three stress channels (-1: nothing elastic)
};

void default_params(sp_sim_params* p);
std::vector<int> closure(const sp_sim_params& p, const std::vector<int>& present);
bool tank_layout(const sp_sim_params& p, const std::vector<int>& rows, Tank& t, bool thermal = false);
// The layout with steam and water added where boiling makes steam (Tank::layout_steam).
bool tank_layout_steam(const sp_sim_params& p, const std::vector<int>& rows, Tank& t, bool thermal);
// ext_speed: the fastest speed a pour imposes this record, cells / s (the queued spawns'
// fastest |vel|; the tool's is 0 for the idle tool this build carries).
int derive(const sp_sim_params& p, int W, int H, const Tank& t, uint32_t n_slots, sp_gparams& g, sp_gmat* mats,
           std::string* why, const std::vector<sp_magnet>& magnets = {}, float ext_speed = 0.0f,
           const std::vector<sp_current>& currents = {}, const sp_tool_state* tool = nullptr);
void levels(int W, int H, sp_levels_info& L);
int viscosity_mode(const sp_sim_params& p, const std::vector<int>& rows, bool thermal = false);
float dt_limit(const sp_sim_params& p, const std::vector<int>& rows, bool thermal = false);
void settle_spawn(const sp_sim_params& p, sp_particle& q);

}  // namespace sp
