// schedule.hip — the substep schedule on a HIP stream (record_substep, mod.rs:2615-2708):
// which kernels run, in which order, with which recorded iteration counts, for one launch
// group of T tanks of the same grid and solver schedule (a Run). Every launch covers all T
// tanks (blockIdx.z = tank), each tank reading its own device state through its context, so
// a group's substep costs the same number of launches as one tank's. A single handle is the
// group of one.
//
// Parts of a substep are captured once per (schedule, recorded iteration count) as HIP graphs and
// replayed: a substep is several hundred dependent launches. The kernels read everything
// that changes between substeps (dt, tables, solve state) from device memory, so a replay
// needs no new arguments.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <vector>

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCHK(x)                                     \
    do {                                              \
        hipError_t e_ = (x);                          \
        if (e_ != hipSuccess) throw HipError{e_, #x}; \
    } while (0)

// One launch group.
struct Run {
    TankBatch tb{};  // the T tanks (kernel argument of every launch)
    int T = 1;
    int W = 0, H = 0, nb0 = 0;  // grid, and its 16 x 16 tiles
    sp_levels_info lv{};
    sp_gparams gp{};    // host copy of the tables: the schedule fields are equal across the group
    hipStream_t stream = nullptr;
    size_t tail_lds = 0;  // dynamic LDS of k_mg_tail
    bool fuse = true;     // coarse multigrid levels two per launch (SPROUT_MG_FUSE=0: one)
    uint32_t slots = 0;   // most particle slots of any tank of the group
};

// Recorded iterations of one substep: first pressure solve, second pressure solve (after
// the viscosity), viscosity solve (record_substep, mod.rs:2615-2623).
struct Recs {
    uint32_t p = 0, p2 = 0, v = 0;
    bool operator==(const Recs& o) const { return p == o.p && p2 == o.p2 && v == o.v; }
};

namespace {

const dim3 TILE(16, 16, 1);
dim3 groups_of(const Run& r, int w, int h) { return dim3((w + 15) / 16, (h + 15) / 16, r.T); }
dim3 per_tank(const Run& r) { return dim3(1, 1, r.T); }
dim3 grid0(const Run& r) { return groups_of(r, r.W, r.H); }

// ---- pressure (record_pressure / record_cg / record_vcycle, mod.rs:2929-3031) -------------

Lv lv_of(const sp_levels_info& L, int l) { return Lv{L.offs[l], L.dims[l][0], L.dims[l][1], L.tile_offs[l]}; }
dim3 grid_t0(const Run& r) { return dim3((r.W + T0 - 1) / T0, (r.H + T0 - 1) / T0, r.T); }
int parts_t0(const Run& r) { return ((r.W + T0 - 1) / T0) * ((r.H + T0 - 1) / T0); }

// The coarse levels 1 .. tail - 1 of a V-cycle, on the way down and on the way up: two
// levels per launch, pairing from the coarsest (level 1, the largest, alone when their
// number is odd), or one per launch without `fuse`. H: the multigrid's storage.
template <bool H>
void launch_coarse_down(const Run& r, int gated, int par) {
    const sp_levels_info& L = r.lv;
    int l = 1;
    if (!r.fuse || (L.tail - 1) % 2 == 1) {
        for (; l < L.tail && (l == 1 || !r.fuse); l++)
            hipLaunchKernelGGL(k_mg_down<H>, groups_of(r, L.dims[l][0], L.dims[l][1]), TILE, 0, r.stream, r.tb, lv_of(L, l),
                               lv_of(L, l + 1), l, gated, par, 0);
    }
    for (; l + 1 < L.tail; l += 2)
        hipLaunchKernelGGL(k_mg_down2<H>, groups_of(r, L.dims[l + 1][0], L.dims[l + 1][1]), dim3(32, 32), DOWN2_LDS, r.stream, r.tb,
                           lv_of(L, l), lv_of(L, l + 1), lv_of(L, l + 2), l, gated, par);
}

template <bool H>
void launch_coarse_up(const Run& r, int gated, int par) {
    const sp_levels_info& L = r.lv;
    const int single = !r.fuse ? L.tail - 1 : (L.tail - 1) % 2;  // levels 1 .. single run one per launch
    for (int l = L.tail - 2; l > single; l -= 2)
        hipLaunchKernelGGL(k_mg_up2<H>, groups_of(r, L.dims[l + 1][0], L.dims[l + 1][1]), dim3(32, 32), UP2_LDS, r.stream, r.tb,
                           lv_of(L, l), lv_of(L, l + 1), lv_of(L, l + 2), l, gated, par);
    for (int l = single; l >= 1; l--)
        hipLaunchKernelGGL(k_mg_up<H>, groups_of(r, L.dims[l][0], L.dims[l][1]), TILE, 0, r.stream, r.tb, lv_of(L, l),
                           lv_of(L, l + 1), l, gated, par);
}

// One V-cycle from the level-0 residual (mgr, or r2 with rsel), ungated: the first of a solve
// (reduce: with the partial sums r.z, r.r) and the debug stage.
template <bool H>
void launch_vcycle_t(const Run& r, int reduce, int rsel) {
    const sp_levels_info& L = r.lv;
    hipLaunchKernelGGL(k_mg_down<H>, grid0(r), TILE, 0, r.stream, r.tb, lv_of(L, 0), lv_of(L, 1), 0, 0, 0, rsel);
    launch_coarse_down<H>(r, 0, 0);
    hipLaunchKernelGGL(k_mg_tail<H>, per_tank(r), dim3(TAIL_NT), r.tail_lds, r.stream, r.tb, 0, 0);
    launch_coarse_up<H>(r, 0, 0);
    hipLaunchKernelGGL(k_up0<H>, grid_t0(r), dim3(T0, T0), 0, r.stream, r.tb, lv_of(L, 0), lv_of(L, 1), parts_t0(r), 0, 0, reduce,
                       rsel);
}

void launch_vcycle(const Run& r, int reduce, int rsel) {
    if (r.gp.mg_half) launch_vcycle_t<true>(r, reduce, rsel);
    else launch_vcycle_t<false>(r, reduce, rsel);
}

template <bool H>
void launch_mg_build_t(const Run& r) {
    const sp_levels_info& L = r.lv;
    for (int l = 1; l <= L.tail; l++)
        hipLaunchKernelGGL(k_mg_build<H>, groups_of(r, L.dims[l][0], L.dims[l][1]), TILE, 0, r.stream, r.tb, l);
    if (L.tail < L.n - 1) hipLaunchKernelGGL(k_mg_tail_build<H>, per_tank(r), dim3(1024), 0, r.stream, r.tb);
}

void launch_mg_build(const Run& r) {
    if (r.gp.mg_half) launch_mg_build_t<true>(r);
    else launch_mg_build_t<false>(r);
}

// The first solve builds the hierarchy and rescales the warm start; the second (after the
// viscosity, `second`) reuses both and starts from what rhs_visc left. Iteration k runs the
// scalar step that ended iteration k - 1, cg_apply, alpha, the update with the way down on
// level 0, the coarse levels, the tail and the way up; its parity k & 1 picks the
// solve-state slot, the residual buffer (mgr / r2) and the direction (cgp / cgp2).
template <bool H>
void launch_pressure_t(const Run& r, uint32_t iters, bool second) {
    const sp_levels_info& L = r.lv;
    const Lv L0 = lv_of(L, 0), L1 = lv_of(L, 1);
    const int nbp = parts_t0(r);
    const dim3 gt0 = grid_t0(r), bt0(T0, T0);
    if (!second) {
        launch_mg_build_t<H>(r);
        hipLaunchKernelGGL(k_cg_scale, grid0(r), TILE, 0, r.stream, r.tb);
    }
    hipLaunchKernelGGL(k_cg_init<H>, gt0, bt0, 0, r.stream, r.tb, L0, nbp);
    launch_vcycle_t<H>(r, 1, 0);
    for (uint32_t k = 0; k < iters; k++) {
        const int par = (int)(k & 1);
        hipLaunchKernelGGL(k_cg_step, per_tank(r), dim3(REDUCE_NT), 0, r.stream, r.tb, nbp, par, k == 0 ? 1 : 0, second ? 1 : 0);
        hipLaunchKernelGGL(k_cg_apply<H>, gt0, bt0, 0, r.stream, r.tb, L0, nbp, par);
        hipLaunchKernelGGL(k_cg_alpha, per_tank(r), dim3(REDUCE_NT), 0, r.stream, r.tb, nbp, par);
        hipLaunchKernelGGL(k_update_down0<H>, gt0, bt0, 0, r.stream, r.tb, L0, L1, par);
        launch_coarse_down<H>(r, 1, par);
        hipLaunchKernelGGL(k_mg_tail<H>, per_tank(r), dim3(TAIL_NT), r.tail_lds, r.stream, r.tb, 1, par);
        launch_coarse_up<H>(r, 1, par);
        hipLaunchKernelGGL(k_up0<H>, gt0, bt0, 0, r.stream, r.tb, L0, L1, nbp, 1, par, 1, par ^ 1);
    }
    hipLaunchKernelGGL(k_cg_end, per_tank(r), dim3(REDUCE_NT), 0, r.stream, r.tb, nbp, (int)iters, second ? 1 : 0);
}

void launch_pressure(const Run& r, uint32_t iters, bool second = false) {
    if (r.gp.mg_half) launch_pressure_t<true>(r, iters, second);
    else launch_pressure_t<false>(r, iters, second);
}

// ---- the implicit viscosity (record_viscosity_split / _solve, mod.rs:2812-2927) ----------

// Weights and active tiles, the explicit step outside them, the Jacobi diagonal, the warm
// start, and the multigrid's matrices.
void launch_visc_split(const Run& r) {
    const dim3 g0 = grid0(r);
    hipLaunchKernelGGL(k_clear_vtiles, per_tank(r), dim3(256), 0, r.stream, r.tb);
    hipLaunchKernelGGL(k_visc_setup, g0, TILE, 0, r.stream, r.tb, 1);
    hipLaunchKernelGGL(k_visc_split, g0, TILE, 0, r.stream, r.tb);
    hipLaunchKernelGGL(k_visc_diag, g0, TILE, 0, r.stream, r.tb);
    if (r.gp.v_precond)
        for (int l = 1; l <= r.gp.v_mg_levels; l++)
            hipLaunchKernelGGL(k_vmg_build, groups_of(r, r.lv.dims[l][0], r.lv.dims[l][1]), TILE, 0, r.stream, r.tb, l);
}

void launch_visc_reduce(const Run& r, int mode) {
    hipLaunchKernelGGL(k_visc_reduce, per_tank(r), dim3(REDUCE_NT), 0, r.stream, r.tb, mode, r.nb0);
}

// The coarse levels of a viscosity V-cycle, down and up (record_visc_coarse, mod.rs:2838-2853).
template <int NU>
void launch_visc_coarse(const Run& r) {
    const int n = r.gp.v_mg_levels;
    for (int l = 1; l <= n; l++)
        hipLaunchKernelGGL(k_vmg_down<NU>, groups_of(r, r.lv.dims[l][0], r.lv.dims[l][1]), TILE, 0, r.stream, r.tb, l,
                           l == n ? 1 : 0);
    for (int l = n; l >= 1; l--)
        hipLaunchKernelGGL(k_vmg_up<NU>, groups_of(r, r.lv.dims[l][0], r.lv.dims[l][1]), TILE, 0, r.stream, r.tb, l,
                           l == n ? 1 : 0);
}

// One V-cycle of the viscosity multigrid from vr: z, A z in vmz, vaz (the debug stage).
template <int NU>
void launch_visc_vcycle(const Run& r) {
    const dim3 g0 = grid0(r);
    hipLaunchKernelGGL(k_vmg_down<NU>, g0, TILE, 0, r.stream, r.tb, 0, 0);
    launch_visc_coarse<NU>(r);
    hipLaunchKernelGGL(k_visc_mgz<NU>, g0, TILE, 0, r.stream, r.tb, 0);
}

// Multigrid-preconditioned CG: z of the starting residual and the first step length, then
// per iteration the update with the way down on level 0, the coarse levels and the way up
// with A z and both step lengths. Residual and A p alternate between (vr, vq) and
// (vr2, vq2) with the iteration's parity.
template <int NU>
void launch_visc_mg(const Run& r, uint32_t iters) {
    const dim3 g0 = grid0(r);
    launch_visc_vcycle<NU>(r);
    launch_visc_reduce(r, VRED_MG_START);
    for (uint32_t k = 0; k < iters; k++) {
        const int par = (int)(k & 1);
        hipLaunchKernelGGL(k_visc_mgu<NU>, g0, TILE, 0, r.stream, r.tb, par);
        launch_visc_reduce(r, VRED_MG_UPDATE);
        launch_visc_coarse<NU>(r);
        hipLaunchKernelGGL(k_visc_mgz<NU>, g0, TILE, 0, r.stream, r.tb, par ^ 1);
        launch_visc_reduce(r, VRED_MG_ITER);
    }
}

// The implicit solve on the active tiles. `restrain`: the first pressure solve changed the
// velocity since the split, so the strain of u* is taken again.
void launch_visc_solve(const Run& r, uint32_t iters, bool restrain) {
    const dim3 g0 = grid0(r);
    if (restrain) hipLaunchKernelGGL(k_visc_strain, g0, TILE, 0, r.stream, r.tb);
    hipLaunchKernelGGL(k_visc_start, g0, TILE, 0, r.stream, r.tb);
    launch_visc_reduce(r, VRED_START);
    if (r.gp.v_warm) {
        hipLaunchKernelGGL(k_visc_warm, g0, TILE, 0, r.stream, r.tb);
        launch_visc_reduce(r, VRED_WARM);
    }
    if (r.gp.v_precond) {
        if (r.gp.v_mg_sweeps == 1) launch_visc_mg<1>(r, iters);
        else launch_visc_mg<2>(r, iters);
    } else {
        for (uint32_t k = 0; k < iters; k++) {
            const int par = (int)(k & 1);
            hipLaunchKernelGGL(k_visc_iter, g0, TILE, 0, r.stream, r.tb, par);
            launch_visc_reduce(r, VRED_ALPHA);
            hipLaunchKernelGGL(k_visc_update, g0, TILE, 0, r.stream, r.tb, par);
            launch_visc_reduce(r, VRED_BETA);
        }
    }
    hipLaunchKernelGGL(k_visc_end, per_tank(r), dim3(64), 0, r.stream, r.tb);
}

bool twice_projected(const sp_gparams& gp) { return gp.visc_mode == 2 && gp.v_twice; }

// The viscosity before the first pressure solve (record_substep, mod.rs:2681-2686).
void launch_visc_before(const Run& r, const Recs& recs) {
    const dim3 g0 = grid0(r);
    if (r.gp.visc_mode == 1) {
        hipLaunchKernelGGL(k_visc_setup, g0, TILE, 0, r.stream, r.tb, 0);
        hipLaunchKernelGGL(k_visc_explicit, g0, TILE, 0, r.stream, r.tb);
    } else if (r.gp.visc_mode == 2) {
        launch_visc_split(r);
        if (!r.gp.v_twice) launch_visc_solve(r, recs.v, false);
    }
}

// After the first projection: the viscosity solve (mod.rs:2690-2691) ...
void launch_visc_after(const Run& r, const Recs& recs) {
    if (twice_projected(r.gp)) launch_visc_solve(r, recs.v, true);
}

// The rigid bodies' five passes after the projections (mod.rs:2701-2707).
void launch_rigid(const Run& r) {
    if (!(r.gp.caps & SP_CAP_RIGID)) return;
    for (int mode = 0; mode < 5; mode++) hipLaunchKernelGGL(k_rigid, grid0(r), TILE, 0, r.stream, r.tb, mode);
}

// ... and the second pressure solve with its projection (mod.rs:2692-2697), then the rigid
// bodies.
void launch_second_projection(const Run& r, const Recs& recs) {
    if (twice_projected(r.gp)) {
        const dim3 g0 = grid0(r);
        hipLaunchKernelGGL(k_rhs_visc, g0, TILE, 0, r.stream, r.tb);
        launch_pressure(r, recs.p2, true);
        hipLaunchKernelGGL(k_project_visc, g0, TILE, 0, r.stream, r.tb);
    }
    launch_rigid(r);
}

// ---- the other passes ------------------------------------------------------------------

void launch_grid_passes(const Run& r, const Recs& recs) {
    const dim3 g0 = grid0(r);
    hipLaunchKernelGGL(k_grid_prep, g0, TILE, 0, r.stream, r.tb);
    // non-Newtonian viscosity (record_substep, mod.rs:2639-2642)
    if (r.gp.caps & SP_CAP_RHEO) hipLaunchKernelGGL(k_rheo, g0, TILE, 0, r.stream, r.tb);
    // the viscoelastic stress's force on the faces (mod.rs:2643-2646)
    if (r.gp.caps & SP_CAP_ELASTIC) hipLaunchKernelGGL(k_elastic, g0, TILE, 0, r.stream, r.tb);
    // settling of suspensions (mod.rs:2647-2650)
    if (r.gp.caps & SP_CAP_SETTLE) hipLaunchKernelGGL(k_settle, g0, TILE, 0, r.stream, r.tb);
    hipLaunchKernelGGL(k_coef_smooth, g0, TILE, 0, r.stream, r.tb);
    hipLaunchKernelGGL(k_surf_curv, g0, TILE, 0, r.stream, r.tb);
    hipLaunchKernelGGL(k_surf_apply, g0, TILE, 0, r.stream, r.tb);
    // the magnetic energy, then its force on the faces (mod.rs:2655-2662)
    if (r.gp.caps & SP_CAP_MAGNET)
        for (int mode = 0; mode < 2; mode++) hipLaunchKernelGGL(k_magnet, g0, TILE, 0, r.stream, r.tb, mode);
    // heat conduction while the tank is thermal (record_substep, mod.rs:2663-2666)
    if ((r.gp.caps & SP_CAP_THERMAL) && r.gp.heat_rate > 0.0f) hipLaunchKernelGGL(k_heat, g0, TILE, 0, r.stream, r.tb);
    launch_visc_before(r, recs);
    hipLaunchKernelGGL(k_rhs, g0, TILE, 0, r.stream, r.tb);
}

void launch_project(const Run& r) { hipLaunchKernelGGL(k_project, grid0(r), TILE, 0, r.stream, r.tb); }

void launch_clear_acc(const Run& r) {
    // a grid-stride fill of each tank's W H stride words; enough groups for the widest stride
    const size_t words = (size_t)r.W * r.H * SP_ACC_MAX;
    const int blocks = (int)std::min<size_t>((words / 4 + 255) / 256, 2048);
    hipLaunchKernelGGL(k_clear_acc, dim3(std::max(blocks, 1), 1, r.T), dim3(256), 0, r.stream, r.tb);
}

}  // namespace

// ---- the particle pass's launch shape, tuned on the launches themselves -----------------
//
// Candidates: work-items per group and the capacity of the LDS scatter box. Every shape
// gives bit-identical results, so each launch of a key may use any of them: while a key is
// tuned, its launches take the candidates in turn (launch by launch, since timings arrive
// only after the launches' stream is synchronised), the timed ones (Timing's events around
// the particle pass) report their durations, and once every candidate has `samples` timings
// the fastest median is kept for the key. Shape 0 is the untuned default.
struct PShape {
    int threads, box, min_groups;
};
constexpr int N_PSHAPES = 9;
constexpr PShape P_SHAPES[N_PSHAPES] = {{256, 12288, 1}, {256, 8192, 1}, {256, 4096, 1}, {128, 2048, 1}, {128, 4096, 1},
                                        {512, 12288, 1}, {512, 8192, 1}, {256, 4096, 5},  {256, 6144, 1}};
using PKernel = void (*)(TankBatch, int);
// [exotic][shape]: the stress update and its scatter, and settling, only in the kernels of
// tanks that are elastic or settle
#define SP_PK(X)                                                                                                      \
    {k_particles<256, 12288, 1, X>, k_particles<256, 8192, 1, X>, k_particles<256, 4096, 1, X>,                        \
     k_particles<128, 2048, 1, X>,  k_particles<128, 4096, 1, X>, k_particles<512, 12288, 1, X>,                       \
     k_particles<512, 8192, 1, X>,  k_particles<256, 4096, 5, X>, k_particles<256, 6144, 1, X>}
const PKernel P_KERNELS[2][N_PSHAPES] = {SP_PK(false), SP_PK(true)};
#undef SP_PK

bool exotic_run(const Run& r) {
    return (r.gp.caps & (SP_CAP_ELASTIC | SP_CAP_SETTLE | SP_CAP_VOLATILE | SP_CAP_SELF_HEAT | SP_CAP_REACT |
                         SP_CAP_CURRENTS | SP_CAP_BOILS | SP_CAP_PUFFS | SP_CAP_STEAM)) != 0u ||
           r.gp.tool_mode != 0;
}

// What a shape is tuned for: the grid, the tanks per launch, the accumulator stride and
// whether the tanks run the elastic or settling code.
uint64_t particle_key(const Run& r) {
    return (uint64_t)(uint32_t)r.W | (uint64_t)(uint32_t)r.H << 14 | (uint64_t)(uint32_t)r.T << 28 |
           (uint64_t)(uint32_t)r.gp.acc_stride << 34 | (uint64_t)exotic_run(r) << 41;
}

class Tuner {
public:
    struct Report {
        uint64_t key;
        int best;
        float best_ms, default_ms;
        uint32_t samples;
    };

    // The shape of the next launch of `key`.
    int pick(uint64_t key) {
        std::lock_guard<std::mutex> g(mu_);
        if (forced_ >= 0) return forced_;
        if (samples_ == 0) return 0;
        Entry& e = map_[key];
        if (e.best >= 0) return e.best;
        return (int)(e.turn++ % (uint32_t)N_PSHAPES);
    }
    // Whether launches of `key` still gather timings.
    bool sampling(uint64_t key) {
        std::lock_guard<std::mutex> g(mu_);
        if (forced_ >= 0 || samples_ == 0) return false;
        auto it = map_.find(key);
        return it == map_.end() || it->second.best < 0;
    }
    void add(uint64_t key, int v, float ms) {
        std::lock_guard<std::mutex> g(mu_);
        if (forced_ >= 0 || samples_ == 0 || v < 0 || v >= N_PSHAPES) return;
        Entry& e = map_[key];
        if (e.best >= 0) return;
        e.ms[v].push_back(ms);
        for (int k = 0; k < N_PSHAPES; k++)
            if (e.ms[k].size() < samples_) return;
        float best = 0.0f;
        for (int k = 0; k < N_PSHAPES; k++) {
            const float m = median(e.ms[k]);
            if (e.best < 0 || m < best) {
                e.best = k;
                best = m;
            }
        }
    }
    // mode -1: tune; -2: forget what was tuned, then tune; 0 .. N_PSHAPES-1: that shape
    // everywhere. samples > 0: timings per candidate before choosing.
    bool set(int mode, uint32_t samples) {
        if (mode < -2 || mode >= N_PSHAPES) return false;
        std::lock_guard<std::mutex> g(mu_);
        if (mode == -2) map_.clear();
        forced_ = mode >= 0 ? mode : -1;
        if (samples > 0) samples_ = samples;
        return true;
    }
    std::vector<Report> report() {
        std::lock_guard<std::mutex> g(mu_);
        std::vector<Report> out;
        for (auto& kv : map_) {
            const Entry& e = kv.second;
            uint32_t n = 0;
            for (int k = 0; k < N_PSHAPES; k++) n += (uint32_t)e.ms[k].size();
            out.push_back({kv.first, e.best, e.best >= 0 ? median(e.ms[e.best]) : 0.0f,
                           e.ms[0].empty() ? 0.0f : median(e.ms[0]), n});
        }
        return out;
    }

private:
    struct Entry {
        int best = -1;
        uint32_t turn = 0;  // launches picked while tuning
        std::vector<float> ms[N_PSHAPES];
    };
    static float median(std::vector<float> v) {
        if (v.empty()) return 0.0f;
        std::sort(v.begin(), v.end());
        return v[v.size() / 2];
    }
    std::mutex mu_;
    int forced_ = -1;
    uint32_t samples_ = 3;
    std::map<uint64_t, Entry> map_;
};

Tuner& tuner() {
    static Tuner t;
    return t;
}

namespace {

// Launches the particle pass in the shape the tuner picks; returns the shape.
int launch_particles(const Run& r, int gather) {
    const int v = tuner().pick(particle_key(r));
    const PShape sh = P_SHAPES[v];
    const uint32_t blocks = std::max<uint32_t>((r.slots + (uint32_t)sh.threads - 1) / (uint32_t)sh.threads, 1);
    hipLaunchKernelGGL(P_KERNELS[exotic_run(r) ? 1 : 0][v], dim3(blocks, 1, r.T), dim3((unsigned)sh.threads), 0, r.stream,
                       r.tb, gather);
    return v;
}

// What the host decides about a substep's launches from the tables (the parameters the
// kernels read themselves stay out of the key).
uint32_t schedule_key(const sp_gparams& g) {
    return (uint32_t)g.visc_mode | (uint32_t)g.v_twice << 2 | (uint32_t)g.v_precond << 3 | (uint32_t)g.v_warm << 4 |
           (uint32_t)g.v_mg_levels << 5 | (uint32_t)g.v_mg_sweeps << 10 | (uint32_t)g.mg_half << 12 |
           (uint32_t)((g.caps & SP_CAP_THERMAL) && g.heat_rate > 0.0f) << 13 | (uint32_t)((g.caps & SP_CAP_RIGID) != 0u) << 14 |
           (uint32_t)((g.caps & SP_CAP_RHEO) != 0u) << 15 | (uint32_t)((g.caps & SP_CAP_ELASTIC) != 0u) << 16 |
           (uint32_t)((g.caps & SP_CAP_SETTLE) != 0u) << 17 | (uint32_t)((g.caps & SP_CAP_MAGNET) != 0u) << 18 |
           (uint32_t)((g.caps & SP_CAP_VOLATILE) != 0u) << 19 | (uint32_t)((g.caps & SP_CAP_SELF_HEAT) != 0u) << 20 |
           (uint32_t)((g.caps & SP_CAP_REACT) != 0u) << 21 | (uint32_t)((g.caps & SP_CAP_CURRENTS) != 0u) << 22;
}

}  // namespace

// ---- graphs and timing ------------------------------------------------------------------

// Instantiated graphs of the parts of a substep: grid passes (with the viscosity before the
// first solve), pressure solve, projection, the viscosity solve after it, and the second
// pressure solve with its projection. Each part is keyed on the schedule and on the one
// recorded iteration count it runs (none for the projection), so a new count recaptures only
// the part that runs it, and a part holds up to GRAPHS_PER_PART graphs, the least recently
// used one making room for a new one.
constexpr int GRAPH_PARTS = 5;
constexpr size_t GRAPHS_PER_PART = 32;
struct GraphEntry {
    uint32_t sched = 0, count = 0;
    uint64_t used = 0;
    hipGraphExec_t exec = nullptr;
};

struct GraphSet {
    std::vector<GraphEntry> part[GRAPH_PARTS];
    uint64_t tick = 0;
    uint64_t captures = 0;
    bool on = true;

    void clear() {
        for (auto& v : part) {
            for (auto& e : v)
                if (e.exec) (void)hipGraphExecDestroy(e.exec);
            v.clear();
        }
    }
    ~GraphSet() { clear(); }
};

// Device time per substep of the phases of a substep, from events at their boundaries:
// [0] total, [1] step + accumulator clear + particle pass, [2] grid passes (with the
// viscosity before the first solve), [3] pressure solve, [4] projection, [5] viscosity
// solve after it, [6] second pressure solve and projection; and the particle kernel alone.
struct Timing {
    static constexpr int SUBSTEPS = 64;
    static constexpr int EV_PER = 9;
    bool on = false;
    std::vector<hipEvent_t> ev;
    uint64_t pkey[SUBSTEPS] = {};  // the particle pass's tuning key and shape per timed substep
    int pshape[SUBSTEPS] = {};
    uint32_t timed = 0, timed_last = 0;
    float phase[7] = {};
    float particle_ms = 0.0f;

    void create() {
        ev.assign((size_t)SUBSTEPS * EV_PER, nullptr);
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        on = true;
    }
    ~Timing() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    hipEvent_t* next() { return on && timed < (uint32_t)SUBSTEPS ? &ev[(size_t)timed++ * EV_PER] : nullptr; }
    void note_shape(uint64_t key, int shape) {
        pkey[timed - 1] = key;
        pshape[timed - 1] = shape;
    }
    // Averages over the substeps recorded since the last call (after the stream finished).
    void collect() {
        if (!on || timed == 0) return;
        double acc[7] = {0, 0, 0, 0, 0, 0, 0}, kp = 0;
        const int from[6] = {0, 1, 2, 3, 4, 7}, to[6] = {1, 2, 3, 4, 7, 8};
        for (uint32_t k = 0; k < timed; k++) {
            hipEvent_t* e = &ev[(size_t)k * EV_PER];
            float t;
            for (int q = 0; q < 6; q++) {
                HIPCHK(hipEventElapsedTime(&t, e[from[q]], e[to[q]]));
                acc[q + 1] += t;
                acc[0] += t;
            }
            HIPCHK(hipEventElapsedTime(&t, e[5], e[6]));
            kp += t;
            tuner().add(pkey[k], pshape[k], t);
        }
        for (int q = 0; q < 7; q++) phase[q] = (float)(acc[q] / timed);
        particle_ms = (float)(kp / timed);
        timed_last = timed;
        timed = 0;
    }
};

namespace {

hipGraphExec_t capture(const Run& r, int part, const Recs& recs) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(r.stream, hipStreamCaptureModeThreadLocal));
    if (part == 0) launch_grid_passes(r, recs);
    else if (part == 1) launch_pressure(r, recs.p);
    else if (part == 2) launch_project(r);
    else if (part == 3) launch_visc_after(r, recs);
    else launch_second_projection(r, recs);
    hipError_t le = hipGetLastError();
    HIPCHK(hipStreamEndCapture(r.stream, &g));
    HIPCHK(le);
    hipGraphExec_t e = nullptr;
    HIPCHK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return e;
}

// The recorded iteration count part `part` of a substep runs (0: none).
uint32_t part_count(const Run& r, int part, const Recs& recs) {
    switch (part) {
    case 0: return r.gp.visc_mode == 2 && !r.gp.v_twice ? recs.v : 0u;
    case 1: return recs.p;
    case 3: return twice_projected(r.gp) ? recs.v : 0u;
    case 4: return twice_projected(r.gp) ? recs.p2 : 0u;
    default: return 0u;
    }
}

hipGraphExec_t graph_for(GraphSet& gs, const Run& r, int part, const Recs& recs) {
    const uint32_t key = schedule_key(r.gp), n = part_count(r, part, recs);
    std::vector<GraphEntry>& v = gs.part[part];
    for (auto& e : v)
        if (e.sched == key && e.count == n) {
            e.used = ++gs.tick;
            return e.exec;
        }
    if (v.size() >= GRAPHS_PER_PART) {
        auto lru = std::min_element(v.begin(), v.end(), [](const GraphEntry& a, const GraphEntry& b) { return a.used < b.used; });
        (void)hipGraphExecDestroy(lru->exec);
        v.erase(lru);
    }
    GraphEntry e;
    e.sched = key;
    e.count = n;
    e.used = ++gs.tick;
    e.exec = capture(r, part, recs);
    gs.captures++;
    v.push_back(e);
    return e.exec;
}

// One whole substep of the group (record_substep, mod.rs:2615-2708). gather = 0 on the
// first substep after a reset (no grid velocities to gather yet).
void launch_substep(const Run& r, GraphSet* gs, Timing* tm, int gather, const Recs& recs) {
    hipEvent_t* e = tm ? tm->next() : nullptr;
    hipGraphExec_t gx[GRAPH_PARTS] = {};
    const bool g = gs && gs->on;
    if (g)
        for (int k = 0; k < GRAPH_PARTS; k++) gx[k] = graph_for(*gs, r, k, recs);
    if (e) HIPCHK(hipEventRecord(e[0], r.stream));
    hipLaunchKernelGGL(k_step, per_tank(r), dim3(64), 0, r.stream, r.tb);
    launch_clear_acc(r);
    if (e) HIPCHK(hipEventRecord(e[5], r.stream));
    const int shape = launch_particles(r, gather);
    if (e) {
        tm->note_shape(particle_key(r), shape);
        HIPCHK(hipEventRecord(e[6], r.stream));
    }
    if (e) HIPCHK(hipEventRecord(e[1], r.stream));
    if (g) HIPCHK(hipGraphLaunch(gx[0], r.stream));
    else launch_grid_passes(r, recs);
    if (e) HIPCHK(hipEventRecord(e[2], r.stream));
    if (g) HIPCHK(hipGraphLaunch(gx[1], r.stream));
    else launch_pressure(r, recs.p);
    if (e) HIPCHK(hipEventRecord(e[3], r.stream));
    if (g) HIPCHK(hipGraphLaunch(gx[2], r.stream));
    else launch_project(r);
    if (e) HIPCHK(hipEventRecord(e[4], r.stream));
    if (g) HIPCHK(hipGraphLaunch(gx[3], r.stream));
    else launch_visc_after(r, recs);
    if (e) HIPCHK(hipEventRecord(e[7], r.stream));
    if (g) HIPCHK(hipGraphLaunch(gx[4], r.stream));
    else launch_second_projection(r, recs);
    if (e) HIPCHK(hipEventRecord(e[8], r.stream));
    HIPCHK(hipGetLastError());
}

}  // namespace
