// sprout.hip — libsprout_hip.so: the device state of one tank, the substep schedule on a
// HIP stream, and the C ABI of sprout_hip.h (the boundary that replaces FluidSim's
// Vulkano dispatch, src/sim/mod.rs:862-3492).
//
// One translation unit: the kernel files are included here so that the kernels and the
// launch code see each other's inline helpers.
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <string>
#include <vector>

#include "grid.hip"
#include "host.h"
#include "particles.hip"
#include "solve.hip"
#include "visc.hip"

namespace {

thread_local std::string g_create_error;

struct HipError {
    hipError_t e;
    const char* what;
};

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) throw HipError{e_, #x}; \
    } while (0)

template <class T>
T* dalloc(size_t n, std::vector<void*>& owned) {
    void* p = nullptr;
    if (n == 0) n = 1;
    HIPCHK(hipMalloc(&p, n * sizeof(T)));
    owned.push_back(p);
    HIPCHK(hipMemset(p, 0, n * sizeof(T)));
    return (T*)p;
}

}  // namespace

// Recorded iterations of one substep: first pressure solve, second pressure solve (after
// the viscosity), viscosity solve (record_substep, mod.rs:2615-2623).
struct Recs {
    uint32_t p = 0, p2 = 0, v = 0;
    bool operator==(const Recs& o) const { return p == o.p && p2 == o.p2 && v == o.v; }
};

// Instantiated graphs of one (recorded iterations, schedule) key: grid passes (with the
// viscosity before the first solve), pressure solve, projection, the viscosity solve after
// it, and the second pressure solve with its projection.
constexpr int GRAPH_PARTS = 5;
struct Graphs {
    Recs recs;
    uint32_t sched = 0xffffffffu;
    hipGraphExec_t exec[GRAPH_PARTS] = {};
};

struct sp_sim {
    int W = 0, H = 0, N = 0;
    uint32_t cap = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> owned;
    std::string err;

    sp_sim_params params{};
    bool params_set = false;
    sp::Tank tank;
    sp_gparams gp{};
    sp_gmat gm[SP_MAX_MATERIALS]{};
    bool tables_dirty = true;
    sp_levels_info lv{};
    int nb0 = 0;  // level-0 work-groups (16 x 16 tiles)
    size_t tail_lds_bytes = 0;  // dynamic LDS of k_mg_tail

    // device
    Ctx c{};
    sp_gparams* d_P = nullptr;
    sp_gmat* d_mats = nullptr;
    sp_particle* d_part_tmp = nullptr;
    sp_particle* d_spawn = nullptr;
    uint32_t* d_counter = nullptr;  // stats / claim counters
    uint32_t* d_count_out = nullptr;
    uint32_t* d_keys = nullptr;
    uint32_t* d_vals = nullptr;
    uint32_t* d_keys2 = nullptr;
    uint32_t* d_vals2 = nullptr;
    float2* d_r2 = nullptr;  // second level-0 residual buffer of the CG
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;

    // host state mirroring FluidSim
    uint32_t n_slots = 0;
    uint32_t next_id = 0;
    bool grid_valid = false;
    uint32_t since_sort = 0;
    uint32_t sort_interval = 8;
    std::vector<sp_particle> pending_spawn;
    double time_base = 0.0;
    double sim_time = 0.0;
    float dt_estimate = 1.0f / 240.0f;
    Recs rec{60, 60, 200};
    sp_solver_stats stats{};
    uint64_t substeps_base = 0;
    uint32_t last_unconv = 0, last_p2_unconv = 0, last_v_unconv = 0;
    uint32_t last_v_iters = 0, last_p2_iters = 0;  // of the latest solves

    // timing: per substep of a batch (up to TIMED_SUBSTEPS), events at the phase
    // boundaries and tightly around the particle kernel
    static constexpr int TIMED_SUBSTEPS = 64;
    static constexpr int EV_PER = 9;
    bool timing = false;
    std::vector<hipEvent_t> ev;
    uint32_t timed = 0;  // substeps of the current batch with events recorded
    uint32_t timed_last = 0;
    float timings[10] = {};
    float particle_ms = 0.0f;
    bool use_graphs = true;
    std::vector<Graphs> graphs;

    Ctx ctx() const { return c; }
};

namespace {

void set_err(sp_sim* s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    s->err = buf;
}

dim3 grid_groups(int w, int h) { return dim3((w + 15) / 16, (h + 15) / 16, 1); }
const dim3 TILE(16, 16, 1);

// ---- the substep schedule (record_substep, mod.rs:2615-2708) ----------------------------

// One V-cycle on the level-0 residual r0 (record_vcycle, mod.rs:2930-2954); with
// `skip_down0` the way down on level 0 has already run (fused into the CG update).
void launch_vcycle(sp_sim* s, int gated, int reduce, const float2* r0, bool skip_down0 = false) {
    const Ctx& c = s->c;
    const sp_levels_info& L = s->lv;
    for (int l = skip_down0 ? 1 : 0; l < L.tail; l++)
        hipLaunchKernelGGL(k_mg_down, grid_groups(L.dims[l][0], L.dims[l][1]), TILE, 0, s->stream, c, l, gated, r0);
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(1024), s->tail_lds_bytes, s->stream, c, gated);
    for (int l = L.tail - 1; l >= 0; l--)
        hipLaunchKernelGGL(k_mg_up, grid_groups(L.dims[l][0], L.dims[l][1]), TILE, 0, s->stream, c, l, gated,
                           l == 0 ? reduce : -1, s->nb0, r0);
    if (reduce >= 0) hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(256), 0, s->stream, c, reduce, s->nb0, gated);
}

void launch_mg_build(sp_sim* s) {
    const Ctx& c = s->c;
    const sp_levels_info& L = s->lv;
    for (int l = 1; l <= L.tail; l++)
        hipLaunchKernelGGL(k_mg_build, grid_groups(L.dims[l][0], L.dims[l][1]), TILE, 0, s->stream, c, l);
    if (L.tail < L.n - 1) hipLaunchKernelGGL(k_mg_tail_build, dim3(1), dim3(1024), 0, s->stream, c);
}

// record_pressure + record_cg (mod.rs:2956-3010). The first solve builds the hierarchy and
// rescales the warm start; the second (after the viscosity, `second`) reuses both and
// starts from what rhs_visc.comp left. The level-0 residual alternates between mgr and r2
// from one iteration to the next (k_update_down0).
void launch_pressure(sp_sim* s, uint32_t iters, bool second = false) {
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    if (!second) {
        launch_mg_build(s);
        hipLaunchKernelGGL(k_cg_scale, g0, TILE, 0, s->stream, c);
    }
    hipLaunchKernelGGL(k_cg_init, g0, TILE, 0, s->stream, c, s->nb0);
    float2* rb[2] = {c.mgr, s->d_r2};
    launch_vcycle(s, 0, second ? (int)RED_START2 : (int)RED_START, rb[0]);
    float2* bufs[2] = {c.cgp, c.cgp2};
    for (uint32_t k = 0; k < iters; k++) {
        const float2* p_old = bufs[k & 1];
        float2* p_new = bufs[(k + 1) & 1];
        hipLaunchKernelGGL(k_cg_apply, g0, TILE, 0, s->stream, c, s->nb0, p_old, p_new);
        hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(256), 0, s->stream, c, 1, s->nb0, 1);
        hipLaunchKernelGGL(k_update_down0, g0, TILE, 0, s->stream, c, (const float2*)rb[k & 1], rb[(k + 1) & 1],
                           (const float2*)p_new);
        launch_vcycle(s, 1, 2, rb[(k + 1) & 1], true);
    }
    hipLaunchKernelGGL(k_cg_end, dim3(1), dim3(64), 0, s->stream, c, second ? 1 : 0);
}

// ---- the implicit viscosity (record_viscosity_split / _solve, mod.rs:2812-2927) ----------

// Weights and active tiles, the explicit step outside them, the Jacobi diagonal, the warm
// start, and the multigrid's matrices.
void launch_visc_split(sp_sim* s) {
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    hipLaunchKernelGGL(k_fill_u32, dim3(std::max((s->nb0 / 4 + 255) / 256, 1)), dim3(256), 0, s->stream, c.vtiles,
                       (size_t)s->nb0, 0u);
    hipLaunchKernelGGL(k_visc_setup, g0, TILE, 0, s->stream, c, 1);
    hipLaunchKernelGGL(k_visc_split, g0, TILE, 0, s->stream, c);
    hipLaunchKernelGGL(k_visc_diag, g0, TILE, 0, s->stream, c);
    if (s->gp.v_precond)
        for (int l = 1; l <= s->gp.v_mg_levels; l++)
            hipLaunchKernelGGL(k_vmg_build, grid_groups(s->lv.dims[l][0], s->lv.dims[l][1]), TILE, 0, s->stream, c, l);
}

void launch_visc_reduce(sp_sim* s, int mode) {
    hipLaunchKernelGGL(k_visc_reduce, dim3(1), dim3(256), 0, s->stream, s->c, mode, s->nb0);
}

// The coarse levels of a viscosity V-cycle, down and up (record_visc_coarse, mod.rs:2838-2853).
template <int NU>
void launch_visc_coarse(sp_sim* s) {
    const Ctx& c = s->c;
    const int n = s->gp.v_mg_levels;
    for (int l = 1; l <= n; l++)
        hipLaunchKernelGGL(k_vmg_down<NU>, grid_groups(s->lv.dims[l][0], s->lv.dims[l][1]), TILE, 0, s->stream, c, l,
                           l == n ? 1 : 0);
    for (int l = n; l >= 1; l--)
        hipLaunchKernelGGL(k_vmg_up<NU>, grid_groups(s->lv.dims[l][0], s->lv.dims[l][1]), TILE, 0, s->stream, c, l,
                           l == n ? 1 : 0);
}

// Multigrid-preconditioned CG: z of the starting residual and the first step length, then
// per iteration the update with the way down on level 0, the coarse levels and the way up
// with A z and both step lengths. Residual and A p alternate between two buffers.
template <int NU>
void launch_visc_mg(sp_sim* s, uint32_t iters) {
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    hipLaunchKernelGGL(k_vmg_down<NU>, g0, TILE, 0, s->stream, c, 0, 0);
    launch_visc_coarse<NU>(s);
    hipLaunchKernelGGL(k_visc_mgz<NU>, g0, TILE, 0, s->stream, c, (const float2*)c.vr, (const float2*)c.vq);
    launch_visc_reduce(s, VRED_MG_START);
    float2* rb[2] = {c.vr, c.vr2};
    float2* qb[2] = {c.vq, c.vq2};
    for (uint32_t k = 0; k < iters; k++) {
        const int a = k & 1, b = a ^ 1;
        hipLaunchKernelGGL(k_visc_mgu<NU>, g0, TILE, 0, s->stream, c, (const float2*)rb[a], (const float2*)qb[a], rb[b], qb[b]);
        launch_visc_reduce(s, VRED_MG_UPDATE);
        launch_visc_coarse<NU>(s);
        hipLaunchKernelGGL(k_visc_mgz<NU>, g0, TILE, 0, s->stream, c, (const float2*)rb[b], (const float2*)qb[b]);
        launch_visc_reduce(s, VRED_MG_ITER);
    }
}

// The implicit solve on the active tiles. `restrain`: the first pressure solve changed the
// velocity since the split, so the strain of u* is taken again.
void launch_visc_solve(sp_sim* s, uint32_t iters, bool restrain) {
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    if (restrain) hipLaunchKernelGGL(k_visc_strain, g0, TILE, 0, s->stream, c);
    hipLaunchKernelGGL(k_visc_start, g0, TILE, 0, s->stream, c);
    launch_visc_reduce(s, VRED_START);
    if (s->gp.v_warm) {
        hipLaunchKernelGGL(k_visc_warm, g0, TILE, 0, s->stream, c);
        launch_visc_reduce(s, VRED_WARM);
    }
    if (s->gp.v_precond) {
        if (s->gp.v_mg_sweeps == 1) launch_visc_mg<1>(s, iters);
        else launch_visc_mg<2>(s, iters);
    } else {
        float2* pb[2] = {c.vp, c.vp2};
        for (uint32_t k = 0; k < iters; k++) {
            const int a = k & 1, b = a ^ 1;
            hipLaunchKernelGGL(k_visc_iter, g0, TILE, 0, s->stream, c, (const float2*)pb[a], pb[b]);
            launch_visc_reduce(s, VRED_ALPHA);
            hipLaunchKernelGGL(k_visc_update, g0, TILE, 0, s->stream, c, (const float2*)pb[b]);
            launch_visc_reduce(s, VRED_BETA);
        }
    }
    hipLaunchKernelGGL(k_visc_end, dim3(1), dim3(64), 0, s->stream, c);
}

// The viscosity before the first pressure solve (record_substep, mod.rs:2681-2686).
void launch_visc_before(sp_sim* s, const Recs& r) {
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    if (s->gp.visc_mode == 1) {
        hipLaunchKernelGGL(k_visc_setup, g0, TILE, 0, s->stream, c, 0);
        hipLaunchKernelGGL(k_visc_explicit, g0, TILE, 0, s->stream, c);
    } else if (s->gp.visc_mode == 2) {
        launch_visc_split(s);
        if (!s->gp.v_twice) launch_visc_solve(s, r.v, false);
    }
}

bool twice_projected(const sp_sim* s) { return s->gp.visc_mode == 2 && s->gp.v_twice; }

// After the first projection: the viscosity solve (mod.rs:2690-2691) ...
void launch_visc_after(sp_sim* s, const Recs& r) {
    if (twice_projected(s)) launch_visc_solve(s, r.v, true);
}

// ... and the second pressure solve with its projection (mod.rs:2692-2697).
void launch_second_projection(sp_sim* s, const Recs& r) {
    if (!twice_projected(s)) return;
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    hipLaunchKernelGGL(k_rhs_visc, g0, TILE, 0, s->stream, c);
    launch_pressure(s, r.p2, true);
    hipLaunchKernelGGL(k_project_visc, g0, TILE, 0, s->stream, c);
}

void launch_grid_passes(sp_sim* s, const Recs& r) {
    const Ctx& c = s->c;
    dim3 g0 = grid_groups(s->W, s->H);
    hipLaunchKernelGGL(k_grid_prep, g0, TILE, 0, s->stream, c);
    hipLaunchKernelGGL(k_coef_smooth, g0, TILE, 0, s->stream, c);
    hipLaunchKernelGGL(k_surf_curv, g0, TILE, 0, s->stream, c);
    hipLaunchKernelGGL(k_surf_apply, g0, TILE, 0, s->stream, c);
    launch_visc_before(s, r);
    hipLaunchKernelGGL(k_rhs, g0, TILE, 0, s->stream, c);
}

void launch_clear_acc(sp_sim* s) {
    size_t n = (size_t)s->N * s->gp.acc_stride;
    int blocks = (int)std::min<size_t>((n / 4 + 255) / 256 + 1, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s->stream, (uint32_t*)s->c.acc, n, 0u);
}

void launch_particles(sp_sim* s, int gather) {
    uint32_t blocks = std::max<uint32_t>((s->n_slots + 255) / 256, 1);
    hipLaunchKernelGGL(k_particles, dim3(blocks), dim3(256), 0, s->stream, s->c, gather);
}

void launch_sort(sp_sim* s, int tiled) {
    if (s->n_slots == 0) return;
    uint32_t blocks = (s->n_slots + 255) / 256;
    // keys below 2^end_bit; dead slots take the largest, so they collect at the end
    uint32_t live_max = tiled ? (uint32_t)(((s->W + 7) / 8) * ((s->H + 7) / 8) * 64) : (uint32_t)s->N;
    int end_bit = 1;
    while (end_bit < 32 && (1u << end_bit) <= live_max) end_bit++;
    uint32_t dead = end_bit >= 32 ? 0xffffffffu : (1u << end_bit) - 1u;
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, s->stream, s->c, s->d_keys, s->d_vals, tiled, dead);
    size_t bytes = s->sort_tmp_bytes;
    HIPCHK(hipcub::DeviceRadixSort::SortPairs(s->d_sort_tmp, bytes, s->d_keys, s->d_keys2, s->d_vals, s->d_vals2,
                                              (int)s->n_slots, 0, end_bit, s->stream));
    hipLaunchKernelGGL(k_sort_gather, dim3(blocks), dim3(256), 0, s->stream, s->c, (const uint32_t*)s->d_vals2,
                       s->d_part_tmp, s->n_slots);
    HIPCHK(hipMemcpyAsync(s->c.part, s->d_part_tmp, sizeof(sp_particle) * s->n_slots, hipMemcpyDeviceToDevice, s->stream));
}

// The grid passes, the pressure solve with `p_iters` recorded iterations, and the
// projection as three HIP graphs, captured once per (iterations, viscosity mode) and
// replayed: a substep is some two hundred dependent launches, which the host could not
// issue as fast as the device retires them. The kernels read everything that changes
// between substeps (dt, tables, solve state) from device memory, so a replay needs no
// new arguments. Events between the graphs time the phases.
void destroy_graphs(sp_sim* s) {
    for (auto& g : s->graphs)
        for (auto& e : g.exec)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    s->graphs.clear();
}

hipGraphExec_t capture(sp_sim* s, int part, const Recs& r) {
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    if (part == 0) launch_grid_passes(s, r);
    else if (part == 1) launch_pressure(s, r.p);
    else if (part == 2) hipLaunchKernelGGL(k_project, grid_groups(s->W, s->H), TILE, 0, s->stream, s->c);
    else if (part == 3) launch_visc_after(s, r);
    else launch_second_projection(s, r);
    hipError_t le = hipGetLastError();
    HIPCHK(hipStreamEndCapture(s->stream, &g));
    HIPCHK(le);
    hipGraphExec_t e = nullptr;
    HIPCHK(hipGraphInstantiate(&e, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    return e;
}

// What the host decides about a substep's launches from the tables (the parameters the
// kernels read themselves stay out of the key).
uint32_t schedule_key(const sp_sim* s) {
    const sp_gparams& g = s->gp;
    return (uint32_t)g.visc_mode | (uint32_t)g.v_twice << 2 | (uint32_t)g.v_precond << 3 | (uint32_t)g.v_warm << 4 |
           (uint32_t)g.v_mg_levels << 5 | (uint32_t)g.v_mg_sweeps << 10;
}

Graphs& graphs_for(sp_sim* s, const Recs& r) {
    const uint32_t key = schedule_key(s);
    for (auto& g : s->graphs)
        if (g.recs == r && g.sched == key) return g;
    if (s->graphs.size() >= 12) destroy_graphs(s);
    s->graphs.emplace_back();
    Graphs& g = s->graphs.back();
    g.recs = r;
    g.sched = key;
    for (int k = 0; k < GRAPH_PARTS; k++) g.exec[k] = capture(s, k, r);
    return g;
}

void launch_substep(sp_sim* s, int gather, const Recs& r) {
    const Ctx& c = s->c;
    hipEvent_t* e = nullptr;
    if (s->timing && s->timed < (uint32_t)sp_sim::TIMED_SUBSTEPS) e = &s->ev[(size_t)s->timed++ * sp_sim::EV_PER];
    Graphs* g = s->use_graphs ? &graphs_for(s, r) : nullptr;
    if (e) HIPCHK(hipEventRecord(e[0], s->stream));
    hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s->stream, c);
    launch_clear_acc(s);
    if (e) HIPCHK(hipEventRecord(e[5], s->stream));
    launch_particles(s, gather);
    if (e) HIPCHK(hipEventRecord(e[6], s->stream));
    if (e) HIPCHK(hipEventRecord(e[1], s->stream));
    if (g) HIPCHK(hipGraphLaunch(g->exec[0], s->stream));
    else launch_grid_passes(s, r);
    if (e) HIPCHK(hipEventRecord(e[2], s->stream));
    if (g) HIPCHK(hipGraphLaunch(g->exec[1], s->stream));
    else launch_pressure(s, r.p);
    if (e) HIPCHK(hipEventRecord(e[3], s->stream));
    if (g) HIPCHK(hipGraphLaunch(g->exec[2], s->stream));
    else hipLaunchKernelGGL(k_project, grid_groups(s->W, s->H), TILE, 0, s->stream, c);
    if (e) HIPCHK(hipEventRecord(e[4], s->stream));
    if (g) HIPCHK(hipGraphLaunch(g->exec[3], s->stream));
    else launch_visc_after(s, r);
    if (e) HIPCHK(hipEventRecord(e[7], s->stream));
    if (g) HIPCHK(hipGraphLaunch(g->exec[4], s->stream));
    else launch_second_projection(s, r);
    if (e) HIPCHK(hipEventRecord(e[8], s->stream));
    HIPCHK(hipGetLastError());
}

// Uploads the grid-unit tables if the parameters or the tank changed.
int upload_tables(sp_sim* s) {
    if (!s->tables_dirty) return SP_OK;
    std::string why;
    int rc = sp::derive(s->params, s->W, s->H, s->tank, s->n_slots, s->gp, s->gm, &why);
    if (rc != SP_OK) {
        set_err(s, "%s", why.c_str());
        return rc;
    }
    HIPCHK(hipMemcpyAsync(s->d_P, &s->gp, sizeof(sp_gparams), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->d_mats, s->gm, sizeof(s->gm), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->tables_dirty = false;
    return SP_OK;
}

// Reads the step and solve state back and folds it into the statistics
// (poll_readbacks, mod.rs:3033-3116).
void readback(sp_sim* s) {
    sp_stepbuf S;
    sp_solve K;
    HIPCHK(hipMemcpyAsync(&S, s->c.S, sizeof(S), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(&K, s->c.K, sizeof(K), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double t = s->time_base + (double)S.advanced - (double)__builtin_bit_cast(float, S.pad0);
    if (t > s->sim_time) s->sim_time = t;
    sp_solver_stats& st = s->stats;
    st.substeps = s->substeps_base + S.substeps;
    st.dt = S.dt;
    st.dt_cfl = S.dt_cfl;
    st.umax = S.umax_seen;
    st.umax_grid = __builtin_bit_cast(float, S.umax_grid);
    st.umax_part = __builtin_bit_cast(float, S.umax_part);
    st.pressure_iters_max = std::max(st.pressure_iters_max, std::max(K.it_max, K.p2_it_max));
    st.viscosity_iters_max = std::max(st.viscosity_iters_max, K.v_it_max);
    const uint32_t p_un = K.unconv - s->last_unconv, p2_un = K.p2_unconv - s->last_p2_unconv,
                   v_un = K.v_unconv - s->last_v_unconv;
    s->last_unconv = K.unconv;
    s->last_p2_unconv = K.p2_unconv;
    s->last_v_unconv = K.v_unconv;
    st.pressure_unconverged += p_un + p2_un;
    st.viscosity_unconverged += v_un;
    st.pressure_residual_max = std::max(st.pressure_residual_max, K.res_max);
    st.viscosity_residual_max = std::max(st.viscosity_residual_max, K.v_res_max);
    st.pressure_iters_last = K.iters_last;
    s->last_v_iters = K.v_iters_last;
    s->last_p2_iters = K.p2_iters_last;
    if (st.dt_cfl > 0.0f) s->dt_estimate = st.dt_cfl;
    // Adapt the recorded iterations (mod.rs:3100-3115): a little more than the batch needed,
    // half as much again when a solve ran out.
    auto adapt = [](uint32_t& rec, uint32_t used, uint32_t unconverged, uint32_t maxi, float margin) {
        if (unconverged > 0) rec = std::min(rec + rec / 2 + 2, maxi);
        else rec = std::min(std::max(used + (uint32_t)((float)used * margin) + 2, 2u), maxi);
    };
    adapt(s->rec.p, K.it_max, p_un, s->params.pressure_max_iters, 0.25f);
    if (s->gp.visc_mode == 2 && s->gp.v_twice) adapt(s->rec.p2, K.p2_it_max, p2_un, s->params.pressure_max_iters, 0.25f);
    else s->rec.p2 = s->rec.p;
    if (s->gp.visc_mode == 2) adapt(s->rec.v, K.v_it_max, v_un, s->params.viscosity_max_iters, 0.5f);
    // the per-batch maxima start again
    uint32_t zero = 0;
    HIPCHK(hipMemcpyAsync(&s->c.K->it_max, &zero, sizeof(zero), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(&s->c.K->p2_it_max, &zero, sizeof(zero), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(&s->c.K->v_it_max, &zero, sizeof(zero), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
}

int record(sp_sim* s, uint32_t substeps, float target) {
    if (!s->params_set) {
        set_err(s, "sp_set_params has not been called");
        return SP_E_INVALID;
    }
    // pending spawns: fresh slots first, then dead slots (mod.rs:2485-2503, claim.comp)
    uint32_t claim = 0;
    if (!s->pending_spawn.empty()) {
        uint32_t n = (uint32_t)s->pending_spawn.size();
        uint32_t fresh = std::min(n, s->cap - s->n_slots);
        if (fresh > 0) {
            HIPCHK(hipMemcpyAsync(s->c.part + s->n_slots, s->pending_spawn.data(), sizeof(sp_particle) * fresh,
                                  hipMemcpyHostToDevice, s->stream));
            s->n_slots += fresh;
            s->tables_dirty = true;
        }
        if (n > fresh) {
            claim = n - fresh;
            HIPCHK(hipMemcpyAsync(s->d_spawn, s->pending_spawn.data() + fresh, sizeof(sp_particle) * claim,
                                  hipMemcpyHostToDevice, s->stream));
        }
        HIPCHK(hipStreamSynchronize(s->stream));
        s->pending_spawn.clear();
    }
    int rc = upload_tables(s);
    if (rc != SP_OK) return rc;
    if (claim > 0) {
        HIPCHK(hipMemsetAsync(s->d_counter, 0, sizeof(uint32_t) * 4, s->stream));
        hipLaunchKernelGGL(k_claim, dim3((s->n_slots + 255) / 256), dim3(256), 0, s->stream, s->c, (const sp_particle*)s->d_spawn,
                           claim, s->d_counter);
    }
    if (substeps > 0) {
        // StepBuf words 0, 2, 3: target time, recorded substeps, done (record, mod.rs:2521-2526)
        uint32_t words[4];
        float tgt = std::max(target, 0.0f);
        memcpy(&words[0], &tgt, 4);
        words[2] = substeps;
        words[3] = 0;
        HIPCHK(hipMemcpyAsync(&s->c.S->frame_left, &words[0], 4, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(&s->c.S->batch_steps, &words[2], 8, hipMemcpyHostToDevice, s->stream));
    }
    for (uint32_t k = 0; k < substeps; k++) {
        if (s->since_sort >= s->sort_interval) {
            launch_sort(s, 1);
            s->since_sort = 0;
        }
        s->since_sort++;
        launch_substep(s, s->grid_valid ? 1 : 0, s->rec);
        s->grid_valid = true;
    }
    return SP_OK;
}

int wait(sp_sim* s) {
    HIPCHK(hipStreamSynchronize(s->stream));
    if (s->timing && s->timed > 0) {
        // averages per substep: [0] total, [1] step + clear + particle pass, [2] grid passes
        // (with the viscosity before the first solve), [3] pressure solve, [4] projection,
        // [5] viscosity solve after it, [6] second pressure solve and projection; [7]-[9]
        // below are iteration counts
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (uint32_t k = 0; k < s->timed; k++) {
            hipEvent_t* e = &s->ev[(size_t)k * sp_sim::EV_PER];
            const int from[6] = {0, 1, 2, 3, 4, 7}, to[6] = {1, 2, 3, 4, 7, 8};
            float t;
            for (int q = 0; q < 6; q++) {
                HIPCHK(hipEventElapsedTime(&t, e[from[q]], e[to[q]]));
                acc[q + 1] += t;
                acc[0] += t;
            }
        }
        for (int q = 0; q < 7; q++) s->timings[q] = (float)(acc[q] / s->timed);
        double kp = 0;
        for (uint32_t k = 0; k < s->timed; k++) {
            hipEvent_t* e = &s->ev[(size_t)k * sp_sim::EV_PER];
            float t;
            HIPCHK(hipEventElapsedTime(&t, e[5], e[6]));
            kp += t;
        }
        s->particle_ms = (float)(kp / s->timed);
        s->timed_last = s->timed;
        s->timed = 0;
    }
    readback(s);
    s->timings[7] = (float)s->stats.pressure_iters_last;
    s->timings[8] = (float)s->last_v_iters;
    s->timings[9] = (float)s->last_p2_iters;
    return SP_OK;
}

template <class F>
int guarded(sp_sim* s, F&& f) {
    if (!s) return SP_E_INVALID;
    try {
        return f();
    } catch (const HipError& e) {
        set_err(s, "%s failed: %s", e.what, hipGetErrorString(e.e));
        return SP_E_DEVICE;
    } catch (const std::bad_alloc&) {
        set_err(s, "host allocation failed");
        return SP_E_NOMEM;
    } catch (...) {
        set_err(s, "unexpected error");
        return SP_E_DEVICE;
    }
}

// The rows the particles carry, with their diluents (note_materials, mod.rs:2051-2081).
bool note_materials(sp_sim* s, const sp_particle* p, uint32_t n, bool reset) {
    std::vector<int> present = reset ? std::vector<int>() : s->tank.rows;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t low = p[k].kind & 255u;
        if (p[k].kind == SP_KIND_DEAD || low == 0u) continue;
        int r = (int)low - 1;
        if (std::find(present.begin(), present.end(), r) == present.end()) present.push_back(r);
    }
    std::vector<int> rows = sp::closure(s->params, present);
    if (rows == s->tank.rows) return true;
    sp::Tank t;
    if (!sp::tank_layout(s->params, rows, t)) return false;
    s->tank = t;
    s->tables_dirty = true;
    return true;
}

}  // namespace

// ---- C ABI ------------------------------------------------------------------------------

extern "C" {

const char* sp_create_error(void) { return g_create_error.c_str(); }

const char* sp_last_error(const sp_sim* sim) { return sim ? sim->err.c_str() : "null handle"; }

void sp_default_params(sp_sim_params* out) {
    if (out) sp::default_params(out);
}

int sp_levels(int32_t grid_w, int32_t grid_h, sp_levels_info* out) {
    if (!out || grid_w < 8 || grid_h < 8) return SP_E_INVALID;
    sp::levels(grid_w, grid_h, *out);
    return SP_OK;
}

int sp_derive_grid_params(const sp_sim_params* params, int32_t grid_w, int32_t grid_h, const int32_t* rows, int32_t n_rows,
                          sp_gparams* out, sp_gmat* mats) {
    if (!params || !out || !mats || (n_rows > 0 && !rows) || n_rows < 0) return SP_E_INVALID;
    std::vector<int> present(rows, rows + n_rows);
    std::vector<int> r = sp::closure(*params, present);
    sp::Tank t;
    if (!sp::tank_layout(*params, r, t)) return SP_E_CAPACITY;
    return sp::derive(*params, grid_w, grid_h, t, 0, *out, mats, nullptr);
}

int sp_create(const sp_config* cfg, sp_sim** out) {
    if (!cfg || !out) {
        g_create_error = "null argument";
        return SP_E_INVALID;
    }
    *out = nullptr;
    if (cfg->grid_w < 8 || cfg->grid_h < 8 || cfg->grid_w > 8192 || cfg->grid_h > 8192) {
        g_create_error = "grid must be between 8 x 8 and 8192 x 8192 cells";
        return SP_E_INVALID;
    }
    sp_sim* s = new (std::nothrow) sp_sim();
    if (!s) {
        g_create_error = "host allocation failed";
        return SP_E_NOMEM;
    }
    try {
        s->W = cfg->grid_w;
        s->H = cfg->grid_h;
        s->N = s->W * s->H;
        s->cap = cfg->capacity;
        s->device = cfg->device;
        HIPCHK(hipSetDevice(s->device));
        HIPCHK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->ev.assign((size_t)sp_sim::TIMED_SUBSTEPS * sp_sim::EV_PER, nullptr);
        for (auto& e : s->ev) HIPCHK(hipEventCreate(&e));
        sp::levels(s->W, s->H, s->lv);
        s->nb0 = ((s->W + 15) / 16) * ((s->H + 15) / 16);
        {
            const sp_levels_info& L = s->lv;
            size_t total = (size_t)(L.offs[L.n - 1] + L.dims[L.n - 1][0] * L.dims[L.n - 1][1] - L.offs[L.tail]);
            size_t nf = (size_t)L.dims[L.tail][0] * L.dims[L.tail][1];
            s->tail_lds_bytes = total * (16 + 8 + 4) + (total - nf) * 8;
            if (total > SP_TAIL_SHARED || total - nf > SP_TAIL_SHARED_COARSE) {
                sp_destroy(s);
                g_create_error = "multigrid tail does not fit its shared-memory budget";
                return SP_E_INVALID;
            }
        }
        const size_t N = (size_t)s->N, T = (size_t)s->lv.total;
        auto& o = s->owned;
        Ctx& c = s->c;
        c.W = s->W;
        c.H = s->H;
        c.N = s->N;
        s->d_P = dalloc<sp_gparams>(1, o);
        s->d_mats = dalloc<sp_gmat>(SP_MAX_MATERIALS, o);
        c.P = s->d_P;
        c.mats = s->d_mats;
        c.S = dalloc<sp_stepbuf>(1, o);
        c.K = dalloc<sp_solve>(1, o);
        c.part = dalloc<sp_particle>(s->cap, o);
        s->d_part_tmp = dalloc<sp_particle>(s->cap, o);
        s->d_spawn = dalloc<sp_particle>(SP_SPAWN_MAX, o);
        s->d_counter = dalloc<uint32_t>(8 + 2 * SP_MAX_MATERIALS, o);
        c.acc = dalloc<int32_t>(N * SP_ACC_MAX, o);
        c.vel = dalloc<float4>(N, o);
        c.cell = dalloc<float4>(N, o);
        c.comp = dalloc<float4>(N, o);
        c.coef = dalloc<float4>(N, o);
        c.surf = dalloc<float4>(N, o);
        c.surf2 = dalloc<float4>(N, o);
        c.sprop = dalloc<float4>(N, o);
        c.aux = dalloc<float4>(N, o);
        c.grho = dalloc<float4>(N, o);
        c.gbase = dalloc<float4>(N, o);
        c.prhs = dalloc<float4>(N, o);
        c.vw = dalloc<float4>(N, o);
        c.vs = dalloc<float4>(N, o);
        c.solid = dalloc<float4>(N, o);
        c.vel0 = dalloc<float2>(N, o);
        c.pjump = dalloc<float2>(N, o);
        c.pres = dalloc<float2>(N, o);
        c.sdf = dalloc<float>((size_t)(s->W + 1) * (s->H + 1), o);
        c.mixb = dalloc<float>(N * SP_MAX_PRESENT, o);
        c.p2keep = dalloc<float>(N, o);
        c.region = dalloc<uint32_t>(N, o);
        c.region_air = dalloc<uint32_t>(1, o);
        c.mga = dalloc<float4>(T, o);
        c.mgd = dalloc<float>(T, o);
        c.mgx = dalloc<float2>(T, o);
        c.mgz = dalloc<float2>(T, o);
        c.mgr = dalloc<float2>(T, o);
        c.cgp = dalloc<float2>(N, o);
        c.cgp2 = dalloc<float2>(N, o);
        c.cgq = dalloc<float2>(N, o);
        s->d_r2 = dalloc<float2>(N, o);
        c.ptile = dalloc<uint32_t>((size_t)s->lv.tiles, o);
        c.part_sums = dalloc<double4>((size_t)2 * s->nb0, o);
        c.vr = dalloc<float2>(N, o);
        c.vp = dalloc<float2>(N, o);
        c.vp2 = dalloc<float2>(N, o);
        c.vq = dalloc<float2>(N, o);
        c.vd = dalloc<float2>(N, o);
        c.vr2 = dalloc<float2>(N, o);
        c.vq2 = dalloc<float2>(N, o);
        c.vaz = dalloc<float2>(N, o);
        c.vtiles = dalloc<uint32_t>((size_t)s->nb0, o);
        c.vtile = dalloc<uint32_t>((size_t)s->lv.tiles, o);
        c.vma = dalloc<float4>(T, o);
        c.vmb = dalloc<float4>(T, o);
        c.vmd = dalloc<float2>(T, o);
        c.vmr = dalloc<float2>(T, o);
        c.vmx = dalloc<float2>(T, o);
        c.vmz = dalloc<float2>(T, o);
        c.vpart = dalloc<double4>((size_t)s->nb0, o);
        int4* lev = dalloc<int4>(SP_MAX_LEVELS, o);
        int4 h_lev[SP_MAX_LEVELS] = {};
        for (int l = 0; l < s->lv.n; l++) h_lev[l] = make_int4(s->lv.offs[l], s->lv.dims[l][0], s->lv.dims[l][1], s->lv.tile_offs[l]);
        HIPCHK(hipMemcpy(lev, h_lev, sizeof(h_lev), hipMemcpyHostToDevice));
        c.lev = lev;
        c.n_levels = s->lv.n;
        c.tail = s->lv.tail;
        s->d_keys = dalloc<uint32_t>(s->cap, o);
        s->d_vals = dalloc<uint32_t>(s->cap, o);
        s->d_keys2 = dalloc<uint32_t>(s->cap, o);
        s->d_vals2 = dalloc<uint32_t>(s->cap, o);
        s->d_count_out = dalloc<uint32_t>(N, o);
        size_t bytes = 0;
        HIPCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, s->d_keys, s->d_keys2, s->d_vals, s->d_vals2,
                                                  (int)std::max<uint32_t>(s->cap, 1), 0, 32, s->stream));
        s->sort_tmp_bytes = bytes;
        s->d_sort_tmp = dalloc<uint8_t>(bytes, o);
        // solid field: closed everywhere until set_solids
        HIPCHK(hipDeviceSynchronize());
        sp::default_params(&s->params);
        s->timing = true;
        // SPROUT_GRAPHS=0 launches the substep's kernels one by one instead of replaying
        // captured graphs (profilers that cannot follow graph launches need it)
        const char* gv = getenv("SPROUT_GRAPHS");
        s->use_graphs = !(gv && gv[0] == '0');
    } catch (const HipError& e) {
        char buf[256];
        snprintf(buf, sizeof(buf), "%s failed: %s", e.what, hipGetErrorString(e.e));
        g_create_error = buf;
        sp_destroy(s);
        return e.e == hipErrorOutOfMemory ? SP_E_NOMEM : SP_E_DEVICE;
    }
    *out = s;
    return SP_OK;
}

void sp_destroy(sp_sim* s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    destroy_graphs(s);
    for (void* p : s->owned) (void)hipFree(p);
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int sp_set_params(sp_sim* s, const sp_sim_params* p) {
    return guarded(s, [&]() -> int {
        if (!p) return SP_E_INVALID;
        const float* f = (const float*)p;
        (void)f;
        float check[] = {p->cell_size, p->gravity, p->ambient_temp, p->cfl, p->dt_max, p->fixed_hz, p->pressure_tol,
                         p->correction_tol, p->correction_time, p->correction_max, p->speed_cap, p->apic_clamp};
        for (float v : check)
            if (v != v) {
                set_err(s, "NaN parameter refused");
                return SP_E_INVALID;
            }
        if (p->cell_size <= 0.0f || p->pressure_max_iters == 0 || p->mat_count < 1 || p->mat_count > SP_MAX_MATERIALS) {
            set_err(s, "parameter out of range");
            return SP_E_INVALID;
        }
        s->params = *p;
        s->params_set = true;
        s->tables_dirty = true;
        s->rec = Recs{p->pressure_max_iters, p->pressure_max_iters, std::max(p->viscosity_max_iters, 1u)};
        // the tank's layout follows the table (diluents may have changed)
        std::vector<int> rows = sp::closure(s->params, s->tank.rows);
        sp::Tank t;
        if (sp::tank_layout(s->params, rows, t)) s->tank = t;
        return SP_OK;
    });
}

int sp_set_solids(sp_sim* s, const float* sdf, const float* solid, const uint32_t* region, const float* temp) {
    return guarded(s, [&]() -> int {
        if (!sdf || !solid || !region) return SP_E_INVALID;
        if (temp) {
            set_err(s, "solids held at a temperature make the tank thermal, which this build does not carry");
            return SP_E_UNSUPPORTED;
        }
        HIPCHK(hipMemcpyAsync((void*)s->c.sdf, sdf, sizeof(float) * (size_t)(s->W + 1) * (s->H + 1), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync((void*)s->c.solid, solid, sizeof(float4) * (size_t)s->N, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync((void*)s->c.region, region, sizeof(uint32_t) * (size_t)s->N, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_reset(sp_sim* s, const sp_particle* particles, uint32_t n) {
    return guarded(s, [&]() -> int {
        if (n > 0 && !particles) return SP_E_INVALID;
        n = std::min(n, s->cap);
        std::vector<sp_particle> ps;
        ps.reserve(n);
        for (uint32_t k = 0; k < n; k++) {
            sp_particle q = particles[k];
            q.id = k;
            sp::settle_spawn(s->params, q);
            uint32_t low = q.kind & 255u;
            if (q.kind == SP_KIND_DEAD || low == 0u || (int)low - 1 >= s->params.mat_count) continue;
            ps.push_back(q);
        }
        if (!note_materials(s, ps.data(), (uint32_t)ps.size(), true)) {
            set_err(s, "the particles hold more materials than the accumulators have channels for");
            return SP_E_CAPACITY;
        }
        s->n_slots = (uint32_t)ps.size();
        s->next_id = n;
        s->tables_dirty = true;
        if (!ps.empty())
            HIPCHK(hipMemcpyAsync(s->c.part, ps.data(), sizeof(sp_particle) * ps.size(), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemsetAsync(s->c.pres, 0, sizeof(float2) * (size_t)s->N, s->stream));
        HIPCHK(hipMemsetAsync(s->c.p2keep, 0, sizeof(float) * (size_t)s->N, s->stream));
        HIPCHK(hipMemsetAsync(s->c.vd, 0, sizeof(float2) * (size_t)s->N, s->stream));
        HIPCHK(hipMemsetAsync(s->c.S, 0, sizeof(sp_stepbuf), s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        s->pending_spawn.clear();
        s->time_base = s->sim_time;
        s->substeps_base = s->stats.substeps;
        s->grid_valid = false;
        s->dt_estimate = 1.0f / 240.0f;
        // a new scene starts from no pressure and no idea how many iterations its solves
        // need: record the most until read-backs say otherwise (mod.rs:2478-2483)
        s->rec = Recs{s->params.pressure_max_iters, s->params.pressure_max_iters, std::max(s->params.viscosity_max_iters, 1u)};
        // a sort before the first substep puts neighbours next to each other
        s->since_sort = s->sort_interval;
        return SP_OK;
    });
}

int sp_spawn(sp_sim* s, const sp_particle* particles, uint32_t n, uint32_t* accepted) {
    return guarded(s, [&]() -> int {
        if (accepted) *accepted = 0;
        if (n > 0 && !particles) return SP_E_INVALID;
        uint32_t room = SP_SPAWN_MAX - (uint32_t)s->pending_spawn.size();
        n = std::min(n, room);
        std::vector<sp_particle> ps(particles, particles + n);
        for (uint32_t k = 0; k < n; k++) {
            uint32_t low = ps[k].kind & 255u;
            if (ps[k].kind == SP_KIND_DEAD || low == 0u || (int)low - 1 >= s->params.mat_count) {
                set_err(s, "spawned particle %u has no material row of the table", k);
                return SP_E_INVALID;
            }
        }
        for (uint32_t k = 0; k < n; k++) {
            ps[k].id = s->next_id + k;
            sp::settle_spawn(s->params, ps[k]);
        }
        if (!note_materials(s, ps.data(), n, false)) return SP_OK;
        for (auto& q : ps) {
            q.id = s->next_id++;
            q.kind |= SP_KIND_FRESH;
            s->pending_spawn.push_back(q);
        }
        if (accepted) *accepted = n;
        return SP_OK;
    });
}

int sp_record(sp_sim* s, uint32_t substeps, float target_s, const sp_tool_state* tool) {
    return guarded(s, [&]() -> int {
        if (tool && tool->mode != 0) {
            set_err(s, "tools are not carried by this build (mode %d)", tool->mode);
            return SP_E_UNSUPPORTED;
        }
        return record(s, substeps, target_s);
    });
}

int sp_wait(sp_sim* s) {
    return guarded(s, [&]() -> int { return wait(s); });
}

int sp_poll(sp_sim* s, sp_solver_stats* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = s->stats;
        return SP_OK;
    });
}

int sp_step_blocking(sp_sim* s, uint32_t n) {
    return guarded(s, [&]() -> int {
        if (n == 0) {
            int rc = record(s, 0, 0.0f);
            if (rc != SP_OK) return rc;
            return wait(s);
        }
        // step_blocking (mod.rs:3316-3338)
        double seconds = (double)n * (1.0 / 240.0);
        double start = s->sim_time;
        double sliver = std::max((double)s->params.min_substep_share, 1e-6);
        for (int guard = 0; guard < 1000000; guard++) {
            double left = seconds - (s->sim_time - start);
            double dt = s->params.fixed_hz > 0.0f ? 1.0 / s->params.fixed_hz : (double)s->dt_estimate;
            if (left <= sliver * dt) break;
            uint32_t k = (uint32_t)std::min(std::max(ceil(left / dt), 1.0), 16.0);
            double target = s->params.fixed_hz > 0.0f ? k * dt : left;
            int rc = record(s, k, (float)target);
            if (rc != SP_OK) return rc;
            rc = wait(s);
            if (rc != SP_OK) return rc;
        }
        return SP_OK;
    });
}

int sp_advance_blocking(sp_sim* s, double seconds) {
    return guarded(s, [&]() -> int {
        // advance_blocking (mod.rs:3284-3312)
        double goal = s->sim_time + seconds;
        double sliver = std::max((double)s->params.min_substep_share, 1e-6);
        for (int guard = 0; guard < 1000000; guard++) {
            double left = goal - s->sim_time;
            double dt = s->params.fixed_hz > 0.0f ? 1.0 / s->params.fixed_hz : (double)s->dt_estimate;
            if (left <= sliver * dt) break;
            double nn = s->params.fixed_hz > 0.0f ? std::max(round(left / dt), 1.0) : ceil(left / dt);
            uint32_t k = (uint32_t)std::min(std::max(nn, 1.0), 16.0);
            double target = s->params.fixed_hz > 0.0f ? k * dt : left;
            int rc = record(s, k, (float)target);
            if (rc != SP_OK) return rc;
            rc = wait(s);
            if (rc != SP_OK) return rc;
            if (s->params.fixed_hz > 0.0f && (goal - s->sim_time) < 0.5 * dt) break;
        }
        return SP_OK;
    });
}

int sp_sim_time(sp_sim* s, double* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = s->sim_time;
        return SP_OK;
    });
}

int sp_download_particles(sp_sim* s, sp_particle* out, uint32_t cap, uint32_t* n) {
    return guarded(s, [&]() -> int {
        if (!n) return SP_E_INVALID;
        std::vector<sp_particle> all(s->n_slots);
        if (s->n_slots)
            HIPCHK(hipMemcpyAsync(all.data(), s->c.part, sizeof(sp_particle) * s->n_slots, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        uint32_t k = 0;
        for (const auto& q : all) {
            if (q.kind == SP_KIND_DEAD || (q.kind & 255u) == 0u) continue;
            if (out) {
                if (k >= cap) break;
                out[k] = q;
            }
            k++;
        }
        *n = k;
        return SP_OK;
    });
}

int sp_download_grid(sp_sim* s, float* vel, float* cell) {
    return guarded(s, [&]() -> int {
        if (vel) HIPCHK(hipMemcpyAsync(vel, s->c.vel, sizeof(float4) * (size_t)s->N, hipMemcpyDeviceToHost, s->stream));
        if (cell) HIPCHK(hipMemcpyAsync(cell, s->c.cell, sizeof(float4) * (size_t)s->N, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_download_pressure(sp_sim* s, float* pres, float* dt) {
    return guarded(s, [&]() -> int {
        if (!pres) return SP_E_INVALID;
        std::vector<float2> p((size_t)s->N);
        HIPCHK(hipMemcpyAsync(p.data(), s->c.pres, sizeof(float2) * (size_t)s->N, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int i = 0; i < s->N; i++) pres[i] = p[i].x;
        if (dt) *dt = s->stats.dt;
        return SP_OK;
    });
}

int sp_cell_counts(sp_sim* s, int32_t x0, int32_t y0, int32_t w, int32_t h, uint32_t* out) {
    return guarded(s, [&]() -> int {
        if (!out || w < 0 || h < 0 || x0 < 0 || y0 < 0 || x0 + w > s->W || y0 + h > s->H) return SP_E_INVALID;
        if (w == 0 || h == 0) return SP_OK;
        HIPCHK(hipMemsetAsync(s->d_count_out, 0, sizeof(uint32_t) * (size_t)w * h, s->stream));
        if (s->n_slots)
            hipLaunchKernelGGL(k_count, dim3((s->n_slots + 255) / 256), dim3(256), 0, s->stream, s->c, x0, y0, w, h, s->d_count_out);
        HIPCHK(hipMemcpyAsync(out, s->d_count_out, sizeof(uint32_t) * (size_t)w * h, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_stats(sp_sim* s, sp_stats_out* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        const int words = 8 + 2 * SP_MAX_MATERIALS;
        HIPCHK(hipMemsetAsync(s->d_counter, 0, sizeof(uint32_t) * words, s->stream));
        uint32_t runs = (s->n_slots + 63) / 64;
        if (runs) hipLaunchKernelGGL(k_stats, dim3((runs + 63) / 64), dim3(64), 0, s->stream, s->c, s->d_counter);
        std::vector<uint32_t> h(words);
        HIPCHK(hipMemcpyAsync(h.data(), s->d_counter, sizeof(uint32_t) * words, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        out->live = h[4];
        for (int m = 0; m < SP_MAX_MATERIALS; m++) {
            out->count[m] = h[8 + 2 * m];
            out->conc[m] = (float)h[9 + 2 * m] / 256.0f;
        }
        return SP_OK;
    });
}

int sp_sort(sp_sim* s) {
    return guarded(s, [&]() -> int {
        int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        launch_sort(s, 0);
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_timings(sp_sim* s, float* out10) {
    return guarded(s, [&]() -> int {
        if (!out10) return SP_E_INVALID;
        for (int k = 0; k < 10; k++) out10[k] = s->timings[k];
        return SP_OK;
    });
}

int sp_kernel_time(sp_sim* s, int32_t kernel, float* avg_ms, uint32_t* launches) {
    return guarded(s, [&]() -> int {
        if (!avg_ms || !launches || kernel != 0) return SP_E_INVALID;
        *avg_ms = s->particle_ms;
        *launches = s->timed_last;
        return SP_OK;
    });
}

static void* field_ptr(sp_sim* s, int32_t field, size_t* bytes) {
    const size_t N = (size_t)s->N, T = (size_t)s->lv.total;
    const Ctx& c = s->c;
    switch (field) {
    case SP_FIELD_PARTICLES: *bytes = sizeof(sp_particle) * s->n_slots; return c.part;
    case SP_FIELD_ACC: *bytes = sizeof(int32_t) * N * s->gp.acc_stride; return c.acc;
    case SP_FIELD_VEL: *bytes = 16 * N; return c.vel;
    case SP_FIELD_CELL: *bytes = 16 * N; return c.cell;
    case SP_FIELD_COMP: *bytes = 16 * N; return c.comp;
    case SP_FIELD_COEF: *bytes = 16 * N; return c.coef;
    case SP_FIELD_SURF: *bytes = 16 * N; return c.surf;
    case SP_FIELD_PJUMP: *bytes = 8 * N; return c.pjump;
    case SP_FIELD_MGA: *bytes = 16 * T; return c.mga;
    case SP_FIELD_MGD: *bytes = 4 * T; return c.mgd;
    case SP_FIELD_PRHS: *bytes = 16 * N; return c.prhs;
    case SP_FIELD_PRES: *bytes = 8 * N; return c.pres;
    case SP_FIELD_STEP: *bytes = sizeof(sp_stepbuf); return c.S;
    case SP_FIELD_SOLVE: *bytes = sizeof(sp_solve); return c.K;
    case SP_FIELD_SPROP: *bytes = 16 * N; return c.sprop;
    case SP_FIELD_AUX: *bytes = 16 * N; return c.aux;
    case SP_FIELD_GRHO: *bytes = 16 * N; return c.grho;
    case SP_FIELD_VEL0: *bytes = 8 * N; return c.vel0;
    case SP_FIELD_MGZ: *bytes = 8 * T; return c.mgz;
    case SP_FIELD_MGR: *bytes = 8 * T; return c.mgr;
    case SP_FIELD_PTILE: *bytes = 4 * (size_t)s->lv.tiles; return c.ptile;
    case SP_FIELD_REGION_AIR: *bytes = 4; return c.region_air;
    case SP_FIELD_SURF2: *bytes = 16 * N; return c.surf2;
    case SP_FIELD_VW: *bytes = 16 * N; return c.vw;
    case SP_FIELD_VS: *bytes = 16 * N; return c.vs;
    case SP_FIELD_GBASE: *bytes = 16 * N; return c.gbase;
    case SP_FIELD_VR: *bytes = 8 * N; return c.vr;
    case SP_FIELD_VP: *bytes = 8 * N; return c.vp;
    case SP_FIELD_VP2: *bytes = 8 * N; return c.vp2;
    case SP_FIELD_VQ: *bytes = 8 * N; return c.vq;
    case SP_FIELD_VD: *bytes = 8 * N; return c.vd;
    case SP_FIELD_VR2: *bytes = 8 * N; return c.vr2;
    case SP_FIELD_VQ2: *bytes = 8 * N; return c.vq2;
    case SP_FIELD_VAZ: *bytes = 8 * N; return c.vaz;
    case SP_FIELD_VTILES: *bytes = 4 * (size_t)s->nb0; return c.vtiles;
    case SP_FIELD_VTILE: *bytes = 4 * (size_t)s->lv.tiles; return c.vtile;
    case SP_FIELD_VMA: *bytes = 16 * T; return c.vma;
    case SP_FIELD_VMB: *bytes = 16 * T; return c.vmb;
    case SP_FIELD_VMD: *bytes = 8 * T; return c.vmd;
    case SP_FIELD_VMR: *bytes = 8 * T; return c.vmr;
    case SP_FIELD_VMX: *bytes = 8 * T; return c.vmx;
    case SP_FIELD_VMZ: *bytes = 8 * T; return c.vmz;
    case SP_FIELD_P2KEEP: *bytes = 4 * N; return c.p2keep;
    default: *bytes = 0; return nullptr;
    }
}

int sp_debug_download(sp_sim* s, int32_t field, void* out, size_t bytes) {
    return guarded(s, [&]() -> int {
        size_t n = 0;
        void* p = field_ptr(s, field, &n);
        if (!p || !out || bytes < n) {
            set_err(s, "bad field %d or buffer too small (%zu < %zu)", field, bytes, n);
            return SP_E_INVALID;
        }
        HIPCHK(hipMemcpyAsync(out, p, n, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_debug_upload(sp_sim* s, int32_t field, const void* data, size_t bytes) {
    return guarded(s, [&]() -> int {
        size_t n = 0;
        if (field == SP_FIELD_PARTICLES) {
            if (!data || bytes % sizeof(sp_particle) != 0 || bytes / sizeof(sp_particle) > s->cap) return SP_E_INVALID;
            s->n_slots = (uint32_t)(bytes / sizeof(sp_particle));
            s->tables_dirty = true;
            int rc = upload_tables(s);
            if (rc != SP_OK) return rc;
            if (bytes) HIPCHK(hipMemcpyAsync(s->c.part, data, bytes, hipMemcpyHostToDevice, s->stream));
            HIPCHK(hipStreamSynchronize(s->stream));
            return SP_OK;
        }
        void* p = field_ptr(s, field, &n);
        if (!p || !data || bytes != n) {
            set_err(s, "bad field %d or size (%zu != %zu)", field, bytes, n);
            return SP_E_INVALID;
        }
        HIPCHK(hipMemcpyAsync(p, data, n, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_debug_gparams(sp_sim* s, sp_gparams* out, sp_gmat* mats) {
    return guarded(s, [&]() -> int {
        int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        if (out) *out = s->gp;
        if (mats) memcpy(mats, s->gm, sizeof(s->gm));
        return SP_OK;
    });
}

int sp_debug_stage(sp_sim* s, int32_t stage, int32_t arg) {
    return guarded(s, [&]() -> int {
        int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        const Ctx& c = s->c;
        dim3 g0 = grid_groups(s->W, s->H);
        // stages run with the most iterations the parameters allow
        const Recs full{s->params.pressure_max_iters, s->params.pressure_max_iters, std::max(s->params.viscosity_max_iters, 1u)};
        switch (stage) {
        case SP_STAGE_STEP: hipLaunchKernelGGL(k_step, dim3(1), dim3(64), 0, s->stream, c); break;
        case SP_STAGE_CLEAR_ACC: launch_clear_acc(s); break;
        case SP_STAGE_PARTICLES: launch_particles(s, arg); break;
        case SP_STAGE_GRID_PREP: hipLaunchKernelGGL(k_grid_prep, g0, TILE, 0, s->stream, c); break;
        case SP_STAGE_COEF: hipLaunchKernelGGL(k_coef_smooth, g0, TILE, 0, s->stream, c); break;
        case SP_STAGE_SURF:
            hipLaunchKernelGGL(k_coef_smooth, g0, TILE, 0, s->stream, c);
            hipLaunchKernelGGL(k_surf_curv, g0, TILE, 0, s->stream, c);
            hipLaunchKernelGGL(k_surf_apply, g0, TILE, 0, s->stream, c);
            break;
        case SP_STAGE_VISC: launch_visc_before(s, full); break;
        case SP_STAGE_VISC_SPLIT: launch_visc_split(s); break;
        case SP_STAGE_VISC_SOLVE: launch_visc_solve(s, full.v, arg != 0); break;
        case SP_STAGE_RHS_VISC: hipLaunchKernelGGL(k_rhs_visc, g0, TILE, 0, s->stream, c); break;
        case SP_STAGE_PRESSURE2: launch_pressure(s, full.p2, true); break;
        case SP_STAGE_PROJECT_VISC: hipLaunchKernelGGL(k_project_visc, g0, TILE, 0, s->stream, c); break;
        case SP_STAGE_VISC_VCYCLE: {
            // One V-cycle from vr, whatever the state of the last solve: the level kernels
            // return at once on a converged solve, so the flag is cleared around it.
            uint32_t done = 0, zero = 0;
            HIPCHK(hipMemcpy(&done, &c.K->v_done, 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&c.K->v_done, &zero, 4, hipMemcpyHostToDevice));
            if (s->gp.v_mg_sweeps == 1) {
                hipLaunchKernelGGL(k_vmg_down<1>, g0, TILE, 0, s->stream, c, 0, 0);
                launch_visc_coarse<1>(s);
                hipLaunchKernelGGL(k_visc_mgz<1>, g0, TILE, 0, s->stream, c, (const float2*)c.vr, (const float2*)c.vq);
            } else {
                hipLaunchKernelGGL(k_vmg_down<2>, g0, TILE, 0, s->stream, c, 0, 0);
                launch_visc_coarse<2>(s);
                hipLaunchKernelGGL(k_visc_mgz<2>, g0, TILE, 0, s->stream, c, (const float2*)c.vr, (const float2*)c.vq);
            }
            HIPCHK(hipStreamSynchronize(s->stream));
            HIPCHK(hipMemcpy(&c.K->v_done, &done, 4, hipMemcpyHostToDevice));
            break;
        }
        case SP_STAGE_RHS: hipLaunchKernelGGL(k_rhs, g0, TILE, 0, s->stream, c); break;
        case SP_STAGE_PRESSURE: launch_pressure(s, full.p); break;
        case SP_STAGE_PROJECT: hipLaunchKernelGGL(k_project, g0, TILE, 0, s->stream, c); break;
        case SP_STAGE_SUBSTEP:
            launch_substep(s, arg, full);
            s->grid_valid = true;
            break;
        case SP_STAGE_MG_BUILD: launch_mg_build(s); break;
        case SP_STAGE_VCYCLE: launch_vcycle(s, 0, -1, s->c.mgr); break;
        default: set_err(s, "unknown stage %d", stage); return SP_E_INVALID;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

}  // extern "C"
