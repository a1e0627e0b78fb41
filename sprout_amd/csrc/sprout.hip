// sprout.hip — libsprout_hip.so: the device state of one tank and the C ABI of
// sprout_hip.h (the boundary that replaces FluidSim's Vulkano dispatch,
// src/sim/mod.rs:862-3492). The substep schedule is in schedule.hip, batches of tanks in
// batch.hip.
//
// One translation unit: the kernel files are included here so that the kernels and the
// launch code see each other's inline helpers.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cstdarg>
#include <mutex>
#include <string>
#include <vector>

#include "bin.hip"
#include "grid.hip"
#include "host.h"
#include "particles.hip"
#include "solve.hip"
#include "visc.hip"
#include "heat.hip"
#include "rigid.hip"
#include "elastic.hip"
#include "settle.hip"
#include "magnet.hip"
#include "schedule.hip"

namespace {

thread_local std::string g_create_error;
thread_local std::string g_call_error;  // the last error of a handle-free entry point

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

struct sp_sim {
    int W = 0, H = 0, N = 0;
    uint32_t cap = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> owned;  // device buffers held for the handle's life
    std::vector<void*> big;    // the arena and the sort's work space, given up when evicted
    std::string err;

    // residency: an evicted tank keeps its arena in pinned host memory
    bool resident = true;
    void* host_arena = nullptr;
    std::vector<std::pair<size_t, size_t>> host_ranges;  // (arena offset, bytes) of what host_arena holds
    size_t arena_bytes = 0;
    uint64_t generation = 0;  // bumped when the arena moves
    uint64_t resets = 0;      // sp_reset calls (FluidSim::resets)
    std::vector<sp_magnet> magnets;  // placed in the world (FluidSim::magnets); survive a reset
    std::vector<sp_current> currents;  // currents and drains (FluidSim::currents); survive a reset

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
    TankRef ref{};            // the tank's arena and sizes
    Ctx c{};                  // its buffers (host copy of what every kernel rebuilds)
    Run run;                  // the launch group of this tank alone
    sp_particle* d_part_tmp = nullptr;
    float4* d_stress_tmp = nullptr;  // the sort's target for the particles' stress
    sp_particle* d_spawn = nullptr;
    uint32_t* d_counter = nullptr;  // stats / claim counters
    uint32_t* d_count_out = nullptr;
    void* d_bin = nullptr;  // the sort's and the claim's work space (bin_space_bytes(cap))

    // host state mirroring FluidSim
    uint32_t n_slots = 0;
    uint32_t next_id = 0;
    bool grid_valid = false;
    uint32_t since_sort = 0;
    uint32_t sort_interval = 8;
    std::vector<sp_particle> pending_spawn;
    float spawn_speed = 0.0f;  // fastest |vel| of the spawns queued for the next record (cells / s)
    sp_tool_state tool{};      // the mouse tool of the latest record (idle: mode 0)
    double time_base = 0.0;
    double sim_time = 0.0;
    float dt_estimate = 1.0f / 240.0f;
    Recs rec{60, 60, 200};
    sp_solver_stats stats{};
    uint64_t substeps_base = 0;
    uint32_t last_unconv = 0, last_p2_unconv = 0, last_v_unconv = 0;
    uint32_t last_v_iters = 0, last_p2_iters = 0;  // of the latest solves

    GraphSet gs;
    Timing tm;
    // Read-backs of the batches sp_record enqueued (poll_readbacks, mod.rs:3030-3115): each
    // batch ends with a copy of the step and solve records into a slot of a ring in pinned
    // host memory and an event, with events around its substeps for substep_cost
    // (mod.rs:3055-3059). Finished slots are taken up in order without blocking; only then
    // does the tank adapt its recorded iterations, so a caller may keep batches in flight.
    static constexpr uint32_t RB_RING = 4;
    struct Readback {
        sp_stepbuf S;
        sp_solve K;
    };
    Readback* rb = nullptr;  // RB_RING slots, pinned
    hipEvent_t rb_start[RB_RING] = {}, rb_end[RB_RING] = {};
    uint32_t rb_substeps[RB_RING] = {};
    uint32_t rb_next = 0, rb_busy = 0;  // oldest slot in flight, slots in flight
    float substep_cost = 0.0f;          // s, smoothed
    sp_gpumat* d_gpumat = nullptr;  // the material table in GpuMat layout (SP_RENDER_GPUMAT)
    sp_gpumat gpumat[SP_MAX_MATERIALS]{};
    bool in_batch = false;  // a batch drives the tank (sp_batch_*)
    bool pooled = false;    // a pool drives the tank (sp_pool_*)

    // heat (FluidSim::thermal, solids_heated, mod.rs:959-962)
    bool thermal = false;        // temperatures carried and conducted since the last reset
    bool solids_heated = false;  // the solids hold temperatures
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

// The counting sort of the particles by cell (record_sort, mod.rs:3223-3242), on `stream`:
// keys are the cell (or the 8 x 8 tile, then the cell, with `tiled`), dead slots last;
// stable by the previous slot (bin.hip).
void launch_sort(sp_sim* s, int tiled, hipStream_t stream) {
    if (s->n_slots == 0) return;
    const uint32_t blocks = (s->n_slots + 255) / 256;
    const int bits = bin_bits(s->W, s->H, tiled);
    BinSpace sp = bin_space_at(s->d_bin, s->cap);
    hipLaunchKernelGGL(k_sort_keys, dim3(blocks), dim3(256), 0, stream, s->run.tb, sp.keys, sp.vals, s->n_slots, tiled,
                       bin_dead_key(bits));
    const uint32_t* order = launch_bin_passes(sp, s->n_slots, bits, stream);
    // the stress moves with its particle while the tank is elastic (record_sort, mod.rs:3238-3240)
    const bool elastic = (s->gp.caps & SP_CAP_ELASTIC) != 0u;
    hipLaunchKernelGGL(k_sort_gather, dim3(blocks), dim3(256), 0, stream, s->run.tb, order, s->d_part_tmp,
                       elastic ? s->d_stress_tmp : nullptr, s->n_slots);
    HIPCHK(hipMemcpyAsync(s->c.part, s->d_part_tmp, sizeof(sp_particle) * s->n_slots, hipMemcpyDeviceToDevice, stream));
    if (elastic)
        HIPCHK(hipMemcpyAsync(s->c.pstress, s->d_stress_tmp, sizeof(float4) * s->n_slots, hipMemcpyDeviceToDevice, stream));
}

// Dead slots claim the `claim` pending spawns in s->d_spawn, in slot order (claim.comp:10-17).
void launch_claim(sp_sim* s, uint32_t claim, hipStream_t stream) {
    if (claim == 0 || s->n_slots == 0) return;
    BinSpace sp = bin_space_at(s->d_bin, s->cap);
    const uint32_t groups = bin_groups(s->n_slots);
    const uint32_t nch = (groups + bin::SCAN_CHUNK - 1) / bin::SCAN_CHUNK;
    hipLaunchKernelGGL(k_claim_count, dim3(groups), dim3(bin::NT), 0, stream, s->run.tb, s->n_slots, sp.hist);
    hipLaunchKernelGGL(k_bin_scan_chunks, dim3(nch), dim3(bin::NT), 0, stream, sp.hist, groups, sp.chunk);
    hipLaunchKernelGGL(k_bin_scan_top, dim3(1), dim3(bin::TOP_NT), 0, stream, sp.chunk, nch);
    hipLaunchKernelGGL(k_claim, dim3(groups), dim3(bin::NT), 0, stream, s->run.tb, s->n_slots, (const sp_particle*)s->d_spawn, claim,
                       (const uint32_t*)sp.hist, (const uint32_t*)sp.chunk);
}

// Uploads the grid-unit tables if the parameters or the tank changed.
// A row of this library's material table in the reference's GpuMat order
// (gpu_materials, mod.rs:2202-2233). Fizz, which only the renderer reads, is not carried (0).
sp_gpumat to_gpumat(const sp_gmat& g) {
    sp_gpumat o;
    memset(&o, 0, sizeof(o));
    const int32_t ia[4] = {g.group, g.diluent, g.chan_w, g.chan_s}, ib[4] = {g.fam, g.rheo, g.temp_model, g.residue};
    const float a[4] = {g.rho, g.ln_nu, g.sigma, g.mix_rate}, b[4] = {g.cos_c, g.sin_c, g.grain, g.rho_solid};
    const float el[4] = {g.gr[1], g.gr[2], g.gr[3], g.residue_share};
    const float misc[4] = {g.freeze_shift, g.load, g.visc_exp, g.expansion};
    const float gr[4] = {g.gr[0], g.fume, 0.0f, 0.0f};
    for (int k = 0; k < 4; k++) {
        o.ia[k] = ia[k];
        o.ib[k] = ib[k];
        o.a[k] = a[k];
        o.b[k] = b[k];
        o.rheo[k] = g.rheo_p[k];
        o.temp[k] = g.temp[k];
        o.th[k] = g.th[k];
        o.th2[k] = g.th2[k];
        o.el[k] = el[k];
        o.misc[k] = misc[k];
        o.ic[k] = g.ic[k];
        o.rx[k] = g.rx[k];
        o.gr[k] = gr[k];
    }
    return o;
}

int upload_tables(sp_sim* s) {
    if (!s->tables_dirty) return SP_OK;
    std::string why;
    int rc = sp::derive(s->params, s->W, s->H, s->tank, s->n_slots, s->gp, s->gm, &why, s->magnets, s->spawn_speed,
                        s->currents, &s->tool);
    if (rc != SP_OK) {
        set_err(s, "%s", why.c_str());
        return rc;
    }
    HIPCHK(hipMemcpyAsync((void*)s->c.P, &s->gp, sizeof(sp_gparams), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync((void*)s->c.mats, s->gm, sizeof(s->gm), hipMemcpyHostToDevice, s->stream));
    for (int r = 0; r < SP_MAX_MATERIALS; r++) s->gpumat[r] = to_gpumat(s->gm[r]);
    if (s->d_gpumat)
        HIPCHK(hipMemcpyAsync(s->d_gpumat, s->gpumat, sizeof(s->gpumat), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->tables_dirty = false;
    s->run.gp = s->gp;
    s->run.slots = s->n_slots;
    return SP_OK;
}

// The tank's evictable device memory: the arena (every field, the particles, the multigrid
// levels, the solve record) and the sort's keys, values, work space and target array.
void alloc_big(sp_sim* s) {
    std::vector<void*> got;
    try {
        TankRef ref = s->ref;
        ref.base = dalloc<char>(s->arena_bytes, got);
        s->d_part_tmp = dalloc<sp_particle>(s->cap, got);
        s->d_stress_tmp = dalloc<float4>(s->cap, got);
        s->d_bin = dalloc<uint8_t>(bin_space_bytes(s->cap), got);
        s->ref = ref;
    } catch (...) {
        for (void* p : got) (void)hipFree(p);
        throw;
    }
    s->big = got;
    s->c = ctx_of(s->ref);
    s->run.tb.t[0] = s->ref;
    s->generation++;
}

void free_big(sp_sim* s) {
    for (void* p : s->big) (void)hipFree(p);
    s->big.clear();
    s->ref.base = nullptr;
}

// Bytes of device memory a tank of this arena and capacity gives back when evicted.
size_t big_bytes(size_t arena_bytes, uint32_t cap) {
    return arena_bytes + (sizeof(sp_particle) + sizeof(float4)) * (size_t)std::max<uint32_t>(cap, 1) + bin_space_bytes(cap);
}
size_t big_bytes(const sp_sim* s) { return big_bytes(s->arena_bytes, s->cap); }

// The byte ranges of the arena that carry a tank from one substep to the next: all of it but
// the work space every substep writes before it reads it (the accumulator, cleared first; the
// Krylov and multigrid iterates and residuals of both solves, their partial sums; the face
// coefficients, the right-hand sides and the pre-projection velocities, which the grid passes
// rewrite), the particle slots past the last live one, the particles' stress in a tank with
// nothing elastic (a tank that turns elastic zeroes it), and the per-family arrays past the
// families the tank has. Sorted, disjoint, merged.
std::vector<std::pair<size_t, size_t>> kept_ranges(const sp_sim* s) {
    TankRef r = s->ref;
    r.base = nullptr;
    const Ctx c = ctx_of(r);
    const size_t N = (size_t)s->N, T = (size_t)s->lv.total, cap = std::max<uint32_t>(s->cap, 1);
    const size_t nb0 = (size_t)s->nb0;
    const int nf = std::max(s->gp.n_fam, 0), nmix = std::max(s->gp.n_fam - 1, 0);
    auto off = [](const void* p) { return (size_t)(uintptr_t)p; };
    std::vector<std::pair<size_t, size_t>> skip;  // (start, end)
    auto drop = [&](const void* p, size_t from, size_t to) {
        if (to > from) skip.push_back({off(p) + from, off(p) + to});
    };
    drop(c.acc, 0, sizeof(int32_t) * N * SP_ACC_MAX);
    for (const void* f : {(const void*)c.coef, (const void*)c.prhs}) drop(f, 0, sizeof(float4) * N);
    for (const void* f : {(const void*)c.vel0, (const void*)c.cgp, (const void*)c.cgp2, (const void*)c.cgq, (const void*)c.r2,
                          (const void*)c.vr, (const void*)c.vp, (const void*)c.vp2, (const void*)c.vq, (const void*)c.vr2,
                          (const void*)c.vq2, (const void*)c.vaz})
        drop(f, 0, sizeof(float2) * N);
    for (const void* f : {(const void*)c.mgx, (const void*)c.mgz, (const void*)c.mgr, (const void*)c.vmr, (const void*)c.vmx,
                          (const void*)c.vmz})
        drop(f, 0, sizeof(float2) * T);
    drop(c.part_sums, 0, sizeof(double4) * 3 * nb0);
    drop(c.vpart, 0, sizeof(double4) * nb0);
    drop(c.part, sizeof(sp_particle) * s->n_slots, sizeof(sp_particle) * cap);
    drop(c.pstress, s->tank.chan_e >= 0 ? sizeof(float4) * s->n_slots : 0, sizeof(float4) * cap);
    drop(c.mixb, sizeof(float) * N * nmix, sizeof(float) * N * SP_MAX_PRESENT);
    drop(c.famocc, sizeof(float) * N * nf, sizeof(float) * N * SP_MAX_PRESENT);
    drop(c.dsettle, sizeof(float2) * N * nf, sizeof(float2) * N * SP_MAX_PRESENT);
    std::sort(skip.begin(), skip.end());
    std::vector<std::pair<size_t, size_t>> keep;  // (offset, bytes)
    size_t at = 0;
    for (const auto& k : skip) {
        if (k.first > at) keep.push_back({at, k.first - at});
        at = std::max(at, k.second);
    }
    if (s->arena_bytes > at) keep.push_back({at, s->arena_bytes - at});
    return keep;
}

// Evicts the tank: the arena's state (kept_ranges) goes to pinned host memory and its
// evictable device memory is freed. Nothing else of the tank changes, so a restored tank
// continues bit for bit.
void evict(sp_sim* s) {
    if (!s->resident) return;
    HIPCHK(hipStreamSynchronize(s->stream));
    std::vector<std::pair<size_t, size_t>> keep = kept_ranges(s);
    size_t total = 0;
    for (const auto& k : keep) total += k.second;
    void* h = nullptr;
    HIPCHK(hipHostMalloc(&h, std::max<size_t>(total, 1), hipHostMallocDefault));
    hipError_t e = hipSuccess;
    size_t at = 0;
    for (const auto& k : keep) {
        if (e == hipSuccess)
            e = hipMemcpyAsync((char*)h + at, s->ref.base + k.first, k.second, hipMemcpyDeviceToHost, s->stream);
        at += k.second;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        (void)hipHostFree(h);
        throw HipError{e, "hipMemcpyAsync (evict)"};
    }
    s->gs.clear();  // captured graphs hold the old addresses
    free_big(s);
    s->host_arena = h;
    s->host_ranges = std::move(keep);
    s->resident = false;
}

// Brings an evicted tank back: fresh device memory, the kept ranges copied in (the work space
// is written before it is read, as in any substep).
void restore(sp_sim* s) {
    if (s->resident) return;
    alloc_big(s);
    size_t at = 0;
    for (const auto& k : s->host_ranges) {
        HIPCHK(hipMemcpyAsync(s->ref.base + k.first, (const char*)s->host_arena + at, k.second, hipMemcpyHostToDevice,
                              s->stream));
        at += k.second;
    }
    HIPCHK(hipStreamSynchronize(s->stream));
    (void)hipHostFree(s->host_arena);
    s->host_arena = nullptr;
    s->host_ranges.clear();
    s->resident = true;
}

// Reads the step and solve state back and folds it into the statistics
// (poll_readbacks, mod.rs:3033-3116).
void readback_apply(sp_sim* s, const sp_stepbuf& S, const sp_solve& K) {
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
    const float margin = std::max(s->params.iters_margin, 0.0f);
    adapt(s->rec.p, K.it_max, p_un, s->params.pressure_max_iters, margin);
    if (s->gp.visc_mode == 2 && s->gp.v_twice) adapt(s->rec.p2, K.p2_it_max, p2_un, s->params.pressure_max_iters, margin);
    else s->rec.p2 = s->rec.p;  // not running: ready for a viscous liquid with what the first solve needs
    adapt(s->rec.v, K.v_it_max, v_un, std::max(s->params.viscosity_max_iters, 1u), 0.5f);
}

// The per-batch iteration maxima start again (after they were read), on `stream`;
// `zero` points at a zero word that lives until the copies ran.
void clear_maxima(sp_sim* s, hipStream_t stream, const uint32_t* zero) {
    HIPCHK(hipMemcpyAsync(&s->c.K->it_max, zero, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(&s->c.K->p2_it_max, zero, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(&s->c.K->v_it_max, zero, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
}

// Takes up the finished read-backs of sp_record's batches, oldest first (poll_readbacks,
// mod.rs:3030-3115); `block`: waits for every one in flight.
void poll_readbacks(sp_sim* s, bool block) {
    while (s->rb_busy > 0) {
        const uint32_t k = s->rb_next;
        if (block) {
            HIPCHK(hipEventSynchronize(s->rb_end[k]));
        } else {
            const hipError_t e = hipEventQuery(s->rb_end[k]);
            if (e == hipErrorNotReady) return;
            HIPCHK(e);
        }
        if (s->rb_substeps[k] > 0) {
            // the batch's device time per substep, smoothed over about ten batches
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, s->rb_start[k], s->rb_end[k]));
            const float cost = ms * 1e-3f / (float)s->rb_substeps[k];
            s->substep_cost = s->substep_cost > 0.0f ? s->substep_cost + 0.1f * (cost - s->substep_cost) : cost;
        }
        readback_apply(s, s->rb[k].S, s->rb[k].K);
        s->rb_next = (k + 1) % sp_sim::RB_RING;
        s->rb_busy--;
    }
}

// A read-back outside the ring (batches read their tanks back themselves), after the
// ring's.
void readback(sp_sim* s) {
    poll_readbacks(s, true);
    sp_stepbuf S;
    sp_solve K;
    HIPCHK(hipMemcpyAsync(&S, s->c.S, sizeof(S), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(&K, s->c.K, sizeof(K), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    readback_apply(s, S, K);
    const uint32_t zero = 0;
    clear_maxima(s, s->stream, &zero);
    HIPCHK(hipStreamSynchronize(s->stream));
}

// This tank's recorded iterations into its solve record, where the reductions stop its
// solves (a tank batched with tanks that record more stops where it would alone).
void write_caps(sp_sim* s, const Recs& rec, hipStream_t stream) {
    const uint32_t caps[3] = {rec.p, rec.p2, rec.v};
    HIPCHK(hipMemcpyAsync(&s->c.K->cap_p, caps, sizeof(caps), hipMemcpyHostToDevice, stream));
}

// The per-tank part of FluidSim::record (mod.rs:2400-2563) before the substeps are
// recorded: pending spawns into fresh or dead slots, the tables, and the batch words of the
// step state. On the tank's own stream.
int record_prepare(sp_sim* s, uint32_t substeps, float target, const sp_tool_state* tool = nullptr) {
    if (!s->params_set) {
        set_err(s, "sp_set_params has not been called");
        return SP_E_INVALID;
    }
    // the record's tool goes into the tables (gpu_params, mod.rs:2300-2305)
    const sp_tool_state t = tool ? *tool : sp_tool_state{};
    if (memcmp(&t, &s->tool, sizeof(t)) != 0) {
        s->tool = t;
        s->tables_dirty = true;
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
    // the spawns' speed bounds this record's steps only (record, mod.rs:2520)
    if (s->spawn_speed != 0.0f) {
        s->spawn_speed = 0.0f;
        s->tables_dirty = true;
    }
    launch_claim(s, claim, s->stream);
    if (substeps > 0) {
        // StepBuf words 0, 2, 3: target time, recorded substeps, done (record, mod.rs:2521-2526)
        uint32_t words[4];
        float tgt = std::max(target, 0.0f);
        memcpy(&words[0], &tgt, 4);
        words[2] = substeps;
        words[3] = 0;
        HIPCHK(hipMemcpyAsync(&s->c.S->frame_left, &words[0], 4, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(&s->c.S->batch_steps, &words[2], 8, hipMemcpyHostToDevice, s->stream));
        write_caps(s, s->rec, s->stream);
    }
    return SP_OK;
}

// The periodic sort before a substep (record_sort, mod.rs:3223-3242), on `stream`.
void maybe_sort(sp_sim* s, hipStream_t stream) {
    if (s->since_sort >= s->sort_interval) {
        launch_sort(s, 1, stream);
        s->since_sort = 0;
    }
    s->since_sort++;
}

// FluidSim::record (mod.rs:2400-2563): the batch on the tank's stream, with the recorded
// iterations that the read-backs finished so far set, and its read-back at its end.
int record(sp_sim* s, uint32_t substeps, float target, const sp_tool_state* tool = nullptr) {
    poll_readbacks(s, false);
    if (s->rb_busy == sp_sim::RB_RING) {
        // every slot in flight: the oldest batch has to finish first
        HIPCHK(hipEventSynchronize(s->rb_end[s->rb_next]));
        poll_readbacks(s, false);
    }
    int rc = record_prepare(s, substeps, target, tool);
    if (rc != SP_OK) return rc;
    // a record of no substeps (the uploads only) has nothing to read back: the next batches
    // keep the iterations they record (record, mod.rs:2559-2561)
    if (substeps == 0) return SP_OK;
    const uint32_t slot = (s->rb_next + s->rb_busy) % sp_sim::RB_RING;
    HIPCHK(hipEventRecord(s->rb_start[slot], s->stream));
    for (uint32_t k = 0; k < substeps; k++) {
        maybe_sort(s, s->stream);
        launch_substep(s->run, &s->gs, &s->tm, s->grid_valid ? 1 : 0, s->rec);
        s->grid_valid = true;
    }
    // the batch's records, then the per-batch iteration maxima start again
    HIPCHK(hipMemcpyAsync(&s->rb[slot].S, s->c.S, sizeof(sp_stepbuf), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(&s->rb[slot].K, s->c.K, sizeof(sp_solve), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemsetAsync(&s->c.K->it_max, 0, sizeof(uint32_t), s->stream));
    HIPCHK(hipMemsetAsync(&s->c.K->p2_it_max, 0, sizeof(uint32_t), s->stream));
    HIPCHK(hipMemsetAsync(&s->c.K->v_it_max, 0, sizeof(uint32_t), s->stream));
    HIPCHK(hipEventRecord(s->rb_end[slot], s->stream));
    s->rb_substeps[slot] = substeps;
    s->rb_busy++;
    return SP_OK;
}

int wait(sp_sim* s) {
    HIPCHK(hipStreamSynchronize(s->stream));
    s->tm.collect();
    poll_readbacks(s, true);
    return SP_OK;
}

// Runs an entry point's body with HIP errors mapped to status codes. `device`: the body
// touches the tank's device memory, so an evicted tank is brought back first.
template <class F>
int guarded(sp_sim* s, F&& f, bool device = true) {
    if (!s) return SP_E_INVALID;
    if (s->pooled) {
        set_err(s, "the tank is driven by a pool");
        return SP_E_BUSY;
    }
    try {
        if (device && !s->resident) {
            HIPCHK(hipSetDevice(s->device));
            restore(s);
        }
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

// Takes a new layout. A tank that turns elastic gives its particles a zeroed stress
// (set_elastic, mod.rs:2082-2100): particles already in it start unstressed.
void adopt_tank(sp_sim* s, const sp::Tank& t) {
    if (t.chan_e >= 0 && s->tank.chan_e < 0 && s->ref.base)
        HIPCHK(hipMemsetAsync(s->c.pstress, 0, sizeof(float4) * (size_t)std::max<uint32_t>(s->cap, 1), s->stream));
    s->tank = t;
}

// The rows the particles carry, with their diluents, and whether the tank becomes thermal
// (note_materials, mod.rs:2051-2081): a particle more than 0.05 K off the ambient
// temperature makes it so.
bool note_materials(sp_sim* s, const sp_particle* p, uint32_t n, bool reset) {
    std::vector<int> present = reset ? std::vector<int>() : s->tank.rows;
    bool thermal = s->thermal;
    for (uint32_t k = 0; k < n; k++) {
        uint32_t low = p[k].kind & 255u;
        if (p[k].kind == SP_KIND_DEAD || low == 0u) continue;
        int r = (int)low - 1;
        // a particle off the ambient temperature, or of a liquid that makes heat (material.rs:455-457)
        const sp_material& m = s->params.mat[r];
        thermal |= fabsf(p[k].t - s->params.ambient_temp) > 0.05f || m.self_heat > 0.0f || (m.react_with >= 0 && m.react_heat != 0.0f);
        if (std::find(present.begin(), present.end(), r) == present.end()) present.push_back(r);
    }
    std::vector<int> rows = sp::closure(s->params, present);
    if (rows == s->tank.rows && thermal == s->thermal && !(reset && s->tank.rows.empty())) return true;
    sp::Tank t;
    if (!sp::tank_layout_steam(s->params, rows, t, thermal)) return false;
    adopt_tank(s, t);
    s->thermal = thermal;
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
                          int32_t thermal, const sp_magnet* magnets, uint32_t n_magnets, const sp_current* currents,
                          uint32_t n_currents, sp_gparams* out, sp_gmat* mats) {
    if (!params || !out || !mats || (n_rows > 0 && !rows) || n_rows < 0 || (n_magnets > 0 && !magnets) ||
        (n_currents > 0 && !currents))
        return SP_E_INVALID;
    std::vector<int> present(rows, rows + n_rows);
    std::vector<int> r = sp::closure(*params, present);
    sp::Tank t;
    if (!sp::tank_layout_steam(*params, r, t, thermal != 0)) return SP_E_CAPACITY;
    std::vector<sp_magnet> mg(magnets, magnets + std::min<uint32_t>(n_magnets, SP_MAX_MAGNETS));
    std::vector<sp_current> cu(currents, currents + std::min<uint32_t>(n_currents, SP_MAX_CURRENTS));
    return sp::derive(*params, grid_w, grid_h, t, 0, *out, mats, nullptr, mg, 0.0f, cu);
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
    if (bin_chunks(cfg->capacity) > (uint32_t)bin::TOP_MAX) {
        g_create_error = "capacity beyond what one sort covers";
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
        s->tm.create();
        for (uint32_t k = 0; k < sp_sim::RB_RING; k++) {
            HIPCHK(hipEventCreate(&s->rb_start[k]));
            HIPCHK(hipEventCreate(&s->rb_end[k]));
        }
        HIPCHK(hipHostMalloc((void**)&s->rb, sizeof(sp_sim::Readback) * sp_sim::RB_RING, hipHostMallocDefault));
        sp::levels(s->W, s->H, s->lv);
        s->nb0 = ((s->W + 15) / 16) * ((s->H + 15) / 16);
        {
            const sp_levels_info& L = s->lv;
            size_t total = (size_t)(L.offs[L.n - 1] + L.dims[L.n - 1][0] * L.dims[L.n - 1][1] - L.offs[L.tail]);
            size_t nf = (size_t)L.dims[L.tail][0] * L.dims[L.tail][1];
            s->tail_lds_bytes = total * (16 + 8 + 8 + 4);
            if (total > SP_TAIL_SHARED || total - nf > SP_TAIL_SHARED_COARSE) {
                sp_destroy(s);
                g_create_error = "multigrid tail does not fit its shared-memory budget";
                return SP_E_INVALID;
            }
        }
        const size_t N = (size_t)s->N;
        auto& o = s->owned;
        // one arena for every buffer the kernels reach through the tank's context
        TankRef& ref = s->ref;
        ref = TankRef{nullptr, s->W, s->H, s->cap, s->lv.total, s->lv.tiles, s->lv.n, s->lv.tail};
        (void)ctx_of(ref, &s->arena_bytes);
        alloc_big(s);
        Ctx& c = s->c;
        s->d_spawn = dalloc<sp_particle>(SP_SPAWN_MAX, o);
        s->d_counter = dalloc<uint32_t>(8 + 2 * SP_MAX_MATERIALS, o);
        int4 h_lev[SP_MAX_LEVELS] = {};
        for (int l = 0; l < s->lv.n; l++) h_lev[l] = make_int4(s->lv.offs[l], s->lv.dims[l][0], s->lv.dims[l][1], s->lv.tile_offs[l]);
        HIPCHK(hipMemcpy((void*)c.lev, h_lev, sizeof(h_lev), hipMemcpyHostToDevice));
        s->d_count_out = dalloc<uint32_t>(N, o);
        s->d_gpumat = dalloc<sp_gpumat>(SP_MAX_MATERIALS, o);
        // solid field: closed everywhere until set_solids
        HIPCHK(hipDeviceSynchronize());
        sp::default_params(&s->params);
        Run& r = s->run;
        r.tb.t[0] = ref;
        r.T = 1;
        r.W = s->W;
        r.H = s->H;
        r.nb0 = s->nb0;
        r.lv = s->lv;
        r.stream = s->stream;
        r.tail_lds = s->tail_lds_bytes;
        // SPROUT_GRAPHS=0 launches the substep's kernels one by one instead of replaying
        // captured graphs (profilers that cannot follow graph launches need it)
        const char* gv = getenv("SPROUT_GRAPHS");
        s->gs.on = !(gv && gv[0] == '0');
        // SPROUT_MG_FUSE=0 runs the coarse multigrid levels one per launch (the same results)
        const char* fv = getenv("SPROUT_MG_FUSE");
        int lds_max = 0;
        HIPCHK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device));
        r.fuse = !(fv && fv[0] == '0') && (size_t)lds_max >= std::max(DOWN2_LDS, UP2_LDS);
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
    s->gs.clear();
    for (uint32_t k = 0; k < sp_sim::RB_RING; k++) {
        if (s->rb_start[k]) (void)hipEventDestroy(s->rb_start[k]);
        if (s->rb_end[k]) (void)hipEventDestroy(s->rb_end[k]);
    }
    if (s->rb) (void)hipHostFree(s->rb);
    free_big(s);
    for (void* p : s->owned) (void)hipFree(p);
    if (s->host_arena) (void)hipHostFree(s->host_arena);
    hipStream_t st = s->stream;
    delete s;  // graphs and events go with it
    if (st) (void)hipStreamDestroy(st);
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
        if (sp::tank_layout_steam(s->params, rows, t, s->thermal)) adopt_tank(s, t);
        return SP_OK;
    });
}

int sp_set_solids(sp_sim* s, const float* sdf, const float* solid, const uint32_t* region, const float* temp) {
    return guarded(s, [&]() -> int {
        if (!sdf || !solid || !region) return SP_E_INVALID;
        HIPCHK(hipMemcpyAsync((void*)s->c.sdf, sdf, sizeof(float) * (size_t)(s->W + 1) * (s->H + 1), hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync((void*)s->c.solid, solid, sizeof(float4) * (size_t)s->N, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync((void*)s->c.region, region, sizeof(uint32_t) * (size_t)s->N, hipMemcpyHostToDevice, s->stream));
        // solid temperatures (mod.rs:2442-2448): held ones make the tank thermal for good
        std::vector<float> st((size_t)s->N, SP_NO_TEMP);
        if (temp) st.assign(temp, temp + s->N);
        HIPCHK(hipMemcpyAsync((void*)s->c.solid_t, st.data(), sizeof(float) * (size_t)s->N, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        s->solids_heated = temp != nullptr;
        if (temp) {
            if (!s->thermal) {  // enable_heat (mod.rs:2163-2177)
                sp::Tank t;
                if (!sp::tank_layout_steam(s->params, s->tank.rows, t, true)) {
                    set_err(s, "the accumulators have no room for the temperature sums");
                    return SP_E_CAPACITY;
                }
                adopt_tank(s, t);
                s->thermal = true;
                s->tables_dirty = true;
            }
        }
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
        // a reset starts from an empty tank, thermal only if the solids are heated (mod.rs:1991-1992)
        s->thermal = s->solids_heated;
        s->tank = sp::Tank();
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
        HIPCHK(hipMemsetAsync(s->c.dheat, 0, sizeof(float4) * (size_t)s->N, s->stream));
        HIPCHK(hipMemsetAsync(s->c.ledger, 0, sizeof(uint32_t) * SP_LEDGER_WORDS, s->stream));
        HIPCHK(hipMemsetAsync(s->c.pstress, 0, sizeof(float4) * (size_t)std::max<uint32_t>(s->cap, 1), s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        poll_readbacks(s, true);  // what ran before the reset counts towards the old bases
        s->pending_spawn.clear();
        s->spawn_speed = 0.0f;
        s->resets++;
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
        // the pour's speed bounds the next record's time step (spawn, mod.rs:2039-2041)
        for (const auto& q : ps) {
            const float v = hypotf(q.vel[0], q.vel[1]);
            if (v > s->spawn_speed) {
                s->spawn_speed = v;
                s->tables_dirty = true;
            }
        }
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
        if (tool && (tool->mode < 0 || tool->mode > 8)) {
            set_err(s, "unknown tool mode %d", tool->mode);
            return SP_E_INVALID;
        }
        return record(s, substeps, target_s, tool);
    });
}

int sp_wait(sp_sim* s) {
    return guarded(s, [&]() -> int { return wait(s); });
}

int sp_poll(sp_sim* s, sp_solver_stats* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        poll_readbacks(s, false);
        *out = s->stats;
        return SP_OK;
    }, false);
}

int sp_step_blocking(sp_sim* s, uint32_t n, const sp_tool_state* tool) {
    return guarded(s, [&]() -> int {
        if (tool && (tool->mode < 0 || tool->mode > 8)) {
            set_err(s, "unknown tool mode %d", tool->mode);
            return SP_E_INVALID;
        }
        if (n == 0) {
            int rc = record(s, 0, 0.0f, tool);
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
            int rc = record(s, k, (float)target, tool);
            if (rc != SP_OK) return rc;
            rc = wait(s);
            if (rc != SP_OK) return rc;
        }
        return SP_OK;
    });
}

int sp_advance_blocking(sp_sim* s, double seconds, const sp_tool_state* tool) {
    return guarded(s, [&]() -> int {
        if (tool && (tool->mode < 0 || tool->mode > 8)) {
            set_err(s, "unknown tool mode %d", tool->mode);
            return SP_E_INVALID;
        }
        // advance_blocking (mod.rs:3284-3312): the tool acts in the first record only, the
        // later ones carry it idle (mode 0)
        sp_tool_state idle{};
        if (tool) {
            idle = *tool;
            idle.mode = 0;
        }
        bool first = true;
        double goal = s->sim_time + seconds;
        double sliver = std::max((double)s->params.min_substep_share, 1e-6);
        for (int guard = 0; guard < 1000000; guard++) {
            double left = goal - s->sim_time;
            double dt = s->params.fixed_hz > 0.0f ? 1.0 / s->params.fixed_hz : (double)s->dt_estimate;
            if (left <= sliver * dt) break;
            double nn = s->params.fixed_hz > 0.0f ? std::max(round(left / dt), 1.0) : ceil(left / dt);
            uint32_t k = (uint32_t)std::min(std::max(nn, 1.0), 16.0);
            double target = s->params.fixed_hz > 0.0f ? k * dt : left;
            int rc = record(s, k, (float)target, tool ? (first ? tool : &idle) : nullptr);
            if (rc != SP_OK) return rc;
            first = false;
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
    }, false);
}

int sp_frame_blocking(sp_sim* s, uint32_t substeps, float target_s, const sp_tool_state* tool, double* covered) {
    return guarded(s, [&]() -> int {
        if (tool && (tool->mode < 0 || tool->mode > 8)) {
            set_err(s, "unknown tool mode %d", tool->mode);
            return SP_E_INVALID;
        }
        const double start = s->sim_time;
        int rc = record(s, substeps, target_s, tool);
        if (rc != SP_OK) return rc;
        rc = wait(s);
        if (rc != SP_OK) return rc;
        if (covered) *covered = s->sim_time - start;
        return SP_OK;
    });
}

int sp_slots(sp_sim* s, uint32_t* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = s->n_slots;
        return SP_OK;
    }, false);
}

int sp_dt_estimate(sp_sim* s, float* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = s->params.fixed_hz > 0.0f ? 1.0f / s->params.fixed_hz : s->dt_estimate;
        return SP_OK;
    }, false);
}

int sp_recorded_iters(sp_sim* s, uint32_t* out3) {
    return guarded(s, [&]() -> int {
        if (!out3) return SP_E_INVALID;
        out3[0] = s->rec.p;
        out3[1] = s->rec.p2;
        out3[2] = s->rec.v;
        return SP_OK;
    }, false);
}

int sp_substep_cost(sp_sim* s, float* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = s->substep_cost;
        return SP_OK;
    }, false);
}

int sp_profile_passes(sp_sim* s, sp_pass_time* out, uint32_t cap, uint32_t* n) {
    return guarded(s, [&]() -> int {
        if (!n || (cap > 0 && !out)) return SP_E_INVALID;
        // what earlier batches timed is collected first, so the phases are this substep's
        HIPCHK(hipStreamSynchronize(s->stream));
        s->tm.collect();
        const float dt = s->params.fixed_hz > 0.0f ? 1.0f / s->params.fixed_hz : s->dt_estimate;
        int rc = record(s, 1, dt);
        if (rc != SP_OK) return rc;
        rc = wait(s);
        if (rc != SP_OK) return rc;
        const float* ph = s->tm.phase;
        const float part = s->tm.particle_ms;
        const char* names[8] = {"step_clear", "particles", "grid", "pressure", "project", "viscosity", "pressure2_project",
                                "total"};
        const double ms[8] = {std::max(ph[1] - part, 0.0f), part, ph[2], ph[3], ph[4], ph[5], ph[6], ph[0]};
        *n = 8;
        for (uint32_t k = 0; k < cap && k < 8; k++) {
            memset(out[k].name, 0, sizeof(out[k].name));
            strncpy(out[k].name, names[k], sizeof(out[k].name) - 1);
            out[k].ms = ms[k];
        }
        return SP_OK;
    });
}

int sp_material_table(sp_sim* s, sp_gpumat* out, uint32_t cap) {
    return guarded(s, [&]() -> int {
        if (cap > 0 && !out) return SP_E_INVALID;
        if (s->tables_dirty) {
            // the table as the next record derives it, without touching the device
            sp_gparams g;
            sp_gmat gm[SP_MAX_MATERIALS];
            std::string why;
            int rc = sp::derive(s->params, s->W, s->H, s->tank, s->n_slots, g, gm, &why, s->magnets, s->spawn_speed,
                                s->currents, &s->tool);
            if (rc != SP_OK) {
                set_err(s, "%s", why.c_str());
                return rc;
            }
            for (uint32_t r = 0; r < cap && r < (uint32_t)SP_MAX_MATERIALS; r++) out[r] = to_gpumat(gm[r]);
            return SP_OK;
        }
        for (uint32_t r = 0; r < cap && r < (uint32_t)SP_MAX_MATERIALS; r++) out[r] = s->gpumat[r];
        return SP_OK;
    }, false);
}

int sp_download_labels(sp_sim* s, int32_t* labels) {
    return guarded(s, [&]() -> int {
        if (!labels) return SP_E_INVALID;
        HIPCHK(hipMemcpyAsync(labels, s->c.rl, sizeof(int2) * (size_t)s->N, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_download_bodies(sp_sim* s, double* out, uint32_t cap, uint32_t* n) {
    return guarded(s, [&]() -> int {
        if (!n || (cap > 0 && !out)) return SP_E_INVALID;
        std::vector<uint32_t> w(SP_RIGID_WORDS);
        HIPCHK(hipMemcpyAsync(w.data(), s->c.rbody, sizeof(uint32_t) * SP_RIGID_WORDS, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        const uint32_t nb = std::min<uint32_t>(w[0], SP_RIGID_BODIES);
        auto get = [&](size_t k) { return (double)(int64_t)((uint64_t)w[k] | (uint64_t)w[k + 1] << 32) / 65536.0; };
        for (uint32_t b = 0; b < nb && b < cap; b++) {
            const size_t o = 16 + (size_t)SP_BODY_WORDS * b;
            for (int q = 0; q < 7; q++) out[7 * (size_t)b + q] = get(o + 2 * q);
        }
        *n = nb;
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

int sp_download_ledger(sp_sim* s, double* energy8) {
    return guarded(s, [&]() -> int {
        if (!energy8) return SP_E_INVALID;
        uint32_t w[SP_LEDGER_WORDS];
        HIPCHK(hipMemcpyAsync(w, s->c.ledger, sizeof(w), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        // download_ledger (mod.rs:3396-3399): LEDGER_SCALE per K of one water-filled cell
        const double unit = 998.0 * 4182.0 * (double)s->params.cell_size * (double)s->params.cell_size;
        for (int e = 0; e < 8; e++) energy8[e] = (double)((uint64_t)w[2 * e] | (uint64_t)w[2 * e + 1] << 32) / 65536.0 * unit;
        return SP_OK;
    });
}

int sp_download_ledger_counts(sp_sim* s, uint32_t* boiled, uint32_t* residue, uint32_t* steam) {
    return guarded(s, [&]() -> int {
        std::vector<uint32_t> w(SP_LEDGER_WORDS);
        HIPCHK(hipMemcpyAsync(w.data(), s->c.ledger, sizeof(uint32_t) * SP_LEDGER_WORDS, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        uint32_t* out[3] = {boiled, residue, steam};
        for (int k = 0; k < 3; k++)
            if (out[k]) memcpy(out[k], w.data() + 16 + k * SP_MAX_MATERIALS, sizeof(uint32_t) * SP_MAX_MATERIALS);
        return SP_OK;
    });
}

int sp_cell_counts(sp_sim* s, int32_t x0, int32_t y0, int32_t w, int32_t h, uint32_t* out) {
    return guarded(s, [&]() -> int {
        if (!out || w < 0 || h < 0 || x0 < 0 || y0 < 0 || x0 + w > s->W || y0 + h > s->H) return SP_E_INVALID;
        if (w == 0 || h == 0) return SP_OK;
        HIPCHK(hipMemsetAsync(s->d_count_out, 0, sizeof(uint32_t) * (size_t)w * h, s->stream));
        if (s->n_slots)
            hipLaunchKernelGGL(k_count, dim3((s->n_slots + 255) / 256), dim3(256), 0, s->stream, s->run.tb, x0, y0, w, h,
                               s->d_count_out);
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
        if (runs) hipLaunchKernelGGL(k_stats, dim3((runs + 63) / 64), dim3(64), 0, s->stream, s->run.tb, s->d_counter);
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

int sp_set_magnets(sp_sim* s, const sp_magnet* m, uint32_t n) {
    return guarded(s, [&]() -> int {
        if (n > 0 && !m) return SP_E_INVALID;
        s->magnets.assign(m, m + std::min<uint32_t>(n, SP_MAX_MAGNETS));
        s->tables_dirty = true;
        return SP_OK;
    }, false);
}

int sp_set_currents(sp_sim* s, const sp_current* cu, uint32_t n) {
    return guarded(s, [&]() -> int {
        if (n > 0 && !cu) return SP_E_INVALID;
        s->currents.assign(cu, cu + std::min<uint32_t>(n, SP_MAX_CURRENTS));
        s->tables_dirty = true;
        return SP_OK;
    }, false);
}

int sp_sort(sp_sim* s) {
    return guarded(s, [&]() -> int {
        int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        launch_sort(s, 0, s->stream);
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

const char* sp_error_text(void) { return g_call_error.c_str(); }

namespace {

// Checks the arguments of the handle-free binning; SP_OK or SP_E_INVALID with the reason.
int bin_args(const sp_particle* particles, uint32_t n, int32_t W, int32_t H, int32_t tiled, const sp_particle* sorted) {
    g_call_error.clear();
    if (W < 1 || H < 1 || W > 8192 || H > 8192 || (tiled != 0 && tiled != 1)) {
        g_call_error = "grid must be between 1 x 1 and 8192 x 8192 cells, tiled 0 or 1";
        return SP_E_INVALID;
    }
    if (n > 0 && (!particles || !sorted)) {
        g_call_error = "null particle array";
        return SP_E_INVALID;
    }
    const uintptr_t a = (uintptr_t)particles, b = (uintptr_t)sorted, bytes = (uintptr_t)n * sizeof(sp_particle);
    if (n > 0 && a < b + bytes && b < a + bytes) {
        g_call_error = "particles and sorted overlap";
        return SP_E_INVALID;
    }
    if (bin_chunks(n) > (uint32_t)bin::TOP_MAX) {
        g_call_error = "too many particles for one sort";
        return SP_E_INVALID;
    }
    return SP_OK;
}

// The binning of caller particles on `stream`, in the work space `ws` (bin_space_bytes(n)).
void launch_bin_particles(const sp_particle* particles, uint32_t n, int32_t W, int32_t H, int32_t tiled, sp_particle* sorted,
                          uint32_t* perm, void* ws, hipStream_t stream) {
    BinSpace sp = bin_space_at(ws, n);
    const int bits = bin_bits(W, H, tiled);
    const uint32_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_bin_keys, dim3(blocks), dim3(256), 0, stream, particles, n, W, H, tiled, bin_dead_key(bits), sp.keys,
                       sp.vals);
    const uint32_t* order = launch_bin_passes(sp, n, bits, stream);
    hipLaunchKernelGGL(k_bin_gather, dim3(blocks), dim3(256), 0, stream, particles, order, sorted, perm, n);
    HIPCHK(hipGetLastError());
}

}  // namespace

int sp_bin_particles(const sp_particle* particles, uint32_t n, int32_t grid_w, int32_t grid_h, int32_t tiled,
                     sp_particle* sorted, uint32_t* perm, void* stream) {
    int rc = bin_args(particles, n, grid_w, grid_h, tiled, sorted);
    if (rc != SP_OK || n == 0) return rc;
    const hipStream_t st = (hipStream_t)stream;
    void* ws = nullptr;
    hipError_t e = hipMalloc(&ws, bin_space_bytes(n));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_call_error = std::string("hipMalloc (binning work space): ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? SP_E_NOMEM : SP_E_DEVICE;
    }
    try {
        launch_bin_particles(particles, n, grid_w, grid_h, tiled, sorted, perm, ws, st);
        HIPCHK(hipStreamSynchronize(st));
    } catch (const HipError& err) {
        g_call_error = std::string(err.what) + " failed: " + hipGetErrorString(err.e);
        rc = SP_E_DEVICE;
    }
    (void)hipFree(ws);
    return rc;
}

// ---- the binning without allocation ------------------------------------------------------
//
// The work space of sp_bin_particles_noalloc: a fixed array in device memory, one per device
// (each device that runs the library holds its own copy until the process ends), sized for
// SP_BIN_NOALLOC_MAX particles. Host code never passes `scratch` itself to a kernel (on the
// host its name is only a stand-in for the device symbol): it takes the array's device
// address from hipGetSymbolAddress once per device and uses that address everywhere.
constexpr size_t BIN_NOALLOC_BYTES = bin_space_bytes(SP_BIN_NOALLOC_MAX);
static __device__ __align__(256) unsigned char scratch[BIN_NOALLOC_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up on the first call on
// that device (with it current) and kept.
hipError_t scratch_of_current_device(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace

int sp_bin_particles_noalloc(const sp_particle* particles, uint32_t n, int32_t grid_w, int32_t grid_h, int32_t tiled,
                             sp_particle* sorted, uint32_t* perm, void* stream) {
    int rc = bin_args(particles, n, grid_w, grid_h, tiled, sorted);
    if (rc != SP_OK || n == 0) return rc;
    if (n > SP_BIN_NOALLOC_MAX) {
        g_call_error = "sp_bin_particles_noalloc sorts at most SP_BIN_NOALLOC_MAX (" + std::to_string(SP_BIN_NOALLOC_MAX) +
                       ") particles; " + std::to_string(n) + " given";
        return SP_E_CAPACITY;
    }
    void* ws = nullptr;
    hipError_t e = scratch_of_current_device(&ws);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        g_call_error = std::string("hipGetSymbolAddress (binning work space): ") + hipGetErrorString(e);
        return SP_E_DEVICE;
    }
    try {
        launch_bin_particles(particles, n, grid_w, grid_h, tiled, sorted, perm, ws, (hipStream_t)stream);
    } catch (const HipError& err) {
        g_call_error = std::string(err.what) + " failed: " + hipGetErrorString(err.e);
        return SP_E_DEVICE;
    }
    return SP_OK;
}

namespace {

// The render buffer `which` of a resident tank: address, element count and size.
int render_buffer(sp_sim* s, int32_t which, void** ptr, uint32_t* count, uint32_t* stride) {
    const Ctx& c = s->c;
    switch (which) {
    case SP_RENDER_PARTICLES: *ptr = c.part; *count = s->n_slots; *stride = sizeof(sp_particle); break;
    case SP_RENDER_SDF: *ptr = (void*)c.sdf; *count = (uint32_t)((s->W + 1) * (s->H + 1)); *stride = sizeof(float); break;
    case SP_RENDER_CELL: *ptr = c.cell; *count = (uint32_t)s->N; *stride = sizeof(float4); break;
    case SP_RENDER_COMPOSITION: *ptr = c.comp; *count = (uint32_t)s->N; *stride = sizeof(float4); break;
    case SP_RENDER_SURFACE: *ptr = c.sprop; *count = (uint32_t)s->N; *stride = sizeof(float4); break;
    case SP_RENDER_MATERIALS: {
        // the table the kernels read, current with the parameters
        const int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        *ptr = (void*)c.mats;
        *count = SP_MAX_MATERIALS;
        *stride = sizeof(sp_gmat);
        break;
    }
    case SP_RENDER_VELOCITY: *ptr = c.vel; *count = (uint32_t)s->N; *stride = sizeof(float4); break;
    case SP_RENDER_GPUMAT: {
        const int rc = upload_tables(s);
        if (rc != SP_OK) return rc;
        *ptr = (void*)s->d_gpumat;
        *count = SP_MAX_MATERIALS;
        *stride = sizeof(sp_gpumat);
        break;
    }
    default: set_err(s, "unknown render buffer %d", which); return SP_E_INVALID;
    }
    return SP_OK;
}

}  // namespace

int sp_render_view(sp_sim* s, int32_t which, sp_device_view* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        void* p = nullptr;
        uint32_t count = 0, stride = 0;
        const int rc = render_buffer(s, which, &p, &count, &stride);
        if (rc != SP_OK) return rc;
        *out = sp_device_view{p, (uint64_t)count * stride, count, stride, s->device, 0u, s->generation};
        return SP_OK;
    });
}

int sp_render_download(sp_sim* s, int32_t which, void* out, size_t bytes) {
    return guarded(s, [&]() -> int {
        void* p = nullptr;
        uint32_t count = 0, stride = 0;
        const int rc = render_buffer(s, which, &p, &count, &stride);
        if (rc != SP_OK) return rc;
        const size_t n = (size_t)count * stride;
        if (!out || bytes < n) {
            set_err(s, "render buffer %d needs %zu bytes (%zu given)", which, n, bytes);
            return SP_E_INVALID;
        }
        if (n) HIPCHK(hipMemcpyAsync(out, p, n, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

int sp_stream(sp_sim* s, void** out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = (void*)s->stream;
        return SP_OK;
    }, false);
}

int sp_resets(sp_sim* s, uint64_t* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        *out = s->resets;
        return SP_OK;
    }, false);
}

// FluidSim::shares_a_group (mod.rs:1872-1880), steam aside.
int sp_shares_a_group(sp_sim* s, int32_t* out) {
    return guarded(s, [&]() -> int {
        if (!out) return SP_E_INVALID;
        int seen[8] = {0};
        *out = 0;
        for (int r : s->tank.rows) {
            if (r == SP_ROW_STEAM) continue;
            const int g = std::min(std::max(s->params.mat[r].group, 0), 7);
            if (++seen[g] >= 2) *out = 1;
        }
        return SP_OK;
    }, false);
}

int sp_evict(sp_sim* s) {
    return guarded(s, [&]() -> int {
        HIPCHK(hipSetDevice(s->device));
        HIPCHK(hipDeviceSynchronize());  // batches run tanks on their own streams
        evict(s);
        return SP_OK;
    }, false);
}

int sp_is_resident(const sp_sim* s, int32_t* out) {
    if (!s || !out) return SP_E_INVALID;
    *out = s->resident ? 1 : 0;
    return SP_OK;
}

int sp_tank_bytes(int32_t grid_w, int32_t grid_h, uint32_t capacity, uint64_t* evictable, uint64_t* fixed) {
    if (grid_w < 8 || grid_h < 8 || grid_w > 8192 || grid_h > 8192 || !evictable || !fixed) return SP_E_INVALID;
    sp_levels_info lv;
    sp::levels(grid_w, grid_h, lv);
    TankRef ref{nullptr, grid_w, grid_h, capacity, lv.total, lv.tiles, lv.n, lv.tail};
    size_t arena = 0;
    (void)ctx_of(ref, &arena);
    *evictable = big_bytes(arena, capacity);
    *fixed = sizeof(sp_particle) * SP_SPAWN_MAX + sizeof(uint32_t) * (8 + 2 * SP_MAX_MATERIALS) +
             sizeof(uint32_t) * (size_t)grid_w * grid_h;
    return SP_OK;
}

int sp_timings(sp_sim* s, float* out10) {
    return guarded(s, [&]() -> int {
        if (!out10) return SP_E_INVALID;
        for (int k = 0; k < 7; k++) out10[k] = s->tm.phase[k];
        out10[7] = (float)s->stats.pressure_iters_last;
        out10[8] = (float)s->last_v_iters;
        out10[9] = (float)s->last_p2_iters;
        return SP_OK;
    }, false);
}

int sp_autotune(int32_t mode, uint32_t samples) { return tuner().set(mode, samples) ? SP_OK : SP_E_INVALID; }

int sp_autotune_shapes(int32_t* threads, int32_t* box_ints, int32_t* min_groups, uint32_t cap, uint32_t* n) {
    if (!n) return SP_E_INVALID;
    *n = N_PSHAPES;
    for (uint32_t k = 0; k < cap && k < (uint32_t)N_PSHAPES; k++) {
        if (threads) threads[k] = P_SHAPES[k].threads;
        if (box_ints) box_ints[k] = P_SHAPES[k].box;
        if (min_groups) min_groups[k] = P_SHAPES[k].min_groups;
    }
    return SP_OK;
}

int sp_autotune_report(sp_tune_entry* out, uint32_t cap, uint32_t* n) {
    if (!n || (cap > 0 && !out)) return SP_E_INVALID;
    const std::vector<Tuner::Report> r = tuner().report();
    *n = (uint32_t)r.size();
    for (uint32_t k = 0; k < cap && k < r.size(); k++) {
        const uint64_t key = r[k].key;
        sp_tune_entry& e = out[k];
        e.grid_w = (int32_t)(key & 0x3fff);
        e.grid_h = (int32_t)(key >> 14 & 0x3fff);
        e.tanks = (int32_t)(key >> 28 & 0x3f);
        e.stride = (int32_t)(key >> 34 & 0xff);
        e.best = r[k].best;
        e.best_ms = r[k].best_ms;
        e.default_ms = r[k].default_ms;
        e.samples = r[k].samples;
    }
    return SP_OK;
}

int sp_kernel_time(sp_sim* s, int32_t kernel, float* avg_ms, uint32_t* launches) {
    return guarded(s, [&]() -> int {
        if (!avg_ms || !launches || kernel != 0) return SP_E_INVALID;
        *avg_ms = s->tm.particle_ms;
        *launches = s->tm.timed_last;
        return SP_OK;
    }, false);
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
    case SP_FIELD_THERM: *bytes = 16 * N; return c.therm;
    case SP_FIELD_DHEAT: *bytes = 16 * N; return c.dheat;
    case SP_FIELD_HCELL: *bytes = 8 * N; return c.hcell;
    case SP_FIELD_SOLID_T: *bytes = 4 * N; return c.solid_t;
    case SP_FIELD_LEDGER: *bytes = 4 * SP_LEDGER_WORDS; return c.ledger;
    case SP_FIELD_RL: *bytes = 8 * N; return c.rl;
    case SP_FIELD_RBODY: *bytes = 4 * (size_t)SP_RIGID_WORDS; return c.rbody;
    case SP_FIELD_RH: *bytes = 16 * N; return c.rh;
    case SP_FIELD_ESTR: *bytes = 16 * N; return c.estr;
    case SP_FIELD_PSTRESS: *bytes = 16 * (size_t)s->n_slots; return c.pstress;
    case SP_FIELD_FAMOCC: *bytes = 4 * N * (size_t)std::max(s->gp.n_fam, 1); return c.famocc;
    case SP_FIELD_DSETTLE: *bytes = 8 * N * (size_t)std::max(s->gp.n_fam, 1); return c.dsettle;
    case SP_FIELD_MAGC: *bytes = 8 * N; return c.magc;
    case SP_FIELD_RMASK: *bytes = 4 * N; return c.rmask;
    case SP_FIELD_MGX: *bytes = 8 * T; return c.mgx;
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
        const Run& r = s->run;
        const dim3 g0 = grid0(r);
        const TankBatch& cs = r.tb;
        // stages run with the most iterations the parameters allow
        const Recs full{s->params.pressure_max_iters, s->params.pressure_max_iters, std::max(s->params.viscosity_max_iters, 1u)};
        write_caps(s, full, s->stream);
        switch (stage) {
        case SP_STAGE_STEP: hipLaunchKernelGGL(k_step, per_tank(r), dim3(64), 0, s->stream, cs); break;
        case SP_STAGE_CLEAR_ACC: launch_clear_acc(r); break;
        case SP_STAGE_PARTICLES: launch_particles(r, arg); break;
        case SP_STAGE_GRID_PREP: hipLaunchKernelGGL(k_grid_prep, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_COEF: hipLaunchKernelGGL(k_coef_smooth, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_SURF:
            hipLaunchKernelGGL(k_coef_smooth, g0, TILE, 0, s->stream, cs);
            hipLaunchKernelGGL(k_surf_curv, g0, TILE, 0, s->stream, cs);
            hipLaunchKernelGGL(k_surf_apply, g0, TILE, 0, s->stream, cs);
            break;
        case SP_STAGE_VISC: launch_visc_before(r, full); break;
        case SP_STAGE_VISC_SPLIT: launch_visc_split(r); break;
        case SP_STAGE_VISC_SOLVE: launch_visc_solve(r, full.v, arg != 0); break;
        case SP_STAGE_RHS_VISC: hipLaunchKernelGGL(k_rhs_visc, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_PRESSURE2: launch_pressure(r, full.p2, true); break;
        case SP_STAGE_PROJECT_VISC: hipLaunchKernelGGL(k_project_visc, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_VISC_VCYCLE: {
            // One V-cycle from vr, whatever the state of the last solve: the level kernels
            // return at once on a converged solve, so the flag is cleared around it.
            uint32_t done = 0, halt = 0, zero = 0;
            HIPCHK(hipStreamSynchronize(s->stream));
            HIPCHK(hipMemcpy(&done, &c.K->v_done, 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&halt, &c.K->v_halt, 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(&c.K->v_done, &zero, 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(&c.K->v_halt, &zero, 4, hipMemcpyHostToDevice));
            if (s->gp.v_mg_sweeps == 1) launch_visc_vcycle<1>(r);
            else launch_visc_vcycle<2>(r);
            HIPCHK(hipStreamSynchronize(s->stream));
            HIPCHK(hipMemcpy(&c.K->v_done, &done, 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(&c.K->v_halt, &halt, 4, hipMemcpyHostToDevice));
            break;
        }
        case SP_STAGE_RHS: hipLaunchKernelGGL(k_rhs, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_HEAT: hipLaunchKernelGGL(k_heat, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_RHEO: hipLaunchKernelGGL(k_rheo, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_ELASTIC: hipLaunchKernelGGL(k_elastic, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_SETTLE: hipLaunchKernelGGL(k_settle, g0, TILE, 0, s->stream, cs); break;
        case SP_STAGE_MAGNET:
            if (arg < 0 || arg > 1) return SP_E_INVALID;
            hipLaunchKernelGGL(k_magnet, g0, TILE, 0, s->stream, cs, arg);
            break;
        case SP_STAGE_RIGID:
            if (arg < 0 || arg > 4) return SP_E_INVALID;
            hipLaunchKernelGGL(k_rigid, g0, TILE, 0, s->stream, cs, arg);
            break;
        case SP_STAGE_PRESSURE: launch_pressure(r, full.p); break;
        case SP_STAGE_PROJECT: launch_project(r); break;
        case SP_STAGE_SUBSTEP:
            launch_substep(r, &s->gs, nullptr, arg, full);
            s->grid_valid = true;
            break;
        case SP_STAGE_MG_BUILD: launch_mg_build(r); break;
        case SP_STAGE_VCYCLE: launch_vcycle(r, 0, 0); break;
        default: set_err(s, "unknown stage %d", stage); return SP_E_INVALID;
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->stream));
        return SP_OK;
    });
}

}  // extern "C"

#include "batch.hip"
#include "pool.hip"
