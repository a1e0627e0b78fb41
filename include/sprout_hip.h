/*
 * sprout_hip.h — C ABI of the MI355X-native Sprout fluid step (libsprout_hip.so).
 *
 * This is the drop-in boundary for the path that Sprout's Rust host drives through
 * `FluidSim` (src/sim/mod.rs:862-3492): a handle owns one tank's device state, the caller
 * hands over particles, solids and parameters, records substeps, and downloads results.
 * Every entry point returns an int status (0 = ok, < 0 = SP_E_* code); the text of the last
 * error of a handle is available through sp_last_error(). No C++ exception or abort crosses
 * this boundary. Calls on one handle are not thread-safe (the reference takes `&mut self`
 * everywhere); different handles may be used from different threads.
 *
 * All types here are plain C: the layouts of sp_particle and the step state follow the
 * reference byte for byte (checked by static assertions in sprout_types.h).
 */
#ifndef SPROUT_HIP_H
#define SPROUT_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "sprout_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define SP_OK 0
#define SP_E_INVALID (-1)      /* bad argument (null pointer, size, NaN parameter ...) */
#define SP_E_DEVICE (-2)       /* a HIP runtime call failed */
#define SP_E_NOMEM (-3)        /* device or host allocation failed */
#define SP_E_UNSUPPORTED (-4)  /* a feature of the reference this build does not carry */
#define SP_E_CAPACITY (-5)     /* more particles / materials than the handle holds */
#define SP_E_BUSY (-6)         /* the handle is driven by a pool; or a request is not done yet */

typedef struct sp_sim sp_sim;

/* Creation parameters: the grid and the particle capacity of `FluidSim::new(gpu, grid,
 * capacity)` (src/sim/mod.rs:1034), plus the HIP device ordinal. */
typedef struct sp_config {
    int32_t grid_w, grid_h;  /* cells; at least 8 x 8 */
    uint32_t capacity;       /* particle slots (the reference sizes W*H*PPC, src/world.rs:155) */
    int32_t device;          /* HIP device ordinal */
} sp_config;

/* Replaces `FluidSim::new` (src/sim/mod.rs:1034). The handle owns all device memory. */
int sp_create(const sp_config* cfg, sp_sim** out);
/* Frees the handle and its device memory (Drop of FluidSim). NULL is accepted. */
void sp_destroy(sp_sim* sim);
/* Text of the last error on this handle ("" if none). Valid until the next call. */
const char* sp_last_error(const sp_sim* sim);
/* Text of the last error of sp_create (no handle exists then). */
const char* sp_create_error(void);

/* Replaces assigning `FluidSim::params` (src/sim/mod.rs:862-1031, field `params`): SI
 * parameters and the material table. The host side derives the grid-unit tables from
 * them at the next record (gpu_params / gpu_materials, src/sim/mod.rs:2179-2330).
 * NaN values are refused (src/sim/params.rs:986-997). */
int sp_set_params(sp_sim* sim, const sp_sim_params* params);

/* Replaces `FluidSim::set_solids(SolidField)` (src/sim/mod.rs:1977, 2437-2448,
 * src/scene.rs:110-122): sdf[(W+1)*(H+1)] node distances, solid[4*W*H] (open u face, open v
 * face, smoothed solid occupancy, open volume), region[W*H] (region number |
 * SP_REGION_VENT), temp[W*H] the temperature (deg C) a solid is held at, SP_NO_TEMP where
 * solids insulate, or NULL when no solid has a temperature. A non-NULL temp makes the tank
 * thermal (enable_heat, src/sim/mod.rs:2163-2177). */
int sp_set_solids(sp_sim* sim, const float* sdf, const float* solid, const uint32_t* region,
                  const float* temp);

/* Replaces `FluidSim::reset(Vec<Particle>)` (src/sim/mod.rs:1982-2001): ids are assigned
 * 0..n in order, the spawn fields are settled, state that warm-starts solves is cleared. */
int sp_reset(sp_sim* sim, const sp_particle* particles, uint32_t n);

/* Replaces `FluidSim::spawn` (src/sim/mod.rs:2027-2048): queues particles for the next
 * record; *accepted receives how many were taken. */
int sp_spawn(sp_sim* sim, const sp_particle* particles, uint32_t n, uint32_t* accepted);

/* Replaces `FluidSim::record(builder, substeps, target, tool)` (src/sim/mod.rs:2400-2563):
 * enqueues `substeps` substeps meant to cover `target_s` seconds on the handle's stream, with
 * the read-back of the batch's records at its end, and returns without waiting (unless four
 * batches are already in flight: then it waits for the oldest). The recorded iterations come
 * from the read-backs finished when it is called, as in the reference's frame loop. tool may
 * be NULL (idle); only mode 0 is carried. */
int sp_record(sp_sim* sim, uint32_t substeps, float target_s, const sp_tool_state* tool);
/* Waits for everything recorded on the handle (submit_and_wait, src/sim/mod.rs:3483-3491). */
int sp_wait(sp_sim* sim);
/* Replaces `FluidSim::poll_readbacks` + `solver` (src/sim/mod.rs:3033-3116, :334-363):
 * takes up the read-backs of the batches finished so far, oldest first (simulated time,
 * statistics, the next batches' recorded iterations) and copies the statistics. Never
 * blocks. */
int sp_poll(sp_sim* sim, sp_solver_stats* out);

/* `FluidSim::advance_blocking(seconds, tool)` (src/sim/mod.rs:3284-3312): records of at most
 * 16 substeps until `seconds` are covered; `tool` (NULL = the idle tool) acts in the first
 * record, the later ones carry it idle. */
int sp_advance_blocking(sp_sim* sim, double seconds, const sp_tool_state* tool);
/* `FluidSim::step_blocking(n, tool)` (src/sim/mod.rs:3316-3338): n ticks of 1/240 s in
 * records of at most 16 substeps, each under `tool` (NULL = the idle tool); n = 0 records no
 * substep and uploads what changed. */
int sp_step_blocking(sp_sim* sim, uint32_t n, const sp_tool_state* tool);
/* `FluidSim::sim_time` (src/sim/mod.rs:1956). */
int sp_sim_time(sp_sim* sim, double* out);
/* `FluidSim::frame_blocking(substeps, target, tool)` (src/sim/mod.rs:3342-3350): one
 * interactive frame, recorded and waited for; *covered receives the simulated time it
 * covered (s). covered may be NULL. */
int sp_frame_blocking(sp_sim* sim, uint32_t substeps, float target_s, const sp_tool_state* tool, double* covered);
/* `FluidSim::slots` (src/sim/mod.rs:1946): particle slots that may hold live particles. */
int sp_slots(sp_sim* sim, uint32_t* out);
/* `FluidSim::dt_estimate` (src/sim/mod.rs:1961): the substep length the next batch will
 * most likely get (1 / fixed_hz when the rate is fixed), s. */
int sp_dt_estimate(sp_sim* sim, float* out);
/* `FluidSim::recorded_iters` (src/sim/mod.rs:1966): the iterations the next substeps
 * record for the first pressure solve, the second and the viscosity solve. */
int sp_recorded_iters(sp_sim* sim, uint32_t* out3);
/* `FluidSim::substep_cost` (src/sim/mod.rs:1972, 3055-3059): device time of one substep,
 * s, from events around each batch recorded by sp_record, taken up by sp_poll or sp_wait,
 * smoothed as x += 0.1 (cost - x); 0 before the first measurement. */
int sp_substep_cost(sp_sim* sim, float* out);
/* `FluidSim::profile_passes` (src/sim/mod.rs:3120-3150): runs one substep (it advances the
 * tank, as the reference's does) with device timestamps between its phases and returns
 * each phase's time (ms) and the total: "step_clear", "particles", "grid", "pressure",
 * "project", "viscosity", "pressure2_project", "total". At most cap entries go to out;
 * *n receives the number of phases. Blocking. */
int sp_profile_passes(sp_sim* sim, sp_pass_time* out, uint32_t cap, uint32_t* n);
/* The material table in the reference's `GpuMat` layout (src/sim/material.rs:508-549),
 * as `FluidSim::material_buffer` holds it after the latest record: at most cap rows of
 * SP_MAX_MATERIALS. Host side; no device work. */
int sp_material_table(sp_sim* sim, sp_gpumat* out, uint32_t cap);

/* `FluidSim::download_particles` (src/sim/mod.rs:3421-3432): the live particles (dead
 * slots dropped), in slot order. out may be NULL to query the count; otherwise at most
 * cap are written and *n receives the number written (or available if out is NULL). */
int sp_download_particles(sp_sim* sim, sp_particle* out, uint32_t cap, uint32_t* n);
/* `FluidSim::download_grid` (src/sim/mod.rs:3436-3450) without the rigid labels:
 * vel[4*W*H] (u, v, state u, state v) and cell[4*W*H] (occupancy, density, viscosity,
 * level). Either pointer may be NULL. */
int sp_download_grid(sp_sim* sim, float* vel, float* cell);
/* The rigid labels of `GridDump` (src/sim/mod.rs:3436-3450, `labels`): labels[2*W*H], per
 * cell the root cell of its body (-1: not rigid) and, at a root, the body's slot. */
int sp_download_labels(sp_sim* sim, int32_t* labels);
/* `FluidSim::download_bodies` (src/sim/mod.rs:3451-3462): per body of the latest substep
 * its mass, mass-weighted offset sums (x, y), momentum (x, y), angular momentum about the
 * root and inertia sum (grid units), 7 doubles each. At most cap bodies are written to out
 * (may be NULL with cap 0); *n receives the bodies the table holds. */
int sp_download_bodies(sp_sim* sim, double* out, uint32_t cap, uint32_t* n);
/* `FluidSim::download_pressure` (src/sim/mod.rs:3464-3475): p*dt per cell, and dt. */
int sp_download_pressure(sp_sim* sim, float* pres, float* dt);
/* `FluidSim::download_ledger` (src/sim/mod.rs:3389-3406): energy that crossed the boundary
 * since the last reset, J per metre of depth: [0] gained from air, [1] lost to air, [2]
 * gained from solids, [3] lost to solids, [4]-[7] heat tool, cool tool (0 here: the tools
 * are not carried), boiled off (boiling, evaporation, what reactions use up), made inside. */
int sp_download_ledger(sp_sim* sim, double* energy8);
/* The counts of `FluidSim::download_ledger` (src/sim/mod.rs:3400-3404): per table row the
 * particles boiled or evaporated away, the residues left and the steam made since the last
 * reset (SP_MAX_MATERIALS each; any pointer may be NULL). Only evaporation counts here. */
int sp_download_ledger_counts(sp_sim* sim, uint32_t* boiled, uint32_t* residue, uint32_t* steam);
/* `FluidSim::cell_counts(lo, hi)` (src/sim/mod.rs:3354-3386, count.comp): live particles
 * per cell in [x0, x0+w) x [y0, y0+h), row by row, into out[w*h]. */
int sp_cell_counts(sp_sim* sim, int32_t x0, int32_t y0, int32_t w, int32_t h, uint32_t* out);
/* `FluidSim::request_stats/poll_stats` (src/sim/mod.rs:3246-3274, stats.comp), blocking:
 * live particles, and per table row the count and the concentration sum. */
int sp_stats(sp_sim* sim, sp_stats_out* out);

/* `FluidSim::set_magnets` (src/sim/mod.rs:2018-2022): replaces the magnets placed in the
 * world (the first SP_MAX_MAGNETS are kept); they survive a reset. A magnet and a liquid
 * with a susceptibility turn the magnetic pass on (magnet.comp). */
int sp_set_magnets(sp_sim* sim, const sp_magnet* magnets, uint32_t n);

/* `FluidSim::set_currents` (src/sim/mod.rs:2008-2012): replaces the currents and drains (the
 * first SP_MAX_CURRENTS are kept); they act from the next record on and survive a reset. A
 * current raises the velocity of the liquid inside it along its direction towards its speed
 * (never braking it), a drain removes the liquid inside it (particles.comp:578-592, 809-829). */
int sp_set_currents(sp_sim* sim, const sp_current* currents, uint32_t n);

/* The periodic counting sort of the particles by cell (record_sort, src/sim/mod.rs:
 * 3223-3242, sort.comp). Runs on the handle's stream; the order within a cell is stable
 * by the previous slot (the reference's is atomic-ordered, sort.comp:64). */
int sp_sort(sp_sim* sim);

/* The same counting sort (sort.comp:26-68, with the order defined as above) over caller
 * device memory, for particles that no handle owns: particles[n] in, sorted[n] out (they
 * must not overlap), by cell_of(pos) of a grid_w x grid_h grid (tiled = 1: by 8 x 8 tile,
 * then by cell within it), dead slots last. perm[n] (device memory, may be NULL) receives
 * the input slot of each output position, for moving per-particle data alongside. Runs on
 * `stream` (a hipStream_t; NULL: the null stream) on the current device, with work space
 * the library allocates for the call, and returns once the sort has finished. Errors:
 * SP_E_INVALID (arguments), SP_E_NOMEM, SP_E_DEVICE; sp_error_text() tells which. */
int sp_bin_particles(const sp_particle* particles, uint32_t n, int32_t grid_w, int32_t grid_h, int32_t tiled,
                     sp_particle* sorted, uint32_t* perm, void* stream);
/* No counterpart in the reference. sp_bin_particles for callers that cannot allocate
 * device memory (work recorded under HIP graph capture, serving loops that forbid
 * hipMalloc): the work space is a fixed array of the library in device memory, one per
 * device (about 17 MB, held until the process ends), instead of memory allocated for the
 * call, so at most SP_BIN_NOALLOC_MAX particles (more: SP_E_CAPACITY). The call enqueues
 * the sort on `stream` and returns without waiting. Every call on one device uses that
 * device's one array: calls whose sorts overlap on a device (other streams, other threads)
 * race and give wrong results; order them (one stream, or events). */
#define SP_BIN_NOALLOC_MAX 1048576u
int sp_bin_particles_noalloc(const sp_particle* particles, uint32_t n, int32_t grid_w, int32_t grid_h, int32_t tiled,
                             sp_particle* sorted, uint32_t* perm, void* stream);
/* Text of the last error of a handle-free entry point on this thread ("" if none). */
const char* sp_error_text(void);

/* ---- render-facing outputs ------------------------------------------------------------- */

/* The buffers a renderer reads (FluidSim::particle_buffer, sdf_buffer, cell_buffer,
 * composition_buffer, surface_buffer, material_buffer; src/sim/mod.rs:1608-1940, read by
 * render.rs:639, 817-819 and record_cell_optics). */
enum {
    SP_RENDER_PARTICLES = 0,    /* sp_particle[n_slots] (particle_buffer; dead slots included) */
    SP_RENDER_SDF = 1,          /* float[(W+1)(H+1)], node SDF of the solids (sdf_buffer) */
    SP_RENDER_CELL = 2,         /* float4[W*H]: occupancy, density, viscosity, level (cell_buffer) */
    SP_RENDER_COMPOSITION = 3,  /* float4[W*H]: density, oil share, family-0 concentration, fill (composition_buffer) */
    SP_RENDER_SURFACE = 4,      /* float4[W*H]: tension, contact normal, group + (row + 1) / 128 (surface_buffer) */
    SP_RENDER_MATERIALS = 5,    /* sp_gmat[SP_MAX_MATERIALS] (material_buffer, this library's row layout) */
    SP_RENDER_VELOCITY = 6,     /* float4[W*H]: u, v and the faces' states (the grid's velocity) */
    SP_RENDER_GPUMAT = 7,       /* sp_gpumat[SP_MAX_MATERIALS] (material_buffer in the reference's GpuMat layout) */
    SP_RENDER_COUNT
};

/* Where a render buffer lives: a device address on `device`, valid until the tank's memory
 * moves. `generation` changes whenever it does (eviction and the restore on the next call,
 * destruction), so a renderer that keeps the address checks it once per frame. */
typedef struct sp_device_view {
    void* ptr;
    uint64_t bytes;
    uint32_t count;      /* elements */
    uint32_t stride;     /* bytes per element */
    int32_t device;
    uint32_t pad;
    uint64_t generation;
} sp_device_view;

/* The device view of render buffer `which` (SP_RENDER_*), restoring an evicted tank first.
 * The contents are what the last finished substep left: read after sp_wait / a completed
 * sp_poll, or on the tank's own stream (sp_stream). Replaces the Subbuffer accessors of
 * src/sim/mod.rs:1608-1940 for a renderer on the same device (HIP, or Vulkan through
 * external memory). */
int sp_render_view(sp_sim* sim, int32_t which, sp_device_view* out);
/* A host copy of render buffer `which`, for a renderer elsewhere: bytes must be at least
 * the view's size. Waits for the tank's work first. */
int sp_render_download(sp_sim* sim, int32_t which, void* out, size_t bytes);
/* The hipStream_t the tank's work runs on, as a void pointer (for ordering a renderer's
 * reads after the step without a host wait). */
int sp_stream(sp_sim* sim, void** out);
/* FluidSim::resets (src/sim/mod.rs:1941-1943): how many times the tank was emptied and
 * refilled; what a renderer remembers from frame to frame is void when it changes. */
int sp_resets(sp_sim* sim, uint64_t* out);
/* FluidSim::shares_a_group (src/sim/mod.rs:1872-1880): whether two liquids of one phase
 * group are in the tank (the renderer then tints from the particles). */
int sp_shares_a_group(sp_sim* sim, int32_t* out);

/* ---- batches of tanks ------------------------------------------------------------------ */

/* No counterpart in the reference, which steps one tank per FluidSim: the batched workload.
 * A batch advances many handles together. Tanks of the same grid size and solver schedule
 * share every kernel launch of a substep (up to 32 tanks per launch), each group of them on
 * a stream of its own so that groups overlap on the device. Each tank's results are those it
 * gets stepped alone. The handles stay owned by the caller, must outlive the batch, be on one
 * device, and must not be used on their own while a batch call on them runs. */
typedef struct sp_batch sp_batch;
int sp_batch_create(sp_sim* const* sims, uint32_t n, sp_batch** out);
void sp_batch_destroy(sp_batch* batch);
const char* sp_batch_last_error(const sp_batch* batch);
/* sp_record for every tank of the batch: `substeps` substeps covering `target_s` seconds,
 * enqueued on the groups' streams (spawns, tables and step words of each tank first). */
int sp_batch_record(sp_batch* batch, uint32_t substeps, float target_s);
/* Waits for the batch and folds each tank's solver statistics in (sp_poll then reads them). */
int sp_batch_wait(sp_batch* batch);
/* The launch groups of the last record: their number, and the tanks of each (up to cap). */
int sp_batch_groups(const sp_batch* batch, uint32_t* n_groups, uint32_t* tanks_per_group, uint32_t cap);
/* Device ms per substep of a group's phases ([0]-[6] as sp_timings) and of its particle
 * kernel ([7]), averaged over the substeps of the last record. */
int sp_batch_timings(const sp_batch* batch, uint32_t group, float* out8);

/* ---- residency ------------------------------------------------------------------------- */

/* No counterpart in the reference (one tank per process, always on the GPU): device memory
 * of many resident tanks. sp_evict moves the tank's arena (every field, the particles, the
 * solver state) to pinned host memory and frees its evictable device memory; the next call
 * that needs the device brings it back, and the tank continues bit for bit as if it had
 * stayed. It waits for the whole device first. */
int sp_evict(sp_sim* sim);
/* 1 if the tank's arena is on the device, 0 if evicted. */
int sp_is_resident(const sp_sim* sim, int32_t* out);
/* Device bytes a tank of this grid and particle capacity holds: `evictable` (given back by
 * sp_evict) and `fixed` (held for the handle's life). No device needed. */
int sp_tank_bytes(int32_t grid_w, int32_t grid_h, uint32_t capacity, uint64_t* evictable, uint64_t* fixed);

/* ---- request records of the pool ---------------------------------------------------- */

typedef struct sp_request_info {
    uint64_t ticket;
    uint32_t scene, substeps;
    int32_t status;          /* SP_OK, or the error that failed it */
    uint32_t tanks;          /* tanks in the launch that ran it */
    double submit_ms, launch_ms, done_ms;   /* pool clock */
    double latency_ms;       /* done - submit */
} sp_request_info;
typedef struct sp_pool_stats {
    uint64_t submitted, completed, failed;
    uint64_t particle_steps;     /* particle slots x substeps of the completed requests */
    uint64_t launches;           /* group launches */
    uint64_t evictions, restores;
    uint64_t resident_bytes;     /* evictable bytes of the resident scenes */
    uint32_t scenes, resident;
    double p50_ms, p99_ms, mean_ms, max_ms;   /* request latency, completed requests */
    double span_s;               /* first submit to last completion */
    double tanks_per_launch;     /* mean */
    uint64_t rounds;             /* rounds run */
} sp_pool_stats;

/* ---- the continuous-batching pool ------------------------------------------------------ */

/* No counterpart in the reference: serving step requests for many resident scenes (tanks),
 * for callers that run a serving loop. Requests are queued from any thread
 * (sp_pool_submit, never blocking) and run in rounds by whichever thread calls sp_pool_step
 * (one round), sp_pool_wait or sp_pool_drain (rounds until done). In a round every scene with
 * an active request advances one substep: scenes of the same grid and solver schedule share
 * each launch (up to 32 tanks, sp_batch's grouping), launches of different groups overlap on
 * `streams` streams, and the round ends when they all have. A request of n substeps is active
 * for n consecutive rounds; a request that arrives while a round runs starts at the next one
 * (continuous batching at substep granularity); a scene's requests run in arrival order, one
 * at a time. Scenes stay on the device while their evictable bytes (sp_tank_bytes) fit
 * `hbm_budget`; over it, idle scenes are evicted to host memory, least recently used first,
 * and brought back when a request of theirs starts, between rounds. Each scene's results are
 * those it gets stepped alone with the same requests (sp_record + sp_wait per request).
 * While added to a pool, a handle must not be used by the caller (its calls return
 * SP_E_BUSY); after sp_pool_destroy it is the caller's again. Request records are kept for
 * the last 2^20 done requests. */
typedef struct sp_pool sp_pool;
int sp_pool_create(int32_t device, uint64_t hbm_budget, uint32_t streams, sp_pool** out);
/* Runs every request already submitted, then gives the handles back. */
void sp_pool_destroy(sp_pool* pool);
const char* sp_pool_last_error(sp_pool* pool);
/* Adds a tank (set up: parameters, solids, particles) as scene *scene (0, 1, ...);
 * SP_E_CAPACITY if its evictable bytes alone exceed the budget. */
int sp_pool_add(sp_pool* pool, sp_sim* sim, uint32_t* scene);
/* Queues a request of `substeps` substeps covering `target_s` seconds (as sp_record).
 * Thread-safe; never waits for a round. */
int sp_pool_submit(sp_pool* pool, uint32_t scene, uint32_t substeps, float target_s, uint64_t* ticket);
/* Runs one round; with nothing to run, first waits up to `wait_ms` for a request.
 * *advanced: the scenes the round advanced (0: none). */
int sp_pool_step(sp_pool* pool, double wait_ms, uint32_t* advanced);
/* The request's state: SP_E_BUSY while it is not done. */
int sp_pool_poll(sp_pool* pool, uint64_t ticket, sp_request_info* out);
/* Runs rounds (or waits for the thread running them) until the request is done. */
int sp_pool_wait(sp_pool* pool, uint64_t ticket, sp_request_info* out);
/* Runs rounds until every submitted request is done. */
int sp_pool_drain(sp_pool* pool);
/* Counters and latency percentiles since the last sp_pool_reset_stats (launch_ms in
 * sp_request_info: the request's first round). */
int sp_pool_stats_get(sp_pool* pool, sp_pool_stats* out);
int sp_pool_reset_stats(sp_pool* pool);

/* ---- launch-shape autotuner ----------------------------------------------------------- */

/* No counterpart in the reference (its work-group shapes are fixed in the shaders): the
 * particle pass's launch shape (work-items per group, capacity of the LDS scatter box) is
 * chosen per key (grid, tanks per launch, accumulator stride) by timing the candidates on the
 * launches themselves, process-wide. Every shape gives bit-identical results, so tuning never
 * changes a result. mode -1: tune (the default); -2: forget what was tuned and tune again;
 * 0 .. n-1: use that shape everywhere (sp_autotune_shapes). samples > 0: timed launches per
 * candidate before the fastest median is kept (default 3). No device needed. */
int sp_autotune(int32_t mode, uint32_t samples);
/* The candidates for k < min(cap, *n): threads[k] work-items per group, box_ints[k] ints of
 * LDS scatter box, min_groups[k] groups per compute unit the registers are limited for (1:
 * no limit). Any output may be NULL. */
int sp_autotune_shapes(int32_t* threads, int32_t* box_ints, int32_t* min_groups, uint32_t cap, uint32_t* n);
typedef struct sp_tune_entry {
    int32_t grid_w, grid_h, tanks, stride;  /* the key */
    int32_t best;                           /* chosen shape, -1 while tuning */
    float best_ms, default_ms;              /* median particle-pass ms: chosen shape, shape 0 */
    uint32_t samples;                       /* timings taken */
} sp_tune_entry;
/* The keys seen so far: out[k] for k < min(cap, *n). */
int sp_autotune_report(sp_tune_entry* out, uint32_t cap, uint32_t* n);

/* ---- host-side derivations (no device needed) ---------------------------------------- */

/* Default SI parameters and the built-in materials this build carries (water, honey, oil,
 * mercury at table rows 0, 1, 2, 8) — SimParams::default (src/sim/params.rs:386-460). */
void sp_default_params(sp_sim_params* out);

/* Tank::layout (src/sim/mod.rs:271-331) + gpu_params / gpu_materials (src/sim/mod.rs:
 * 2179-2330) for a tank holding the table rows `rows[0..n_rows)` (ascending), isothermal
 * or (thermal = 1) carrying temperatures, with magnets[0..n_magnets) placed: fills the
 * grid-unit parameter block and the material table (SP_MAX_MATERIALS rows). */
int sp_derive_grid_params(const sp_sim_params* params, int32_t grid_w, int32_t grid_h,
                          const int32_t* rows, int32_t n_rows, int32_t thermal, const sp_magnet* magnets,
                          uint32_t n_magnets, const sp_current* currents, uint32_t n_currents,
                          sp_gparams* out, sp_gmat* mats);

/* Levels::new (src/sim/mod.rs:727-753): the multigrid hierarchy of a W x H grid. */
int sp_levels(int32_t grid_w, int32_t grid_h, sp_levels_info* out);

/* ---- stage-level access for parity tests and diagnostics ----------------------------- */

/* Copy an internal device field to / from host memory (SP_FIELD_* ids, sizes in bytes).
 * These synchronise the handle's stream. */
int sp_debug_download(sp_sim* sim, int32_t field, void* out, size_t bytes);
int sp_debug_upload(sp_sim* sim, int32_t field, const void* data, size_t bytes);
/* Runs one stage (SP_STAGE_*) of the substep on the current device state, synchronously. */
int sp_debug_stage(sp_sim* sim, int32_t stage, int32_t arg);
/* The grid-unit tables the handle currently uses. */
int sp_debug_gparams(sp_sim* sim, sp_gparams* out, sp_gmat* mats);
/* Device time per substep of the last recorded batch's phases (ms, averaged over its
 * substeps): [0] total, [1] step + accumulator clear + particle pass, [2] grid passes
 * before the solve (with the viscosity that precedes it), [3] pressure solve, [4]
 * projection, [5] implicit viscosity solve after the projection, [6] second pressure solve
 * and projection ([5] and [6] are 0 unless the tank takes them); then the iterations of the
 * latest solves: [7] pressure, [8] viscosity, [9] second pressure. Valid after sp_wait. */
int sp_timings(sp_sim* sim, float* out10);
/* Average device time (ms) of one launch of a hot kernel over the substeps of the last
 * batch (kernel 0: the particle pass, k_particles), timed with HIP events on the handle's
 * stream, and the number of launches averaged. */
int sp_kernel_time(sp_sim* sim, int32_t kernel, float* avg_ms, uint32_t* launches);

#ifdef __cplusplus
}
#endif

#endif
