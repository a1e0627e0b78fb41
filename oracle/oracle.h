/*
 * oracle.h — CPU restatement of the Sprout fluid step. TEST INFRASTRUCTURE ONLY: this code
 * is the checker the HIP path is compared against, never the thing measured or shipped.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 */
#ifndef SPROUT_ORACLE_H
#define SPROUT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/sprout_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_sim or_sim;

or_sim* or_create(int32_t W, int32_t H, uint32_t capacity);
void or_destroy(or_sim* s);
/* Grid-unit tables (as sp_derive_grid_params makes them). */
void or_set_gparams(or_sim* s, const sp_gparams* P, const sp_gmat* mats);
void or_set_solids(or_sim* s, const float* sdf, const float* solid, const uint32_t* region, const float* temp);
/* Replaces the particles; clears the warm-start state and the step state (reset). */
void or_reset(or_sim* s, const sp_particle* p, uint32_t n);
/* Sets the batch words of StepBuf (record: frame_left, batch_steps, batch_done = 0). */
void or_begin_batch(or_sim* s, float target, uint32_t substeps);
/* One substep (record_substep). gather = 0 on the first substep after a reset. */
void or_substep(or_sim* s, int gather);
/* A single stage (SP_STAGE_* ids). */
void or_stage(or_sim* s, int stage, int arg);
/* Field access (SP_FIELD_* ids); returns bytes copied, or -1. */
long or_get(or_sim* s, int field, void* out, size_t bytes);
long or_put(or_sim* s, int field, const void* in, size_t bytes);
uint32_t or_n_slots(or_sim* s);
/* Canonical counting sort by cell, stable by slot (sort.comp with a defined order). */
void or_sort(or_sim* s);
/* The spawn step of record: fresh slots first, then dead slots in slot order (claim.comp). */
void or_spawn(or_sim* s, const sp_particle* spawn, uint32_t n);
/* Live particles per cell in a window (count.comp). */
void or_cell_counts(or_sim* s, int32_t x0, int32_t y0, int32_t w, int32_t h, uint32_t* out);
/* Live count and per-row counts and concentration sums (stats.comp). */
void or_stats(or_sim* s, uint32_t* live, uint32_t* count, uint32_t* conc256);
/* Levels::new restated (mod.rs:727-753). */
void or_levels(int32_t W, int32_t H, sp_levels_info* out);
/* The pressure iterations of the last solve. */
uint32_t or_last_iters(or_sim* s);
/* The air's velocity at (x, y) for a puff leaning `lean` and the liquid occupancy there
 * (air.glsl:148-175), at the step state's time. */
void or_air_velocity(or_sim* s, float x, float y, float lean, float* vel2, float* occ);

#ifdef __cplusplus
}
#endif

#endif
